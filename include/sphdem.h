/*
 * sphdem.h -- C ABI of libsphdem.so, the MI355X (gfx950) implementation of
 * the coupled WCSPH-DEM time step and of the Eulerian ISPH cavity step.
 *
 * Reference counterparts are cited as file:line of hojin9908/Warp_SPH_GPT.
 * The reference binds its path through NVIDIA Warp (wp.launch of @wp.kernel
 * functions with SoA structs and wp.HashGrid ids); this ABI replaces that
 * operator layer.  Python binds it with ctypes (warp_sph_gpt_amd/_lib.py).
 *
 * Conventions
 *  - Every REAL buffer is void*: float when `precision` == 4 (the reference's
 *    FP32 path) or double when `precision` == 8 (the FP64 shadow of the same
 *    kernels).  vec3 fields are packed (N,3) arrays, as Warp's wp.vec3 arrays
 *    and torch (N,3) tensors are.  Index/offset buffers are int32.
 *  - The caller owns every buffer (device memory, e.g. torch tensors).  The
 *    library allocates nothing and borrows pointers only for the work it
 *    enqueues.  All work is enqueued on `stream` (a hipStream_t; NULL is the
 *    default stream) in call order; no call synchronises the host.
 *  - Every call returns SPHDEM_OK or an error code; sphdem_last_error()
 *    returns the message of the last failure on the calling thread.
 *  - Particle outputs are written at the particle's original index, as the
 *    reference writes P.field[i] with i = hash_grid_point_id(grid, tid).
 */
#ifndef SPHDEM_H
#define SPHDEM_H

#ifdef __cplusplus
extern "C" {
#endif

#define SPHDEM_OK 0
#define SPHDEM_E_ARG 1      /* invalid argument: maps to ValueError */
#define SPHDEM_E_HIP 2      /* HIP runtime/launch failure: RuntimeError */
#define SPHDEM_E_OVERFLOW 3 /* CSR capacity exceeded: OverflowError */

/* Solver constants of one call.  Fields mirror input/Config.py:11-36 `Solv`
 * (h, rho0, c0, gamma, mu, g, dt, shepard_step, dem_porosity_min/max) and
 * the cached dem_search_scales tuple (input/input_reader.py:163-168). */
typedef struct {
    int precision;               /* 4 = FP32, 8 = FP64 */
    double h, support;           /* Solv.h, Solv.support = 2h */
    double rho0, c0, gamma, mu, g, dt;
    double eps_min, eps_max;     /* Solv.dem_porosity_min / _max */
    int shepard_step;            /* Solv.shepard_step */
    double radius_max;           /* search_scales[0]: max moving DEM radius */
    double boundary_radius_max;  /* search_scales[1]: max DEM-wall radius */
    double h_max;                /* search_scales[2]: max DEM coupling h */
    double h_uniform;            /* search_scales[3]: common DEM h or 0 */
    int fast_math;               /* 0: W, dW/dr with IEEE division and square
                                    root, as the reference writes them;
                                    1: in the porosity term and the force pass
                                    the hardware square root and reciprocal,
                                    1/h multiplies, fused multiply-adds
                                    (densities stay exact) */
} sphdem_params_t;

/* input/struct.py:3-21 SPHptl */
typedef struct {
    void *pos, *vel, *rho_raw, *rho, *pres, *acc, *m, *flt, *porosity, *pgf, *acc_dem;
    int n;
} sphdem_sph_t;

/* input/struct.py:23-36 BNDptl */
typedef struct {
    void *pos, *vel, *rho_raw, *rho, *pres, *acc, *m;
    int n;
} sphdem_bnd_t;

/* input/struct.py:39-82 DEMptl.  CSR id/tang buffers hold `cap_*` entries;
 * the logical edge count E is offset[n] on the device. */
typedef struct {
    void *pos, *vel, *acc, *force, *omega, *torque, *radius, *rho, *m, *volume,
         *inertia, *K, *eta, *mu, *h, *flt_s, *porosity, *drag, *pressure_force;
    int *contact_dem_offset_old, *contact_dem_id_old;
    void *tang_dem_old;
    int *contact_dem_offset_new, *contact_dem_id_new;
    void *tang_dem_new;
    int *contact_bnd_offset_old, *contact_bnd_id_old;
    void *tang_bnd_old;
    int *contact_bnd_offset_new, *contact_bnd_id_new;
    void *tang_bnd_new;
    int cap_dem_new, cap_bnd_new;   /* capacity of the *_new id/tang buffers */
    int n;
} sphdem_dem_t;

/* input/struct.py:85-104 DEMBNDptl */
typedef struct {
    void *pos, *vel, *omega, *radius, *rho, *m, *volume, *inertia, *K, *eta, *mu;
    int n;
} sphdem_dembnd_t;

/* A built hash grid: replaces wp.HashGrid (constructor main.py:117-122,
 * build main.py:66,73,85,87) and the handle kernels receive as grid.id.
 * Buffers are owned by the caller; sizes in brackets. */
typedef struct {
    int precision;        /* element type of `sorted` */
    int dim_x, dim_y, dim_z;
    double cell_width;    /* the `radius` of the last build */
    int n;                /* points in the last build */
    int *point_cells;     /* [n]   bucket key of original point i */
    int *point_ids;       /* [n]   sorted slot -> original index (stable) */
    int *rank;            /* [n]   after a build: the sorted slot of original point i */
    int *cell_lo;         /* [ncell] first sorted slot of each bucket */
    int *cell_hi;         /* [ncell] one past the last slot of each bucket */
    void *sorted;         /* [n*4] (x, y, z, w) in sorted order, w from `w` */
    int *scan_tmp;        /* [sphdem_scan_scratch(ncell)] */
    /* Batched instances (no reference counterpart): independent scenes share
     * one grid, each with its own dim_x*dim_y*dim_z buckets, so ncell in the
     * sizes above is dim_x*dim_y*dim_z*max(n_inst, 1).  inst[i] is the
     * instance of point i and must be
     * nondecreasing, so instance k owns one contiguous range of points, both
     * in input and in sorted order.  A query by a point of instance k sees
     * instance k's buckets only.  n_inst <= 1 and inst = NULL: one scene. */
    int n_inst;
    const int *inst;      /* [n] or NULL */
} sphdem_grid_t;

/* Per-step working arrays of the fused step, in grid-sorted order.
 * Sizes: fluid arrays >= fluid count, wall arrays >= wall count. */
typedef struct {
    void *fl_rho, *fl_pres, *fl_V;  /* [nf] density, pressure, m/rho */
    void *fl_vel, *fl_pgf;          /* [nf*4] (vel, rho), (-grad p, rho) */
    void *wl_rho, *wl_pres;         /* [nb] */
    void *wl_vel;                   /* [nb*4] (vel, rho) */
    int *scan_tmp;                  /* [sphdem_scan_scratch(n_dem + 1)] */
    int *flags;                     /* [4] bit 0: DEM CSR overflow */
    int *fl_flag;                   /* [nf] near-DEM flags */
    int *fl_list;                   /* [2*nf] see fl_sum */
    int *fl_count;                  /* [4] see fl_chunk */
    int *fl_scan;                   /* [sphdem_scan_scratch(nf)] */
    /* [cap_d] or NULL: per-DEM flags, all zero between steps.  In the fused
     * coupled step the porosity mark sets dem_near[b] = 1 where some fluid
     * point lies within 2 h_b, and the DEM interpolation skips the fluid
     * search of the other DEMs (it would find no term) and clears the flags
     * it read.  NULL (or cap_d below the DEM count): no skipping. */
    int *dem_near;
    /* [cap_b] or NULL: for callers whose walls never move (an engine that
     * builds the wall grid once).  Per wall slot the density sum over wall
     * neighbours alone, NaN-filled by the caller and whenever the wall grid,
     * B->pos or B->m change.  A wall particle with no fluid term takes its
     * density from here (the same bits), and the pass fills a NaN on first
     * use.  NULL: every wall sum is evaluated. */
    void *wl_static;
    /* [6 * 257] or NULL: scratch of the fused step's fluid bounding box
     * (min x, y, z and max x, y, z of the fluid grid's sorted positions,
     * all instances; 256 partial boxes, then the box).  The wall pass skips
     * the fluid query of a wall particle that lies farther than the support
     * from this box along some axis: that sum has no term (+0).  Used for
     * wall sets of 4M particles or more, where it pays.  NULL: every wall
     * particle runs its fluid query. */
    void *fl_box;
    /* [cap_f] REAL or NULL: with fast_math on FP32 and a uniform DEM h, the
     * porosity's fluid sums are evaluated symmetrically (a pair of near-DEM
     * subjects once, for both sums: the same terms, other rounding) and
     * added in fixed point: fl_list is read as cap_f uint64 accumulators
     * that must be zero between steps (zero-filled once at allocation; the
     * step clears what it used).  The pass leaves each near-DEM subject's
     * unscaled sum here.  NULL: every subject sums its own terms. */
    void *fl_sum;
    /* [7 * cap_f] ints or NULL: scratch of the porosity's work list (the
     * near-DEM slots of each bucket row, compacted and cut every 64; one
     * wavefront per entry; counts in fl_count[0..1], the stream's item
     * counter in fl_count[2]), rebuilt every step.
     * NULL: one wavefront per 64 consecutive slots (the same sums; with the
     * reference's arithmetic the same bits). */
    int *fl_chunk;
    int cap_f, cap_b, cap_d;
} sphdem_work_t;

/* Library identity and diagnostics. */
int sphdem_version(void);
const char *sphdem_last_error(void);
int sphdem_device_count(void);

/* Scratch ints an exclusive scan of n entries needs. */
int sphdem_scan_scratch(int n);

/* Exclusive prefix sum of a[0..n) in place: wp.utils.array_scan(a, a,
 * inclusive=False), source/Simulation.py:148-149, 181-182. */
int sphdem_exclusive_scan(int *a, int n, int *scratch, void *stream);

/* No reference counterpart: sphdem_exclusive_scan for callers that cannot
 * allocate device memory (code under hipGraph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The tile totals live in a device array
 * the library holds per GPU until the process ends (32,769 ints, resolved on
 * the device's first call or by sphdem_device_init) instead of a caller's
 * scratch, so n is at most SPHDEM_SCAN_NOALLOC_MAX (larger: SPHDEM_E_ARG,
 * nothing enqueued).  It allocates nothing.  Every call on one GPU shares the
 * array: calls that overlap on a GPU (other streams or threads) race. */
#define SPHDEM_SCAN_NOALLOC_MAX (1 << 26)
int sphdem_exclusive_scan_noalloc(int *a, int n, void *stream);

/* wp.HashGrid.build(points, radius) (main.py:66,73,85,87): bucket keys with
 * 1/radius cells folded into dim_x*dim_y*dim_z buckets, a stable counting
 * sort of the points by key, bucket ranges, and a sorted copy of the points
 * (w column from `w`, or 0 when w is NULL).  `points`/`w` use g->precision. */
int sphdem_grid_build(sphdem_grid_t *g, const void *points, int n, double radius,
                      const void *w, void *stream);

/* Refill the w column of the sorted copy from w[original index] (the fluid
 * and wall kernels read neighbour masses from it). */
int sphdem_grid_set_w(sphdem_grid_t *g, const void *w, void *stream);

/* Warp-convention bucket bounds (0,0 for empty buckets) for parity views. */
int sphdem_grid_warp_bounds(const sphdem_grid_t *g, int *cell_starts, int *cell_ends,
                            void *stream);

/* Number of true neighbours (r^2 < radius^2, self included) of each subject
 * point of grid gs in grid gn, written at the subject's original index. */
int sphdem_neighbour_counts(const sphdem_grid_t *gs, const sphdem_grid_t *gn,
                            double radius, int *counts, void *stream);

/* kernel/Kernel_KNL.py:7-41 sampled at r[0..n): W and dW/dr. */
int sphdem_wendland(int precision, const void *r, int n, double h, void *w, void *dw,
                    void *stream);

/* kernel/KERNEL_SPHDEM_interaction.py:11-40 DEM_drag_beta sampled elementwise. */
int sphdem_drag_beta(int precision, const void *eps, const void *rho, const void *mu,
                     const void *diameter, const void *speed, int n, void *beta,
                     void *stream);

/* ---- one entry per reference kernel ------------------------------------ */
/* kernel/KERNEL_rho.py:6-40 */
int sphdem_shepard_sph(const sphdem_params_t *p, sphdem_sph_t *P, sphdem_bnd_t *B,
                       const sphdem_grid_t *gs, const sphdem_grid_t *gb, void *stream);
/* kernel/KERNEL_rho.py:45-82 */
int sphdem_density_sph(const sphdem_params_t *p, sphdem_sph_t *P, sphdem_bnd_t *B,
                       const sphdem_grid_t *gs, const sphdem_grid_t *gb, void *stream);
/* kernel/KERNEL_rho.py:87-126 */
int sphdem_density_bnd(const sphdem_params_t *p, sphdem_sph_t *P, sphdem_bnd_t *B,
                       const sphdem_grid_t *gs, const sphdem_grid_t *gb, void *stream);
/* kernel/KERNEL_pres.py:6-23 and 26-45: p = B*max((rho/rho0)^gamma - 1, 0) */
int sphdem_pres(const sphdem_params_t *p, const void *rho, void *pres, int n, void *stream);
/* kernel/KERNEL_force.py:7-71 (uses w as scratch) */
int sphdem_force_sph(const sphdem_params_t *p, sphdem_sph_t *P, sphdem_bnd_t *B,
                     const sphdem_grid_t *gs, const sphdem_grid_t *gb,
                     const sphdem_work_t *w, void *stream);
/* kernel/KERNEL_step.py:6-22 */
int sphdem_step_sph(const sphdem_params_t *p, sphdem_sph_t *P, void *stream);
/* kernel/KERNEL_DEM_force.py:9-22 */
int sphdem_count_dem_contacts(const sphdem_params_t *p, sphdem_dem_t *D,
                              const sphdem_grid_t *gd, double radius_max, void *stream);
/* kernel/KERNEL_DEM_force.py:77-135 (flags[0] bit 0 set on capacity overflow) */
int sphdem_force_dem(const sphdem_params_t *p, sphdem_dem_t *D, const sphdem_grid_t *gd,
                     double radius_max, double g, int *flags, void *stream);
/* kernel/KERNEL_DEM_BC.py:10-24 */
int sphdem_count_bnd_contacts(const sphdem_params_t *p, sphdem_dem_t *D,
                              const sphdem_dembnd_t *W, const sphdem_grid_t *gw,
                              double radius_max, void *stream);
/* kernel/KERNEL_DEM_BC.py:27-84 */
int sphdem_bc_dem(const sphdem_params_t *p, sphdem_dem_t *D, const sphdem_dembnd_t *W,
                  const sphdem_grid_t *gw, double radius_max, int *flags, void *stream);
/* kernel/KERNEL_DEM_step.py:6-30 */
int sphdem_step_dem(const sphdem_params_t *p, sphdem_dem_t *D, void *stream);
/* kernel/KERNEL_SPHDEM_interaction.py:43-115; `support`/`h` are the coupling
 * support (2*h_max) and the uniform DEM h (0 when heterogeneous). */
int sphdem_prep_sphdem(const sphdem_params_t *p, sphdem_sph_t *P, sphdem_bnd_t *B,
                       sphdem_dem_t *D, const sphdem_grid_t *gs, const sphdem_grid_t *gb,
                       const sphdem_grid_t *gd, double support, double h,
                       const sphdem_work_t *w, void *stream);
/* kernel/KERNEL_SPHDEM_interaction.py:118-187 */
int sphdem_interaction_dem(const sphdem_params_t *p, sphdem_sph_t *P, sphdem_dem_t *D,
                           const sphdem_grid_t *gs, const sphdem_grid_t *gd,
                           const sphdem_work_t *w, void *stream);
/* kernel/KERNEL_SPHDEM_interaction.py:190-226 */
int sphdem_interaction_sph(const sphdem_params_t *p, sphdem_sph_t *P, sphdem_dem_t *D,
                           const sphdem_grid_t *gs, const sphdem_grid_t *gd,
                           double support, void *stream);

/* ---- fused steps (the hot path) ---------------------------------------- */
/* source/Simulation.py:17-99 SPH_OneStep: Shepard (at step 0 and every
 * shepard_step), density and EOS of fluid and wall, force, and (integrate)
 * semi-implicit Euler.  Grids must be built from the current positions. */
int sphdem_sph_step(const sphdem_params_t *p, sphdem_sph_t *P, sphdem_bnd_t *B,
                    const sphdem_grid_t *gs, const sphdem_grid_t *gb, int step,
                    int integrate, const sphdem_work_t *w, void *stream);

/* Per-device set-up (no reference counterpart; Warp creates its streams on
 * wp.init()): the side streams and events the fused step's DEM contact work,
 * its wall density pass and its force pass run on (default priority), and the
 * device address of sphdem_exclusive_scan_noalloc's array, for the current
 * device.  Call once per device
 * before capturing a coupled step into a hipGraph, outside capture; a first
 * step outside capture does the same.  Until then a captured step records a
 * single-stream step (the same results, without the overlap). */
int sphdem_device_init(void *stream);

/* source/Simulation.py:102-221 SPHDEM_OneStep, split at the two points where
 * the reference reads the CSR totals back to the host:
 *   stage 0: SPH fields and forces, DEM-DEM counts + scan (E_dem = offset[n])
 *   stage 1: DEM-DEM fill, DEM-wall counts + scan
 *   stage 2: DEM-wall fill, coupling, integration of SPH and DEM
 * The caller swaps old/new CSR sides after stages 1 and 2 exactly as the
 * reference does.  stage -1 runs all three with no host round trip; it needs
 * *_new capacities large enough (flags bit 0 reports overflow) and copies the
 * new CSR back into the old buffers instead of swapping. */
int sphdem_coupled_step(const sphdem_params_t *p, sphdem_sph_t *P, sphdem_bnd_t *B,
                        sphdem_dem_t *D, sphdem_dembnd_t *W, const sphdem_grid_t *gs,
                        const sphdem_grid_t *gb, const sphdem_grid_t *gd,
                        const sphdem_grid_t *gw, int step, int stage,
                        const sphdem_work_t *w, void *stream);

/* ---- per-pass timing (no reference counterpart) ------------------------ */
/* When enabled, every pass records a pair of HIP events on its stream;
 * sphdem_profile_read waits for them and adds the elapsed milliseconds and
 * call counts per pass id (array index = id).  Not for use under graph
 * capture. */
#define SPHDEM_PASS_GRID 0          /* hash-grid build */
#define SPHDEM_PASS_FLUID_DENSITY 1 /* Shepard + density + EOS, fluid */
#define SPHDEM_PASS_WALL_DENSITY 2  /* density + EOS, wall */
#define SPHDEM_PASS_FLUID_FORCE 3   /* acceleration (+ -grad p) */
#define SPHDEM_PASS_DEM_CSR 4       /* contact counts + scan */
#define SPHDEM_PASS_DEM_CONTACT 5   /* DEM-DEM / DEM-wall loads + CSR fill */
#define SPHDEM_PASS_POROSITY 6      /* coupling-scale porosity */
#define SPHDEM_PASS_INTERACTION_DEM 7
#define SPHDEM_PASS_FLUID_FINAL 8   /* drag reaction + integration */
#define SPHDEM_PASS_STEP_DEM 9
#define SPHDEM_PASS_OTHER 10
#define SPHDEM_NUM_PASSES 11
/* Passes of the Eulerian cavity step (profiled, not tunable). */
#define SPHDEM_PASS_EISPH_ADVEC 11
#define SPHDEM_PASS_EISPH_PPE 12
#define SPHDEM_PASS_EISPH_SOLVE 13
#define SPHDEM_PASS_EISPH_CORRECT 14
#define SPHDEM_PASS_EISPH_FINISH 15
#define SPHDEM_PASS_EISPH_STEPS 16
#define SPHDEM_NUM_PROFILED_PASSES 17
int sphdem_profile_enable(int on);
int sphdem_profile_read(double *ms, int *counts, int n);

/* ---- launch shapes (no reference counterpart) -------------------------- */
/* Process-wide launch shapes of the per-subject passes, set by the
 * autotuner.  They change which lanes and workgroups evaluate which
 * subjects, never the terms of a subject's sums or their order, so results
 * are the same bits under every valid setting.  Honoured by: GRID,
 * FLUID_DENSITY, WALL_DENSITY, FLUID_FORCE, POROSITY and FLUID_FINAL.
 *   block[pass]      threads per workgroup: 64, 128 or 256 (default 256)
 *   xcd_remap[pass]  1: hand consecutive subject tiles to one XCD, so that
 *                    neighbouring tiles share that XCD's L2 (default 0)
 *   por_x_tile       log2 of the buckets along x that one porosity stream
 *                    sub-group may span: 2..6 (default 4)
 *   dem_beside       1: in the fused coupled step the DEM interpolation runs
 *                    on the side stream beside the porosity stream kernel;
 *                    0: after it, on the caller's stream (default 1)
 *   por_parts        with the porosity work list and the symmetric FP32
 *                    sums: waves per work item, each streaming every
 *                    por_parts-th bucket row of the item's query: 1..16
 *                    (default 4).  The terms are the same; their grouping
 *                    into fixed-point partial sums changes with it, so the
 *                    sums may differ in the last units of that form.
 *   por_queue        how the porosity's waves take their (item, part) work
 *                    (work list only): 0 wave w takes items w, w + waves,
 *                    ...; 1 every wave takes the next item from one counter
 *                    when it finishes one (default 1).  The same partial
 *                    sums under both values.
 *   wall_beside      1: in the fused coupled step, off a Shepard step, the
 *                    wall density pass runs on a stream of its own beside the
 *                    fluid density pass (neither reads what the other
 *                    writes); 0: after it on the caller's stream (default 1)
 *   force_beside     1: in the fused coupled step the force pass runs on a
 *                    stream of its own beside the porosity passes (the
 *                    force writes accelerations, densities and pressure
 *                    gradients, the porosity reads positions, m/rho and the
 *                    DEM state); 0: before them on the caller's stream
 *                    (default 1) */
typedef struct {
    int block[SPHDEM_NUM_PASSES];
    int xcd_remap[SPHDEM_NUM_PASSES];
    int por_x_tile;
    int dem_beside;
    int por_parts;
    int por_queue;
    int wall_beside;
    int force_beside;
} sphdem_tuning_t;
int sphdem_tuning_get(sphdem_tuning_t *t);
int sphdem_tuning_set(const sphdem_tuning_t *t);

/* ---- FP64 diagnostics (no reference counterpart) ----------------------- */
/* out[0] = total fluid mass, out[1] = fluid KE 1/2 sum m|v|^2,
 * out[2] = DEM KE incl. 1/2 I|w|^2, out[3] = DEM mass, out[4] = surge front
 * (max fluid x), all accumulated in FP64.  `out` is device memory [5]. */
int sphdem_reduce_f64(int precision, const void *sph_pos, const void *sph_vel,
                      const void *sph_m, int nf, const void *dem_vel, const void *dem_omega,
                      const void *dem_m, const void *dem_inertia, int nd, double *out,
                      void *stream);

/* ---- learned surrogate (no reference counterpart) ---------------------- */
/* A two-round message-passing network over the SPH neighbourhoods of the
 * fluid and wall grids (width 64, SiLU), BF16 operands on MFMA with FP32
 * accumulation; it stands in for ten WCSPH substeps.  All node arrays are in
 * the fluid grid's sorted order, row-major [N x 64] (bf16 arrays as uint16
 * bit patterns).  Weights are FP32 [K x 64]; sphdem_mpnn_pack writes their
 * bf16 MFMA B-operand fragments for W (Bp) and W^T (BTp), each
 * (K/16)*4*64*4 uint16.  FP32 grids and params only. */
int sphdem_mpnn_pack(const float *W, int K, void *Bp, void *BTp, void *stream);
int sphdem_mpnn_to_bf16(const float *x, int n, void *y, void *stream);
/* Z = X W + b, X = [X0 (K0 wide) | X1 (K1 wide)] bf16.  act nonzero: Yf
 * (FP32) receives Z, the pre-activation the backward reads, and Yb (bf16)
 * silu(Z) (resid unused).  act zero: Y = resid + Z goes to Yf and/or Yb.
 * b, resid, X1, Yf, Yb may be NULL. */
int sphdem_mpnn_dense_fwd(const void *X0, int K0, const void *X1, int K1, int rows,
                          const void *Bp, const float *b, int act, const float *resid,
                          float *Yf, void *Yb, void *stream);
/* dX (+)= dZ W^T with dZ = dY * silu'(Ymask), Ymask the layer's stored
 * pre-activation (NULL: no activation, dZ = dY). */
int sphdem_mpnn_dense_bwd_x(const float *dY, const float *Ymask, int rows, const void *BTp,
                            float *dX0, int K0, float *dX1, int K1, int accumulate,
                            void *stream);
/* gW += X^T dZ, gb += column sums of dZ (gb may be NULL). */
int sphdem_mpnn_dense_bwd_w(const float *dY, const float *Ymask, const void *X0, int K0,
                            const void *X1, int K1, int rows, float *gW, float *gb,
                            void *stream);
/* s_i = (1/64) sum_j silu(silu(P_i + Q_j + e_ij C + c1) W2 + c2) over the
 * fluid and wall neighbours j of fluid node i (1e-12 < r^2 < support^2);
 * Q_j = Qw for walls.  fvel: [N x 4] sorted fluid velocity (from
 * sphdem_mpnn_features); wvel: [nb x 4] sorted wall velocity or NULL. */
int sphdem_mpnn_edge_fwd(const sphdem_params_t *p, const sphdem_grid_t *gs,
                         const sphdem_grid_t *gb, const void *fvel, const void *wvel,
                         const float *P, const void *Q, const void *Qw, const void *Cp,
                         const float *c1, const void *W2p, const float *c2, float *s,
                         void *s_bf, void *stream);
/* Backward of the edge pass for ds = dL/ds: writes dP, adds into dQ (zeroed
 * by the caller) and into the parameter gradients gC, gc1, gW2, gc2, gQw. */
int sphdem_mpnn_edge_bwd(const sphdem_params_t *p, const sphdem_grid_t *gs,
                         const sphdem_grid_t *gb, const void *fvel, const void *wvel,
                         const float *P, const void *Q, const void *Qw, const void *Cp,
                         const float *c1, const void *W2p, const void *W2Tp, const float *c2,
                         const float *ds, float *dP, float *dQ, float *gC, float *gc1,
                         float *gW2, float *gc2, float *gQw, void *stream);
/* Node features [N x 16] bf16: v/U, a/A, 100 (rho/rho0 - 1) (U = 1 m/s,
 * A = 9.81 m/s^2) and the sorted velocity [N x 4]. */
int sphdem_mpnn_features(const sphdem_params_t *p, const sphdem_sph_t *P, const sphdem_grid_t *gs,
                         void *nfeat, void *fvel, void *stream);
/* Training target [N x 6] (sorted): ((v1 - v0) / (A T),
 * (x1 - x0 - T v0) / (A T^2 / 2)); arrays at original index (a0 unused). */
int sphdem_mpnn_targets(const sphdem_grid_t *gs, const float *x0, const float *v0,
                        const float *a0, const float *x1, const float *v1, double T,
                        float *target, void *stream);
/* The outputs of one explicit step of length T from a0 -- the method without
 * the network -- in the network's output layout: (a0 / A, a0 / A). */
int sphdem_mpnn_explicit_out(const sphdem_grid_t *gs, const float *a0, float *out,
                             void *stream);
/* loss += mean of (out[:, 0:6] - target)^2 (FP64, device), dout = its gradient. */
int sphdem_mpnn_loss(const float *out, const float *target, int n, double *loss, float *dout,
                     void *stream);
/* One surrogate call: v += A T out[0:3], x += T v + A T^2 out[3:6] / 2. */
int sphdem_mpnn_apply(const sphdem_params_t *p, sphdem_sph_t *P, const sphdem_grid_t *gs,
                      const float *out, double T, void *stream);
/* Start of a surrogate call: the fused SPH passes of step `step` (Shepard
 * cadence, density, EOS, acceleration, -grad p), the DEM coupling sums of
 * Kernel_interaction_dem and the drag reaction added to P.acc (with the
 * fluid porosity of the last plain step).  D may be NULL (no DEM). */
int sphdem_surrogate_prep(const sphdem_params_t *p, sphdem_sph_t *P, sphdem_bnd_t *B,
                          sphdem_dem_t *D, const sphdem_grid_t *gs, const sphdem_grid_t *gb,
                          const sphdem_grid_t *gd, int step, const sphdem_work_t *w,
                          void *stream);
/* n_sub DEM substeps of a surrogate call, all in one workgroup: contacts
 * (DEM-DEM candidates from gd as built at the call's start, queried with the
 * radius widened by `skin` and read at their current positions; DEM-wall
 * from the static gw), the drag and pressure force held from the call's
 * start, and the integration of Kernel_step_dem.  The contact histories go
 * back to the old and new CSR as stage -1 of sphdem_coupled_step leaves them.
 * `scratch` holds sphdem_dem_substeps_scratch(n, precision) bytes.  flags
 * bit 0: the CSR capacity was exceeded; bit 1: a DEM moved farther than
 * `skin` from where gd was built (the candidate sets may then be short);
 * bit 2: a DEM had more than 64 contacts of one kind. */
long sphdem_dem_substeps_scratch(int n, int precision);
int sphdem_dem_substeps(const sphdem_params_t *p, sphdem_dem_t *D, sphdem_dembnd_t *W,
                        const sphdem_grid_t *gd, const sphdem_grid_t *gw, int n_sub, double skin,
                        void *scratch, const sphdem_work_t *w, void *stream);
/* Adam step `step` (>= 1) on FP32 parameters. */
int sphdem_adam(float *param, const float *grad, float *m, float *v, int n, double lr,
                double beta1, double beta2, double eps, int step, void *stream);

/* ---- Eulerian ISPH lid-driven cavity ------------------------------------ */
/* input/struct_eisph.py:4-28 EISPHptl: fixed fluid or ghost points.  vec3
 * fields [n*3]; kgc [n*4] the row-major 2x2 correction; mirror int32 [n]
 * (a ghost's stable fluid row).  residual[0] holds the right-hand-side sum
 * after the pressure assembly, as in the reference. */
typedef struct {
    void *pos, *vel, *vel_star, *vel_bc, *acc, *kgc, *rho, *m, *pres, *Aij, *bi,
         *Aijpj, *divergence, *residual;
    int *mirror;
    int n;
} sphdem_eisph_t;

/* Solver constants of the cavity step (input/Config.py:11-36: h, rho0, nu,
 * dt; support = 2h). */
typedef struct {
    int precision;   /* 4 = FP32, 8 = FP64 */
    double h, support, rho0, nu, dt;
} sphdem_eisph_params_t;

/* Fixed neighbour list of the fluid points (no reference counterpart: the
 * reference re-walks both HashGrids in every pass, KERNEL_EISPH_*.py; the
 * points never move, so the walk is done once).  Row i lists, in the
 * HashGrid visiting order, the fluid and then the ghost neighbours j of
 * fluid point i with 1e-12 < |r_ij|^2 < support^2: id = j for a fluid
 * point, -1 - bj for ghost bj, and geo = (gx, gz, A_ij, V_j), the
 * KGC-corrected grad_i W_ij (KERNEL_EISPH_KNL.py:52-64), the Laplacian
 * weight 2 V_j (r_ij . g)/r^2 (KERNEL_EISPH_KNL.py:67-71) and V_j = m/rho.
 * Layout: slices of 64 consecutive rows (one wavefront), entry k of row i
 * at offset[i / 64] + 64 k + i % 64, so the 64 lanes of a wavefront read
 * entry k of their rows from consecutive addresses; a slice is as wide as
 * its longest row (slots past len[i] are never read).
 * The list holds the kgc, m and rho of its build: rebuild it after changing
 * any of them or the positions. */
typedef struct {
    int n;           /* fluid points of the build */
    int *len;        /* [n+1] entries of each row */
    int *offset;     /* [slices+1] first slot of each slice; offset[slices] = slots */
    int *id;         /* [cap] */
    void *geo;       /* [cap*4] REAL */
    int cap;
    int *width;      /* [slices+1] build scratch */
    int *scan_tmp;   /* [sphdem_scan_scratch(slices+1)] */
    double *part;    /* [sphdem_eisph_part_size(n_fluid, n_ghost)]: the FP64
                        right-hand-side sum, the raw bi[0], per-workgroup sums */
} sphdem_eisph_nbr_t;
/* Slices of the list of n fluid points: (n + 63) / 64. */
int sphdem_eisph_slices(int n);
long sphdem_eisph_part_size(int n_fluid, int n_ghost);

/* KERNEL_EISPH_KNL.py:74-132 Kernel_prepare_kgc: P->kgc from both grids
 * (built from P->pos and B->pos with radius = support). */
int sphdem_eisph_kgc(const sphdem_eisph_params_t *p, sphdem_eisph_t *P, const sphdem_eisph_t *B,
                     const sphdem_grid_t *gs, const sphdem_grid_t *gb, void *stream);
/* Neighbour list build, in two calls: _count writes L->len and L->offset;
 * the caller reads offset[slices] back, sizes id/geo to cap >= that, then
 * _fill writes id/geo (it never writes a slot at or beyond cap). */
int sphdem_eisph_nbr_count(const sphdem_eisph_params_t *p, const sphdem_eisph_t *P,
                           const sphdem_eisph_t *B, const sphdem_grid_t *gs,
                           const sphdem_grid_t *gb, sphdem_eisph_nbr_t *L, void *stream);
int sphdem_eisph_nbr_fill(const sphdem_eisph_params_t *p, const sphdem_eisph_t *P,
                          const sphdem_eisph_t *B, const sphdem_grid_t *gs,
                          const sphdem_grid_t *gb, sphdem_eisph_nbr_t *L, void *stream);
/* KERNEL_EISPH_BC.py:6-31 Kernel_Dirichlet_BC: B->vel (predicted == 0) or
 * B->vel_star = 2 vel_bc - (fluid field)[mirror]. */
int sphdem_eisph_dirichlet(int precision, const sphdem_eisph_t *P, sphdem_eisph_t *B,
                           int predicted, void *stream);
/* KERNEL_EISPH_BC.py:34-47 Kernel_pressure_boundary_eisph: B->pres = P->pres[mirror]. */
int sphdem_eisph_pressure_boundary(int precision, const sphdem_eisph_t *P, sphdem_eisph_t *B,
                                   void *stream);
/* The pass kernels read the ghost fields as stored in B, as the reference's
 * kernels do.  KERNEL_EISPH_force.py:12-72 Kernel_advec_vis_eisph. */
int sphdem_eisph_advec_vis(const sphdem_eisph_params_t *p, sphdem_eisph_t *P,
                           const sphdem_eisph_t *B, const sphdem_eisph_nbr_t *L, void *stream);
/* KERNEL_EISPH_PPE.py:12-79 Kernel_build_ppe_eisph, plus the FP64 sum of
 * the raw bi (EISPH.py:72) into L->part and P->residual[0]. */
int sphdem_eisph_build_ppe(const sphdem_eisph_params_t *p, sphdem_eisph_t *P,
                           const sphdem_eisph_t *B, const sphdem_eisph_nbr_t *L, void *stream);
/* KERNEL_EISPH_PPE.py:82-125 Kernel_remove_bi_mean + Kernel_solve_ppe_eisph,
 * from the sums of the last sphdem_eisph_build_ppe on L. */
int sphdem_eisph_solve_ppe(const sphdem_eisph_params_t *p, sphdem_eisph_t *P,
                           const sphdem_eisph_nbr_t *L, void *stream);
/* KERNEL_EISPH_step.py:12-61 Kernel_correct_eisph. */
int sphdem_eisph_correct(const sphdem_eisph_params_t *p, sphdem_eisph_t *P,
                         const sphdem_eisph_t *B, const sphdem_eisph_nbr_t *L, void *stream);
/* KERNEL_EISPH_step.py:64-121 Kernel_diagnostics_eisph. */
int sphdem_eisph_diagnostics(const sphdem_eisph_params_t *p, sphdem_eisph_t *P,
                             const sphdem_eisph_t *B, const sphdem_eisph_nbr_t *L, void *stream);
/* source/EISPH.py:26-90 EISPH_OneStep in five launches: the ghost values
 * (2 vel_bc - u[mirror], p[mirror]) are formed where they are read, and the
 * ghost fields of B are published by the pass after their source is final,
 * so every field ends the step as the reference's ten launches leave it. */
int sphdem_eisph_step(const sphdem_eisph_params_t *p, sphdem_eisph_t *P, sphdem_eisph_t *B,
                      const sphdem_eisph_nbr_t *L, int diagnostics, void *stream);
/* source/EISPH.py:131-139, the run_eisph loop between output steps: `steps`
 * calls of sphdem_eisph_step, diagnostics (if asked) on the last one only,
 * with the same results field for field.  A cloud for which
 * sphdem_eisph_one_group returns 1 runs all the steps in one launch of one
 * workgroup (fields in LDS, barriers between the passes); larger clouds take
 * the five-launch step `steps` times.  The sum buffer L->part ends as the
 * last step leaves it. */
int sphdem_eisph_steps(const sphdem_eisph_params_t *p, sphdem_eisph_t *P, sphdem_eisph_t *B,
                       const sphdem_eisph_nbr_t *L, int steps, int diagnostics, void *stream);
/* 1 when sphdem_eisph_steps runs n_fluid points and n_ghost ghosts of this
 * precision (4 or 8) in one workgroup: its LDS share (per point 11 values and
 * 2 ints, per ghost 3 values and an int) fits in 60 KiB. */
int sphdem_eisph_one_group(int precision, int n_fluid, int n_ghost);

/* ---- ASCII output (host only) ------------------------------------------ */
/* The VTK XML frame writer's number formatting (output/output.py:14-24):
 * "%.6g" floats / "%d" integers, per_line values on a line, single spaces,
 * '\n' between lines, no trailing newline.  Return the bytes written (without
 * the NUL) or -1 when `cap` is too small. */
long sphdem_format_g6(const double *v, long n, int per_line, char *out, long cap);
long sphdem_format_int(const int *v, long n, int per_line, char *out, long cap);

#ifdef __cplusplus
}
#endif
#endif
