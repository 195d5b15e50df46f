/*
 * CPU oracle for the MLS-MPM particle-grid transfer loop.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this code, and only as the checker or
 * the timed CPU baseline.  The product library (physgrad_amd) never calls it.
 *
 * Three parts:
 *   1. ref-compat step (float): a restatement of the reference CPU solver
 *      MPMSolver<float,64>::step (src/material_point_method.h:479-666) with the
 *      same loop nesting and floating-point operation order, so its results are
 *      bit-identical to the reference compiled with -ffp-contract=off.
 *   2. MLS-MPM step (double and float): APIC transfers, deformation-gradient
 *      update, fixed-corotated and neo-Hookean elasticity, sticky walls, plus the
 *      hand-derived reverse (adjoint) step in double.  The reference has no
 *      executable MLS-MPM, so this part is pinned by closed forms, invariants and
 *      central finite differences (tests/test_oracle_mls.py).
 *   3. cell keys + stable counting sort: a restatement of
 *      SpatialHashGrid::mortonEncode3D/getCellID/buildHashTable
 *      (src/sparse_data_structures.h:121-190).
 *   4. neighbour lists of the hash (getNeighbors, :193-228) and the SPH-MPM
 *      coupling pairs and forces (src/sph_mpm_coupling.h:454-545).
 */
#ifndef PHYSGRAD_ORACLE_H
#define PHYSGRAD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- ref-compat ---------------- */
typedef struct {
    int nx, ny, nz;
    float dx, dy, dz;
    float ox, oy, oz;
    float dt;
    float gx, gy, gz;
    int order; /* 1 linear, 2 quadratic, 3 cubic (reference SimulationParams::interpolation_order) */
} orc_ref_params;

/* Runs nsteps reference steps in place.  x, v: [3n] xyz-interleaved; m: [n].
 * grid_mass [G], grid_mom [3G], grid_vel [3G] receive the grid of the last step
 * (node index k*nx*ny + j*nx + i, as MPMGrid::getNodeIndex). */
void orc_ref_run(const orc_ref_params* p, int n, float* x, float* v, const float* m,
                 int nsteps, float* grid_mass, float* grid_mom, float* grid_vel);

/* ---------------- MLS-MPM ---------------- */
enum { ORC_FIXED_COROTATED = 0, ORC_NEO_HOOKEAN = 1, ORC_FLUID = 2 };

typedef struct {
    int nx, ny, nz;       /* grid nodes per axis */
    double dx;            /* cell size (isotropic) */
    double ox, oy, oz;    /* grid origin */
    double dt;
    double gx, gy, gz;
    int bc;               /* sticky wall thickness in nodes */
    int material;         /* ORC_FIXED_COROTATED or ORC_NEO_HOOKEAN */
    double nu;            /* Poisson ratio */
    double p_vol;         /* initial particle volume V0 */
} orc_mls_params;

/* State: x,v [3n]; C,F [9n] row-major; mass [n]; E [n] (Young's modulus).
 * grid_m [G], grid_v [3G] (optional, may be NULL) receive the last step's grid
 * mass and post-update grid velocity. */
void orc_mls_run_f64(const orc_mls_params* p, int n, double* x, double* v, double* C, double* F,
                     const double* mass, const double* E, int nsteps,
                     double* grid_m, double* grid_v);
void orc_mls_run_f32(const orc_mls_params* p, int n, float* x, float* v, float* C, float* F,
                     const float* mass, const float* E, int nsteps,
                     float* grid_m, float* grid_v);

/* Kirchhoff stress tau = P(F) F^T for one particle (double). */
void orc_kirchhoff_f64(int material, double mu, double lam, const double* F, double* tau);
/* Polar factor R of F (F = R S). Returns the Newton iteration count. */
int orc_polar_f64(const double* F, double* R, double* S);

/* One reverse step.  Inputs: the state (x,v,C,F) at the START of the step and
 * the adjoints of the state at the END of the step (ax, av, aC, aF).  On return
 * ax..aF hold the adjoints of the state at the start of the step and aE
 * accumulates dL/dE. */
void orc_mls_adjoint_step_f64(const orc_mls_params* p, int n,
                              const double* x, const double* v, const double* C, const double* F,
                              const double* mass, const double* E,
                              double* ax, double* av, double* aC, double* aF, double* aE);

/* The FP64 step with per-particle materials and per-node boundary flags
 * (mlsmpm_ext.c): model [n] (ORC_*), nu, visc, vol [n] (NULL: the scene's
 * material / Poisson ratio / 0 / p_vol), flags [G] node BoundaryType bits
 * (NULL: none).  Identical to orc_mls_run_f64 when all of them are NULL. */
void orc_mls_run_ext_f64(const orc_mls_params* p, int n, double* x, double* v, double* C, double* F,
                         const double* mass, const double* E, const int* model, const double* nu,
                         const double* visc, const double* vol, const unsigned* flags, int nsteps,
                         double* grid_m, double* grid_v);
/* tau of one particle for any model: the Newtonian fluid gives J visc (C + C^T) */
void orc_kirchhoff_ext_f64(int model, double mu, double lam, double visc, const double* F, const double* C,
                           double* tau);

/* FP32 MLS-MPM on all host cores (OpenMP; nthreads <= 0: the OpenMP default);
 * the CPU baseline timed beside the GPU path */
void orc_mls_run_f32_mt(const orc_mls_params* p, int n, float* x, float* v, float* C, float* F,
                        const float* mass, const float* E, int nsteps, int nthreads);
int orc_num_threads(void);

/* ---------------- reference launchers (float4 particles) ---------------- */
void orc_launch_p2g(const float* pos, const float* vel, float* gm, float* gp, const float* o, const float* h,
                    const int* d, int n, int order);
void orc_launch_g2p(float* pos, float* vel, const float* gm, const float* gv, const float* o, const float* h,
                    const int* d, int n, float dt, float flip, int order);
void orc_launch_fused(float* pos, float* vel, float* gm, float* gp, float* gv, const float* o, const float* h,
                      const int* d, int n, float dt, const float* grav, float flip, int order);
void orc_launch_g2p2g(float* pos, float* vel, float* gm, float* gp, float* gv, const float* o, const float* h,
                      const int* d, int n, float dt, const float* grav, int use_affine_pic, float flip, int order);

/* ---------------- cell keys + counting sort ---------------- */
unsigned orc_morton3(unsigned x, unsigned y, unsigned z);
/* particle cell id = morton(clamped truncated coords) % total_cells */
void orc_cell_ids(const float* pos, unsigned n, float cell_size, const float* domain_min,
                  const unsigned* res, unsigned* ids);
/* stable counting sort (count, exclusive scan, scatter in index order) */
void orc_counting_sort(const unsigned* ids, unsigned n, unsigned total_cells,
                       unsigned* particle_indices, unsigned* cell_start, unsigned* cell_count);

/* ---------------- neighbour lists and SPH-MPM coupling (coupling.c) ---------------- */
/* SpatialHashGrid::getNeighbors for nq queries over a built hash: counts [nq]
 * (may be NULL), out the lists back to back (NULL: count only).  Returns the
 * total. */
unsigned long long orc_hash_neighbors(const float* queries, unsigned nq, float radius, float cell_size,
                                      const float* dmin, const unsigned* res, const unsigned* particle_indices,
                                      const unsigned* cell_start, const unsigned* cell_count, unsigned* counts,
                                      unsigned* out);
/* findCouplingPairs: row_counts [ns] (may be NULL); col / penetration the
 * pairs in SPH-major, MPM-ascending order (NULL: count only).  Returns the
 * pair count. */
unsigned long long orc_coupling_pairs(const float* sph_x, unsigned ns, const float* mpm_x, unsigned nm, float radius,
                                      unsigned* row_counts, unsigned* col, float* penetration);
/* computeCouplingForces over CSR pairs (row_ptr [ns+1]): solid_force,
 * fluid_force [3 nnz]; two_way adds each fluid force onto sph_force [3 ns]. */
void orc_coupling_forces(const float* sph_x, const float* sph_v, unsigned ns, const float* mpm_x,
                         const float* mpm_v, const unsigned* row_ptr, const unsigned* col, const float* penetration,
                         float stiffness, float damping, float friction, int two_way, float* solid_force,
                         float* fluid_force, float* sph_force);

/* SparseMatrix::multiply over a CSR matrix: y [nrows] */
void orc_csr_spmv(const unsigned* row_ptr, const unsigned* col, const float* val, const float* x, unsigned nrows,
                  float* y);

#ifdef __cplusplus
}
#endif
#endif
