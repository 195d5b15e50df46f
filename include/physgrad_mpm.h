/*
 * physgrad_mpm.h - C-ABI of the MI355X MLS-MPM transfer loop.
 *
 * Every entry point returns an int status (PG_OK = 0); pg_mpm_last_error()
 * gives the message of the last failure on a context (or, for NULL, of the
 * last failure on this host thread).  Pointers are plain host or device
 * pointers as the `memory` argument says; streams are hipStream_t passed as
 * void* (NULL = the null stream).  A context is single-stream and not
 * thread-safe; distinct contexts are independent.
 *
 * Reference interfaces these entry points replace (pedroscortes/physgrad):
 *   pg_mpm_create / pg_mpm_destroy
 *       MPMSolver<T,64>::MPMSolver(grid_dims, cell_size, origin, SimulationParams)
 *           src/material_point_method.h:450-452
 *       MPMSolver<T>::MPMSolver(const MPMSolverConfig&) / initialize() / cleanup()
 *           src/mpm_solver.h:155-187, :506-532
 *   pg_mpm_set_particles
 *       MPMSolver<T,64>::addParticles   src/material_point_method.h:454-477
 *       MPMSolver<T>::addParticles      src/mpm_solver.h:190-225
 *   pg_mpm_step
 *       MPMSolver<T,64>::step           src/material_point_method.h:479-500
 *       MPMSolver<T>::step              src/mpm_solver.h:228-285
 *       G2P2GKernelLauncher<T>::launchG2P2G  src/mpm_g2p2g_kernels.h:494-515
 *   pg_mpm_get_particles / pg_mpm_get_grid / pg_mpm_get_time
 *       getParticles/getGrid/getCurrentTime/getStepCount
 *           src/material_point_method.h:503-506, src/mpm_solver.h:323-328
 *   pg_mpm_forward_checkpointed / pg_mpm_backward
 *       DifferentiableMPMSystem<T>::simulationStep / computeGradients
 *           src/differentiable_mpm.h:748-783 (its backward is empty; this is the
 *           working replacement), and the torch-side launchers
 *           launch_mpm_timestep_forward/backward src/pytorch_autograd.cu:371-402
 *   pg_spatial_hash_build
 *       SpatialHashGrid<T,uint32_t>::buildHashTable src/sparse_data_structures.h:155-190
 *   pg_hash_neighbors_count / pg_hash_neighbors_fill
 *       SpatialHashGrid<T,uint32_t>::getNeighbors   src/sparse_data_structures.h:193-228
 *   pg_coupling_create / pg_coupling_destroy
 *       SPHMPMCouplingSystem<T>::SPHMPMCouplingSystem(const CouplingConfig<T>&)
 *           src/sph_mpm_coupling.h:391-427
 *   pg_coupling_find_pairs / pg_coupling_get_pairs
 *       SPHMPMCouplingSystem<T>::findCouplingPairs    src/sph_mpm_coupling.h:454-486
 *   pg_coupling_compute_forces
 *       SPHMPMCouplingSystem<T>::computeCouplingForces src/sph_mpm_coupling.h:488-545
 *   pg_csr_spmv
 *       SparseMatrix<T,IndexType>::multiply src/sparse_data_structures.h:392-410
 */
#ifndef PHYSGRAD_MPM_H
#define PHYSGRAD_MPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PG_OK 0
#define PG_ERR_INVALID 1   /* bad argument or size */
#define PG_ERR_HIP 2       /* HIP runtime error */
#define PG_ERR_NOMEM 3     /* device allocation failed or over budget */
#define PG_ERR_STATE 4     /* call out of order (e.g. backward without forward) */

enum pg_mode { PG_MODE_REF_COMPAT = 0, PG_MODE_MLS_MPM = 1 };
enum pg_material { PG_MAT_FIXED_COROTATED = 0, PG_MAT_NEO_HOOKEAN = 1, PG_MAT_FLUID = 2 };
enum pg_impl { PG_IMPL_PLAIN = 0, PG_IMPL_TILED = 1 };
enum pg_memory { PG_HOST = 0, PG_DEVICE = 1 };

typedef struct pg_mpm_ctx pg_mpm_ctx;

/* Scene and solver parameters: a superset of the reference's
 * MPMSolver::SimulationParams (src/material_point_method.h:440-448) and
 * MPMSolverConfig (src/mpm_solver.h:29-66). */
typedef struct pg_mpm_params {
    int mode;                 /* pg_mode */
    int impl;                 /* pg_impl (MLS-MPM only) */
    int nx, ny, nz;           /* grid nodes per axis */
    float dx, dy, dz;         /* cell size; MLS-MPM needs dx == dy == dz */
    float ox, oy, oz;         /* grid origin */
    float dt;                 /* time step */
    float gx, gy, gz;         /* gravity */
    int interpolation_order;  /* ref-compat: 1 linear, 2 quadratic, 3 cubic */
    int material;             /* pg_material (MLS-MPM) */
    float youngs_modulus;     /* E, used when no per-particle E is given */
    float poisson_ratio;      /* nu */
    float density;            /* rho: default particle mass = rho * particle_volume */
    float particle_volume;    /* V0 */
    int boundary_thickness;   /* sticky wall thickness in nodes (MLS-MPM) */
    int max_particles;        /* capacity */
} pg_mpm_params;

/* Defaults reproduce the reference's SimulationParams: dt 1e-4, g (0,-9.81,0),
 * quadratic interpolation, 64^3 nodes of 1/64, E 1e6, nu 0.3, rho 1000. */
void pg_mpm_default_params(pg_mpm_params* p);

int pg_mpm_create(const pg_mpm_params* p, pg_mpm_ctx** out);
int pg_mpm_destroy(pg_mpm_ctx* ctx);
const char* pg_mpm_last_error(const pg_mpm_ctx* ctx);

/* Replace the particle set.  x, v: [3n] xyz-interleaved; C, F: [9n] row-major
 * (NULL: C = 0, F = I); mass: [n] (NULL: density * particle_volume);
 * E: [n] per-particle Young's modulus (NULL: params.youngs_modulus).
 * ref-compat mode uses x, v and mass only. */
int pg_mpm_set_particles(pg_mpm_ctx* ctx, int n, const float* x, const float* v, const float* C, const float* F,
                         const float* mass, const float* E, int memory, void* stream);
/* Read the particle set back in the caller's original particle order.  Any
 * pointer may be NULL. */
int pg_mpm_get_particles(pg_mpm_ctx* ctx, float* x, float* v, float* C, float* F, int memory, void* stream);
int pg_mpm_num_particles(const pg_mpm_ctx* ctx);
/* Grid of the last step: mass [G], momentum [3G] (ref-compat and plain MLS;
 * pre-update momentum), velocity [3G] (post-update).  Any pointer may be NULL. */
int pg_mpm_get_grid(pg_mpm_ctx* ctx, float* mass, float* momentum, float* velocity, int memory, void* stream);
/* Advance nsteps steps on the stream.  On the tiled MLS-MPM path the last
 * step of a call also scatters the new state to the grid windows and counts
 * its binning, which the next pg_mpm_step call starts from (so that stepping
 * one step per call costs what a step inside a long call does); the calls
 * that replace or re-simulate the particles complete that work first. */
int pg_mpm_step(pg_mpm_ctx* ctx, int nsteps, void* stream);
/* Simulated time and step count (reference getCurrentTime/getStepCount). */
int pg_mpm_get_time(const pg_mpm_ctx* ctx, double* time, long long* steps);
/* Sets the clock, for resuming from a checkpoint (call after set_particles,
 * which resets it).  No reference counterpart: the reference solver has no
 * resume (its outputData, src/mpm_solver.h:501-504, only prints). */
int pg_mpm_set_time(pg_mpm_ctx* ctx, double time, long long steps);

/* Differentiable rollout (MLS-MPM, tiled or plain implementation).
 * pg_mpm_forward_checkpointed advances nsteps steps like pg_mpm_step and keeps
 * the state every checkpoint_every steps in device memory (the reference's
 * CheckpointStrategy::UNIFORM, src/differentiable_mpm.h:63-68).
 * pg_mpm_backward then runs the exact reverse of those steps, segment by
 * segment from the last checkpoint: each segment is recomputed into a tape,
 * then reversed (G2P, grid update with sticky walls, P2G with elasticity).
 * Inputs: adjoints dL/d(final x, v, C, F) in the caller's particle order
 * ([3n],[3n],[9n],[9n]; NULL = zero).  Outputs: dL/d(initial x, v, C, F) and
 * dL/dE per particle ([n]); NULL outputs are skipped (with gE NULL the E
 * adjoint is not computed at all).  Arrays may be in host
 * or device memory; when any is host memory the call returns after the
 * outputs are written, otherwise it is asynchronous on the stream.  The
 * reference's computeGradients (src/differentiable_mpm.h:772-783) reverses an
 * empty tape; this is the working replacement. */
int pg_mpm_forward_checkpointed(pg_mpm_ctx* ctx, int nsteps, int checkpoint_every, void* stream);
int pg_mpm_backward(pg_mpm_ctx* ctx, const float* ax, const float* av, const float* aC, const float* aF, float* gx,
                    float* gv, float* gC, float* gF, float* gE, void* stream);

/* Per-kernel timing with HIP events recorded on the launch stream around each
 * kernel (replaces the cudaEventRecord timing of MPMSolver::step,
 * src/mpm_solver.h:236-284, and its MPMPerformanceMetrics :72-101).
 * pg_mpm_set_profiling resets the totals.  Kernel ids: 0 tiled P2G, 1 tiled
 * fused G2P+P2G, 2 tiled G2P, 3 grid update, 4 binning, 5 plain P2G, 6 plain
 * G2P, 7 grid clear, 8 ref-compat step, 9 reverse G2P, 10 reverse grid
 * update, 11 reverse P2G, 12 reverse P2G of a step fused with the reverse G2P
 * of the step before it (the tiled backward runs 9 and 11 once per segment
 * and 12 for the steps between). */
int pg_mpm_set_profiling(pg_mpm_ctx* ctx, int on);
int pg_mpm_kernel_time(pg_mpm_ctx* ctx, int kernel, double* total_ms, long long* launches);

/* State validation.  The reference steps on whatever its particles hold and
 * has no failure detection on this path (MPMSolver<T>::step,
 * src/mpm_solver.h:228-285); these calls add it.
 * pg_mpm_check_state scans the particle state: *first_nonfinite receives the
 * lowest (caller-order) index of a particle whose x, v, C or F is not finite
 * and *first_outside the lowest index of a particle outside its grid box
 * [origin, origin + (n-1) dx] per axis (-1: none; either pointer may be NULL).
 * It also reads the tiled path's binning error word, which the binning kernels
 * set instead of writing past a buffer.  Returns PG_ERR_STATE, with a message
 * naming the first bad particle, for non-finite state or a binning error;
 * particles outside the grid are reported but are not an error (the tiled
 * path clamps their tile and skips out-of-grid stencil nodes).  Synchronises
 * the stream.  pg_mpm_set_validation(ctx, 1) makes every pg_mpm_step call end
 * with this check (and so synchronise). */
int pg_mpm_check_state(pg_mpm_ctx* ctx, int* first_nonfinite, int* first_outside, void* stream);
int pg_mpm_set_validation(pg_mpm_ctx* ctx, int on);

/* Per-particle materials (MLS-MPM, tiled or plain, forward and adjoint).
 * Replaces the reference's MaterialDatabase lookup per particle
 * (src/mpm_data_structures.h:539-612; MaterialParameters :63-104; the kernel's
 * per-particle model switch src/mpm_kernels.cu:295, :405-421).  table[m]
 * describes material m; ids [n] (the caller's particle order, host or device
 * memory) gives each particle's material, all < nmat <= 256.  A particle then
 * uses its material's model, Poisson ratio, viscosity and Young's modulus
 * (a per-particle E given to set_particles takes precedence) and, when the
 * material has a density, the initial volume mass / density (else the
 * context's particle_volume).  PG_MAT_FLUID is the reference's Newtonian
 * fluid (src/mpm_kernels.cu:131-149): Cauchy stress 2 viscosity sym(grad v),
 * grad v = the particle's APIC C, entering P2G like the elastic stresses as
 * J sigma.  table NULL or nmat 0 restores the scenes' materials.
 * set_particles resets to the scenes' materials, so call this after it. */
typedef struct pg_material_desc {
    int model;              /* pg_material */
    float youngs_modulus;   /* E (elastic models) */
    float poisson_ratio;    /* nu (elastic models), in (-1, 0.5) */
    float density;          /* rho > 0: particle volume = mass / rho; 0: the context's particle_volume */
    float viscosity;        /* dynamic viscosity (PG_MAT_FLUID) */
} pg_material_desc;
int pg_mpm_set_materials(pg_mpm_ctx* ctx, const pg_material_desc* table, int nmat, const uint8_t* ids, int memory,
                         void* stream);

/* Per-node boundary flags (MLS-MPM, tiled or plain, forward and adjoint):
 * flags [G] (node index (k ny + j) nx + i; the concatenated grids of a batch)
 * with the reference's BoundaryType bits (src/mpm_data_structures.h:617-627),
 * applied by the grid update after v = p/m + dt g as the reference's
 * updateGridKernel does (src/mpm_kernels.cu:536-539) and extended to the bits
 * it declares but never applies:
 *   PG_BC_DIRICHLET_X/Y/Z  zero that velocity component
 *   PG_BC_STICKY           zero the whole velocity (no slip)
 *   PG_BC_SLIP             remove each component that points out of the domain
 *                          through the nearer wall of its axis (node index <
 *                          n/2: negative components; otherwise positive ones)
 *   PG_BC_NEUMANN_X/Y/Z    accepted, no effect (none in the reference either)
 * The sticky band of boundary_thickness nodes applies as well.  NULL clears
 * the flags. */
#define PG_BC_DIRICHLET_X 1u
#define PG_BC_DIRICHLET_Y 2u
#define PG_BC_DIRICHLET_Z 4u
#define PG_BC_NEUMANN_X 8u
#define PG_BC_NEUMANN_Y 16u
#define PG_BC_NEUMANN_Z 32u
#define PG_BC_STICKY 64u
#define PG_BC_SLIP 128u
int pg_mpm_set_boundary(pg_mpm_ctx* ctx, const uint32_t* flags, int memory, void* stream);

/* Batched independent scenes in one context (tiled MLS-MPM).  The reference
 * runs one scene per MPMSolver (src/mpm_solver.h:150-171, one grid and one
 * particle AoSoA each); a batch concatenates the scenes' particles and grids
 * and advances all of them with the same launches.  Each scene has its own grid,
 * cell size, origin, gravity, time step, material and elastic constants;
 * pg_mpm_get_time reports scene 0's clock.  A batch context accepts every call
 * a tiled MLS-MPM context
 * does (step, get_particles, forward_checkpointed, backward, graph, autotune);
 * get_grid returns the concatenated grids. */
typedef struct pg_scene_desc {
    int nx, ny, nz;           /* grid nodes per axis */
    float dx;                 /* cell size */
    float ox, oy, oz;         /* grid origin */
    float dt;                 /* time step */
    float gx, gy, gz;         /* gravity */
    int material;             /* pg_material */
    float youngs_modulus;
    float poisson_ratio;
    float density;
    float particle_volume;
    int boundary_thickness;
    int max_particles;        /* this scene's capacity */
} pg_scene_desc;

/* Tiled contexts (single scene or batch) take grids of at most 256 nodes per
 * axis and return PG_ERR_INVALID for larger ones. */
int pg_mpm_batch_create(const pg_scene_desc* scenes, int nscenes, pg_mpm_ctx** out);
/* counts[s] particles of scene s, scenes back to back in the arrays (same
 * layouts as pg_mpm_set_particles); mass NULL uses each scene's
 * density * particle_volume, E NULL each scene's youngs_modulus. */
int pg_mpm_batch_set_particles(pg_mpm_ctx* ctx, const int* counts, const float* x, const float* v, const float* C,
                               const float* F, const float* mass, const float* E, int memory, void* stream);
int pg_mpm_num_scenes(const pg_mpm_ctx* ctx);

/* Replay pg_mpm_step's launches from a captured hipGraph (tiled MLS-MPM).  One
 * graph per step count is captured on first use and replayed afterwards;
 * set_particles, set_launch and autotune drop the captured graphs.  Replaces
 * the reference's per-step kernel launches of MPMSolver::step
 * (src/mpm_solver.h:236-284). */
int pg_mpm_set_graph(pg_mpm_ctx* ctx, int enable);

/* Launch shape of the tiled path: the workgroups of the persistent tile
 * kernels, and the work-item size (a tile's particles are split into work
 * items of at most that many particles, [640, 5120], default 1280: smaller
 * spreads dense tiles over more workgroups, larger costs less per item), and
 * the threads of a tile workgroup (256: 4 waves, 3 waves per SIMD; or 512: 8
 * waves, 2 per SIMD).  The autotuner picks all three by timing trial_steps
 * steps per candidate (workgroups 256..8192 x work items 640, 1280, 2560 x
 * 256 or 512 threads) on the current particle set (the state is restored
 * afterwards).  Results are cached in a text file ($PG_AUTOTUNE_CACHE, else
 * ~/.cache/physgrad_amd_autotune.txt) keyed by tile slots, grid blocks,
 * particle count bucket and scene count; use_cache reads that file first.  The
 * tile edge (4 cells) is compile-time: an 8-cell tile's P2G windows (10^3
 * nodes, 16 KB each, two per wave) do not fit the 160 KB of LDS of a CU at
 * more than one workgroup.  pg_mpm_get_tuning reports all four.  The
 * reference fixes its launch shapes at compile time (KernelConfig,
 * src/mpm_g2p2g_kernels.h:68-75). */
int pg_mpm_get_launch(const pg_mpm_ctx* ctx, int* tile_workgroups);
int pg_mpm_set_launch(pg_mpm_ctx* ctx, int tile_workgroups);
int pg_mpm_set_work_item(pg_mpm_ctx* ctx, int particles);
int pg_mpm_set_block(pg_mpm_ctx* ctx, int threads);
int pg_mpm_get_tuning(const pg_mpm_ctx* ctx, int* tile_workgroups, int* work_item, int* tile_edge, int* block_threads);
int pg_mpm_autotune(pg_mpm_ctx* ctx, int trial_steps, int use_cache, void* stream);

/* The reference's standalone transfer launchers, same arguments and meaning
 * as launchOptimizedP2G / launchOptimizedG2P / launchFullyFusedG2P2G
 * (src/mpm_g2p2g_kernels.cu:506-617; kernels :72-271, :277-500) with the
 * float3/int3 by-value arguments passed as 3-element host arrays.
 *   positions  device float4[n] (x, y, z, mass); velocities float4[n] (vx, vy, vz, 0)
 *   grid_mass  device float[G]; grid_momentum / grid_velocity float3[G]
 *   order      2 (quadratic, 27 nodes) or 3 (cubic, 64 nodes)
 * P2G accumulates into the caller's grid (no clear).  G2P normalises by the
 * weight of nodes with mass > 1e-10 and blends v = flip v_old + (1-flip) v_grid.
 * The fused launch clears the grid, scatters with gravity added to the
 * particle velocity, forms v = p/m and gathers; the reference's cooperative
 * grid.sync() phases become stream-ordered kernels. */
int pg_launch_optimized_p2g(const float* positions, const float* velocities, float* grid_mass, float* grid_momentum,
                            const float* origin, const float* cell, const int* dims, int n, int order, void* stream);
int pg_launch_optimized_g2p(float* positions, float* velocities, const float* grid_mass, const float* grid_velocity,
                            const float* origin, const float* cell, const int* dims, int n, float dt, float flip_ratio,
                            int order, void* stream);
int pg_launch_fully_fused_g2p2g(float* positions, float* velocities, float* grid_mass, float* grid_momentum,
                                float* grid_velocity, const float* origin, const float* cell, const int* dims, int n,
                                float dt, const float* gravity, float flip_ratio, int order, void* stream);

/* G2P2GKernelLauncher<T>::launchG2P2G (src/mpm_g2p2g_kernels.h:494-547) with
 * the header kernels it launches (optimizedP2GKernel :343-413,
 * calculateGridVelocityKernel :419-437, fusedG2P2GKernel :211-335): clear the
 * grid, scatter mass and momentum with gravity added to the particle velocity,
 * v = p/m on nodes with mass > 1e-10 (0 elsewhere), gather the plain weighted
 * sum of node velocities (not normalised), v = flip v_old + (1-flip) v_grid
 * when use_affine_pic is set and v = v_grid otherwise, x += v dt.
 * Arrays as for pg_launch_fully_fused_g2p2g; memory = PG_HOST stages host
 * arrays through device copies and returns after the results are back
 * (the reference passes host containers to this launcher). */
int pg_launch_g2p2g(float* positions, float* velocities, float* grid_mass, float* grid_momentum,
                    float* grid_velocity, const float* origin, const float* cell, const int* dims, int n, float dt,
                    const float* gravity, int use_affine_pic, float flip_ratio, int order, int memory,
                    void* stream);

/* Spatial hash: cell ids (Morton code of the clamped truncated cell coordinate,
 * modulo the cell count), stable sort of particle indices by cell id, per-cell
 * start and count; bit-identical to SpatialHashGrid::buildHashTable.
 * pos: device [3n]; res[a] = ceil((domain_max[a]-domain_min[a])/cell_size).
 * Outputs are device arrays: cell_ids [n], particle_indices [n],
 * cell_start [total], cell_count [total], total = res0*res1*res2.
 * workspace: device scratch of pg_spatial_hash_workspace_size(n) bytes. */
size_t pg_spatial_hash_workspace_size(unsigned n);
int pg_spatial_hash_build(const float* pos, unsigned n, float cell_size, const float* domain_min,
                          const unsigned* res, unsigned* cell_ids, unsigned* particle_indices, unsigned* cell_start,
                          unsigned* cell_count, void* workspace, size_t workspace_bytes, void* stream);

/* pg_spatial_hash_build without a workspace argument and without any device
 * allocation: no counterpart in the reference.  It is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The radix sort's workspace is a
 * library-owned device array of 12.25 MiB per GPU, held until the process
 * ends, sized for at most PG_SPATIAL_HASH_NOALLOC_MAX particles; larger n
 * returns PG_ERR_INVALID.  Same arguments (minus the workspace), results and
 * errors as pg_spatial_hash_build.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (concurrent streams or host threads)
 * race and must be serialised by the caller. */
#define PG_SPATIAL_HASH_NOALLOC_MAX (1u << 20)
int pg_spatial_hash_build_noalloc(const float* pos, unsigned n, float cell_size, const float* domain_min,
                                  const unsigned* res, unsigned* cell_ids, unsigned* particle_indices,
                                  unsigned* cell_start, unsigned* cell_count, void* stream);

/* ---------------------------------------------------------------- neighbour lists
 * SpatialHashGrid::getNeighbors (src/sparse_data_structures.h:193-228) for nq
 * query points at once, over a hash built by pg_spatial_hash_build with the
 * same cell_size, domain_min and res: the (2R+1)^3 cells around each query's
 * cell, R = ceil(radius / cell_size), x outermost and z innermost, cells whose
 * coordinates reach res skipped, all particles of each visited hash slot in
 * sorted order (no distance test and no de-duplication, as the reference).
 * Result in CSR form: row_ptr [nq+1], neighbours [row_ptr[nq]].
 * Cell coordinates saturate where the reference's conversion is undefined
 * (a query below domain_min has coordinate 0 on that axis).
 *   pg_hash_neighbors_count: row_ptr (device), *total = row_ptr[nq] (host;
 *       synchronises the stream); workspace of
 *       pg_hash_neighbors_workspace_size(nq) device bytes.  PG_ERR_INVALID if
 *       the total reaches 2^32.
 *   pg_hash_neighbors_fill: neighbours (device, row_ptr[nq] entries) from the
 *       row_ptr of the count call. */
size_t pg_hash_neighbors_workspace_size(unsigned nq);
int pg_hash_neighbors_count(const float* queries, unsigned nq, float radius, float cell_size, const float* domain_min,
                            const unsigned* res, const unsigned* cell_count, unsigned* row_ptr,
                            unsigned long long* total, void* workspace, size_t workspace_bytes, void* stream);
int pg_hash_neighbors_fill(const float* queries, unsigned nq, float radius, float cell_size, const float* domain_min,
                           const unsigned* res, const unsigned* particle_indices, const unsigned* cell_start,
                           const unsigned* cell_count, const unsigned* row_ptr, unsigned* neighbors, void* stream);

/* CSR sparse matrix-vector product y = A x (SparseMatrix::multiply,
 * src/sparse_data_structures.h:392-410): row_ptr [nrows+1], col and val
 * [row_ptr[nrows]], x [ncols], y [nrows], all device arrays; each row's
 * products are summed in its stored order, in float, as the reference does.
 * A column index >= ncols makes its row NaN (no read outside x). */
int pg_csr_spmv(const unsigned* row_ptr, const unsigned* col, const float* val, const float* x, unsigned ncols,
                float* y, unsigned nrows, void* stream);

/* ---------------------------------------------------------------- SPH-MPM coupling
 * SPHMPMCouplingSystem (src/sph_mpm_coupling.h:391-579) on the GPU.  The
 * reference tests every (SPH, MPM) pair on the host; here both particle sets
 * are sorted by cell with the spatial hash's radix sort (cells of
 * coupling_radius (1 + 2^-8) over [domain_min, domain_max], linear cell ids,
 * so the 27 cells around an SPH particle's cell hold every MPM particle within
 * the radius, also for particles outside the domain, which the binning clamps
 * into its border cells) and each SPH particle tests the candidates of its 27
 * cells.  Results are identical to the reference's:
 *   pairs     all (s, m) with |x_m - x_s| < coupling_radius, in the
 *             reference's order (SPH index major, MPM index ascending), in CSR
 *             form: row_ptr [ns+1], col [npairs] (MPM index),
 *             penetration [npairs] = coupling_radius - distance;
 *   forces    per pair the reference's spring-damper normal force plus
 *             friction, solid_force = -fluid_force [3 npairs]; with
 *             enable_two_way_coupling each fluid force is added to the SPH
 *             particle's force in pair order.  contact_area is
 *             pi coupling_radius^2 for every pair.
 * Positions, velocities and outputs are device arrays ([3n] xyz interleaved).
 * The pairs of the last find_pairs stay in the context. */
typedef struct pg_coupling_config {
    float coupling_stiffness;    /* reference default 1e5 */
    float coupling_damping;      /* 1e3 */
    float friction_coefficient;  /* 0.3 */
    float coupling_radius;       /* 0.15 */
    int enable_two_way_coupling; /* 1 */
    float domain_min[3];         /* binning box of the MPM particles */
    float domain_max[3];
} pg_coupling_config;
typedef struct pg_coupling pg_coupling;
/* Reference defaults (CouplingConfig, src/sph_mpm_coupling.h:391-411) and the unit box. */
void pg_coupling_default_config(pg_coupling_config* cfg);
/* PG_ERR_INVALID for a radius <= 0, an empty box or more than 2^28 binning cells. */
int pg_coupling_create(const pg_coupling_config* cfg, pg_coupling** out);
void pg_coupling_destroy(pg_coupling* c);
const char* pg_coupling_last_error(const pg_coupling* c);
/* findCouplingPairs: bins mpm_x, finds the pairs, *npairs (host) = their
 * count; synchronises the stream.  PG_ERR_INVALID if the count reaches 2^32. */
int pg_coupling_find_pairs(pg_coupling* c, const float* sph_x, unsigned ns, const float* mpm_x, unsigned nm,
                           unsigned long long* npairs, void* stream);
/* Copies the pairs of the last find_pairs into caller arrays (any may be
 * NULL): row_ptr [ns+1], col, penetration [npairs]. */
int pg_coupling_get_pairs(pg_coupling* c, unsigned* row_ptr, unsigned* col, float* penetration, void* stream);
/* computeCouplingForces for the pairs of the last find_pairs, with the current
 * positions and velocities (same particle counts).  solid_force, fluid_force
 * [3 npairs] (may be NULL).  sph_force [3 ns] (in/out) receives the fluid
 * forces when enable_two_way_coupling is set (may be NULL otherwise). */
int pg_coupling_compute_forces(pg_coupling* c, const float* sph_x, const float* sph_v, const float* mpm_x,
                               const float* mpm_v, float* sph_force, float* solid_force, float* fluid_force,
                               void* stream);

#ifdef __cplusplus
}
#endif
#endif
