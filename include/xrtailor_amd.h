/*
 * xrtailor_amd.h -- C ABI of the MI355X (gfx950) cloth-substep solver.
 *
 * The drop-in boundary for XRTailor's per-substep solver.  Each entry point names the
 * reference interface it replaces (paths relative to the XRTailor repository root).
 * Data handles change from the reference's `Node**` / `Edge**` / `Face**` device
 * pointer arrays to an opaque context that owns the device state of one or more cloth
 * instances (an instance = one cloth mesh plus its optional obstacle mesh).
 *
 * Conventions
 *  - Every function returns an int status: XRT_OK (0) or a negative XRT_ERR_* code.
 *    Nothing calls exit(); the last error text is available from xrt_last_error().
 *  - Inputs are host arrays copied in; outputs go to caller-owned host buffers.
 *    Vectors are packed float[3*n] (x,y,z), triangles int32[3*nt].
 *  - All device work is asynchronous on the context's own HIP stream (default
 *    priority).  Getters synchronise that stream.  One context per host thread; there
 *    is no internal locking.
 */
#ifndef XRTAILOR_AMD_H
#define XRTAILOR_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XRT_ABI_VERSION 5

enum {
  XRT_OK = 0,
  XRT_ERR_INVALID_ARG = -1,  /* bad pointer / size / index                        */
  XRT_ERR_HIP = -2,          /* a HIP runtime call failed                         */
  XRT_ERR_STATE = -3,        /* call out of order (e.g. solve before xrt_start)   */
  XRT_ERR_CAPACITY = -4,     /* a fixed-capacity buffer overflowed                */
  XRT_ERR_NOT_FINITE = -5,   /* NaN/Inf detected in the state                     */
  XRT_ERR_UNSUPPORTED = -6   /* requested feature not available on this device    */
};

/* Simulation parameters; mirrors the fields of SimParams the substep reads
 * (xrtailor/include/xrtailor/core/Common.hpp:31-168) plus the scene constants the
 * reference keeps in its config (frame time: Timer::FixedDeltaTime, utils/Timer.hpp:129-138). */
typedef struct xrt_params {
  float gravity[3];              /* SimParams::gravity                         */
  float damping;                 /* SimParams::damping (0.98)                  */
  float relaxation_factor;       /* SimParams::relaxation_factor (0.25)        */
  float long_range_stretchiness; /* SimParams::long_range_stretchiness (1.0)   */
  float bvh_tolerance;           /* SimParams::bvh_tolerance (swift 0.003)     */
  float ground_height;           /* plane y (SDFCollider type 1)               */
  float ground_friction;         /* SimParams::sdf_friction on the plane       */
  float pbd_thickness;           /* self-contact thickness h                   */
  float repulsion_relaxation;    /* PBD repulsion relaxation_rate              */
  float frame_time;              /* seconds per frame (1/60)                   */
  int32_t num_substeps;          /* SimParams::num_substeps                    */
  int32_t num_iterations;        /* SimParams::num_iterations                  */
  int32_t num_collision_passes;  /* SimParams::num_collision_passes            */
  int32_t solve_bending;         /* SimParams::solve_bending                   */
  int32_t enable_ground;
  int32_t enable_obstacle_contact;
  int32_t enable_self_contact;
  int32_t pair_update_interval;  /* iterations between pair refreshes (10)     */
  /* XRT_BEND_REFERENCE: SolveBend_Kernel as written (BasicSolverHelper.cu:8-72): d clamped to
   * [0,1], rest angle from the other normal convention, step along +q.
   * XRT_BEND_CONSISTENT: rest angle pi - theta0 (the solver's own normals), d clamped to
   * [-1,1], PBD step along -q (Mueller 2007).  Same cost; stable for small compliance. */
  int32_t bend_model;
  /* constraint coloring at xrt_start: XRT_COLOR_GREEDY (first fit in constraint order,
   * fewest colors on structured meshes) or XRT_COLOR_PALETTE (GraphColoring::Paint,
   * xrtailor/src/physics/graph_coloring/Base.cu:163-317, made deterministic from the seed) */
  int32_t coloring;
  /* capacity of the repulsion face-pair buffer per cloth face (>= 1; a change on a started
   * context re-allocates the lists).  The reference sizes its pair vector from the count
   * (BVH.cuh:139-152); here a pair count above the capacity leaves the excess pairs
   * unevaluated and makes the next xrt_sync / xrt_time_* call return XRT_ERR_CAPACITY,
   * naming the per-face capacity that would have held them */
  int32_t pair_capacity_per_face;
  /* capacities of the batched path's repulsion test lists (VF and EE items per cloth face,
   * defaults 64 and 128), rebuilt from the pairs at each refresh; an instance whose items
   * exceed a list evaluates that half per pair instead, with the same results.  0 turns
   * the list off.  No counterpart in the reference, which evaluates per pair. */
  int32_t repulsion_vf_items_per_face;
  int32_t repulsion_ee_items_per_face;
  /* swift-mode particle self-collision (SimParams::enable_self_collision,
   * ParticleCollisionSettings, xrtailor/include/xrtailor/config/Base.hpp:59-66): a spatial
   * hash of the predicted positions every `inter_leaved_hash` substeps of a frame, then
   * CollideParticles at the start of every iteration.  Each instance's particle diameter is
   * particle_diameter x the length of its first triangle's first edge (ClothObject.hpp:
   * 36-37); hash cells are diameter x hash_cell_size wide; friction is ground_friction
   * (SimParams::sdf_friction, shared with the plane as in the reference).  Positions at
   * xrt_start are the "initial positions": pairs closer than the diameter there never
   * collide (SpatialHashGPU::SetInitialPositions). */
  int32_t enable_particle_collision;
  int32_t max_num_neighbors;     /* ParticleCollisionSettings::max_neighbor_size (64)  */
  int32_t inter_leaved_hash;     /* ParticleCollisionSettings::inter_leaved_hash (3)   */
  float particle_diameter;       /* ParticleCollisionSettings::particle_diameter (1.3) */
  float hash_cell_size;          /* ParticleCollisionSettings::hash_cell_size (1.5)    */
  /* constraint model of every iteration.  XRT_CONSTRAINTS_XPBD: the colored XPBD stretch and
   * dihedral bend constraints (swift mode).  XRT_CONSTRAINTS_FEM: the quality mode's FEM
   * constraints in their place (ClothSolverBase.cpp:314-323, set up at :465-485): the
   * orthotropic Green-strain triangle constraint over faces colored at xrt_start, then, with
   * solve_bending, the isometric bending constraint over the bend hinges (XPBD, multiplier
   * reset each substep).  Rest data come from the registered positions.  Everything else
   * in the iteration (contacts, LRA, ground, bindings) runs as configured. */
  int32_t constraint_model;
  /* FEMFabricSettings (xrtailor/include/xrtailor/config/Base.hpp:47-57): Young's moduli
   * along the two material axes and the shear modulus, the Poisson ratios (xy * yx != 1),
   * and the isometric bending stiffness (> 0); defaults 1, 1, 1, 0.3, 0.3, 1 */
  float fem_xx_stiffness;
  float fem_yy_stiffness;
  float fem_xy_stiffness;
  float fem_xy_poisson_ratio;
  float fem_yx_poisson_ratio;
  float fem_bending_stiffness;
  /* friction of the body-cloth vertex-face contacts (VF-DCD and VF-CCD solves), >= 0: each
   * correction removes the tangential part of the vertex's substep displacement relative to
   * the obstacle face (its vertices' mean frame displacement / substeps), at most
   * contact_friction x |correction| (ComputeFriction, ClothSolverHelper.cu:156-169, applied
   * to the contact correction).  The reference applies no friction on mesh contacts: 0
   * (default) is its behaviour.  Edge-edge contacts and self-contact repulsion have none. */
  float contact_friction;
} xrt_params;

/* Launch shapes of the batched substep (no counterpart in the reference, whose launch
 * shapes are fixed).  None of them changes a result: they change how many threads share
 * an instance's colors or how many independent constraints of one color a thread loads
 * before solving them. */
typedef struct xrt_tuning {
  int32_t persistent_block;  /* threads per workgroup of the persistent substep (no contacts):
                              * a multiple of 64 in [64, 1024]; default 1024 */
  int32_t persistent_unroll; /* constraints per thread per round in the persistent substep:
                              * 0 default (2 stretch / 1 bend in LDS, 4 / 2 in global memory,
                              * 4 FEM), 1 light (1 / 1, 2 / 1, 2), 2 heavy (4 / 2, 8 / 4, 8) */
  int32_t wide_per_thread;   /* stretch / bend elements per thread of the wide launches (the
                              * batched substep with contacts): 1 (default) or 2 */
  int32_t path;              /* 0 persistent substep, 1 wide substep (one launch per color of
                              * every instance, each instance over as many blocks as the
                              * color needs); contexts with contacts, particles or bindings
                              * always take the wide substep */
  float best_ms;             /* set by xrt_autotune: ms per substep of the chosen shapes */
  float default_ms;          /* set by xrt_autotune: ms per substep of the defaults */
  int32_t candidates;        /* set by xrt_autotune: configurations timed */
} xrt_tuning;

enum { XRT_BEND_REFERENCE = 0, XRT_BEND_CONSISTENT = 1 };
enum { XRT_COLOR_GREEDY = 0, XRT_COLOR_PALETTE = 1 };
enum { XRT_CONSTRAINTS_XPBD = 0, XRT_CONSTRAINTS_FEM = 1 };

typedef struct xrt_ctx xrt_ctx;

/* ------------------------------------------------------------------ context */
int xrt_abi_version(void);
const char* xrt_last_error(void);
void xrt_default_params(xrt_params* p);
/* creates a context on HIP device `device` with its own stream */
int xrt_create(xrt_ctx** out, int device);
int xrt_destroy(xrt_ctx* ctx);
/* SetSimulationParams (xrtailor/src/physics/ClothSolverHelper.cu:16-23) */
int xrt_set_params(xrt_ctx* ctx, const xrt_params* p);
int xrt_get_params(xrt_ctx* ctx, xrt_params* p);
/* waits for the context's stream; XRT_ERR_CAPACITY if an instance's face pairs exceeded
 * pair_capacity_per_face since the last check (the condition is reported once) */
int xrt_sync(xrt_ctx* ctx);
/* device properties: [0]=compute units, [1]=max LDS bytes per workgroup,
 * [2]=wavefront size, [3]=L2 bytes */
int xrt_device_info(xrt_ctx* ctx, int64_t info[4]);

/* ------------------------------------------------- host-only setup helpers */
/* These run on the host and need no device; xrt_add_cloth / xrt_start use the same code. */
/* stretch edges (PrepareBasicStretchData, xrtailor/src/physics/ClothSolverHelper.cpp:90-109);
 * *ne_out receives the edge count, XRT_ERR_CAPACITY if it exceeds max_edges */
int xrt_build_edges(const int32_t* tris, int32_t nt, int32_t* edges_out, int32_t max_edges,
                    int32_t* ne_out);
/* bend elements (GenerateBendingElements, ClothSolverHelper.cpp:4-88) */
int xrt_build_bends(const int32_t* tris, int32_t nt, const int32_t* edges, int32_t ne,
                    int32_t* bends_out, int32_t max_bends, int32_t* nb_out);
/* constraint coloring (GraphColoring::Paint, xrtailor/src/physics/graph_coloring/Base.cu:163-317) */
int xrt_color_constraints(const int32_t* cons, int32_t arity, int32_t n, int32_t nv,
                          uint64_t seed, int32_t mode, int32_t* colors_out,
                          int32_t* ncolors_out);

/* -------------------------------------------------------------- scene setup */
/* ClothSolverUniversal::AddCloth (xrtailor/src/pipeline/impl/universal/ClothSolver.cpp:247-350):
 * registers a cloth instance, builds its stretch edges (PrepareBasicStretchData,
 * xrtailor/src/physics/ClothSolverHelper.cpp:90-109) and bend elements
 * (GenerateBendingElements, ClothSolverHelper.cpp:4-88) with rest lengths / angles from
 * `pos`, and sets uniform stretch and bend compliance (BasicSolver.cu:42-70,109-125). */
int xrt_add_cloth(xrt_ctx* ctx, const float* pos, int32_t nv, const int32_t* tris, int32_t nt,
                  float stretch_compliance, float bend_compliance, int32_t* instance_out);
/* pinning: PhysicsMesh::SetInvMass(0) + SetFree(false) (universal/ClothSolver.cpp:304-307) */
int xrt_pin_vertices(xrt_ctx* ctx, int32_t inst, const int32_t* idx, int32_t n);
/* per-edge stretch compliance, in the edge order xrt_get_edges reports */
int xrt_set_stretch_compliance(xrt_ctx* ctx, int32_t inst, const float* compliance);
/* geodesic long-range attachments from slot vertices: Geodesic::ComputeGeodesicDistance +
 * ClothObjectBase::GenerateGeodesicLongRangeConstraints
 * (xrtailor/src/utils/Geodesic.cpp:6-79; xrtailor/src/pipeline/base/ClothObjectBase.cpp:80-94) */
int xrt_add_lra_slots(xrt_ctx* ctx, int32_t inst, const int32_t* slots, int32_t n);
/* ClothSolverUniversal::AddObstacle (xrtailor/src/pipeline/impl/universal/ClothSolver.cpp:189-245):
 * the instance's obstacle mesh (nodes not free, inverse mass 1e-3) */
int xrt_set_obstacle(xrt_ctx* ctx, int32_t inst, const float* pos, int32_t nv, const int32_t* tris,
                     int32_t nt);
/* ClothSolverBase::Start / SetupInternalDynamics / SetupRTriangle (ClothSolverBase.cpp:453-525):
 * colors the stretch and bend constraints of every instance from `seed`, builds the contact
 * topology (R-triangles, edge-face adjacency) and uploads the scene. */
int xrt_start(xrt_ctx* ctx, uint64_t seed);
/* UpdateObstacleAnimationFrame: the obstacle's target positions for the next frame (x); the
 * substeps interpolate from the frame's start positions (PhysicsMesh::Interpolate) */
int xrt_update_obstacle(xrt_ctx* ctx, int32_t inst, const float* pos);
/* contact records of the last collision pass: VF-DCD (face, nearest point) and VF-CCD (face,
 * hit point) per cloth vertex, EE-DCD per cloth edge as 16 floats (dist, s, t, -, p2, -, q2,
 * -, fn, -); stats = obstacle pairs, total pairs, pair capacity, overflow flag.  Any pointer
 * may be NULL. */
int xrt_get_contacts(xrt_ctx* ctx, int32_t inst, int32_t* vf_face, float* vf_q, int32_t* rt_face,
                     float* rt_q, float* ee16, int32_t stats[4]);

/* ------------------------------------ reference-shaped collider calls (one instance)
 * For an engine that keeps its own SimulateSwift / SimulateQuality loop
 * (xrtailor/src/pipeline/base/ClothSolverBase.cpp:171-313).  Each launches on the context's
 * stream exactly the kernels xrt_substep_plain launches for that step, so the reference's
 * call order through these entry points reproduces xrt_substep_plain bit for bit.  They act
 * on the instance's contact state (created by xrt_start when the instance has an obstacle
 * or self contact is on); XRT_ERR_STATE otherwise, or without an obstacle for the calls that
 * need one.  The obstacle (the reference's obstacle_bvh_ / PhysicsMesh) is the instance's
 * own, set by xrt_set_obstacle. */
/* obstacle_bvh_->UpdateData(faces, 0, false, n) + Construct() from the obstacle's frame-start
 * positions (xrtailor/include/xrtailor/physics/broad_phase/lbvh/BVH.cuh:90-92) */
int xrt_obstacle_bvh_construct(xrt_ctx* ctx, int32_t inst);
/* cloth_bvh_->UpdateData(faces, 0, false, n) + Construct() from the cloth's x0 (BVH.cuh:90-92) */
int xrt_cloth_bvh_construct(xrt_ctx* ctx, int32_t inst);
/* obstacle_->Interpolate(cnt, num_passes): x1 = x0 + (x - x0) (cnt + 1) / num_passes
 * (xrtailor/include/xrtailor/physics/PhysicsMesh.cuh:128) */
int xrt_obstacle_interpolate(xrt_ctx* ctx, int32_t inst, int32_t cnt, int32_t num_passes);
/* obstacle_->UpdateMidstepNormals() (PhysicsMesh.cuh:84) */
int xrt_obstacle_update_midstep_normals(xrt_ctx* ctx, int32_t inst);
/* obstacle_bvh_->RefitMidstep(faces, tolerance) (BVH.cuh:94) */
int xrt_obstacle_bvh_refit_midstep(xrt_ctx* ctx, int32_t inst, float tolerance);
/* obstacle_->PredictedToPositions(): obstacle x0 = x at frame end (PhysicsMesh.cuh:126) */
int xrt_obstacle_predicted_to_positions(xrt_ctx* ctx, int32_t inst);
/* VFConstraint::Generate / Solve (DCD, offset 1e-3),
 * xrtailor/include/xrtailor/physics/pbd_collision/Constraint.cuh:15-30 */
int xrt_vf_dcd_generate(xrt_ctx* ctx, int32_t inst);
int xrt_vf_dcd_solve(xrt_ctx* ctx, int32_t inst);
/* RTConstraint::Generate / Solve (VF-CCD, offset 0), Constraint.cuh:54-69 */
int xrt_vf_ccd_generate(xrt_ctx* ctx, int32_t inst);
int xrt_vf_ccd_solve(xrt_ctx* ctx, int32_t inst);
/* EEConstraint::Generate / Solve, Constraint.cuh:32-48: the solve accumulates into the
 * instance's Jacobi sums (the reference's deltas / delta_counts); xrt_apply_deltas applies
 * them */
int xrt_ee_dcd_generate(xrt_ctx* ctx, int32_t inst);
int xrt_ee_dcd_solve(xrt_ctx* ctx, int32_t inst);
/* PBDRepulsionSolver::UpdatePairs(cloth, obstacle, cloth_bvh, obstacle_bvh, thickness):
 * cloth tree Update(ccd), Traverse(cloth, obstacle) then Traverse(cloth, cloth)
 * (xrtailor/include/xrtailor/physics/repulsion/PBDRepulsion.cuh:19-20).  Needs the cloth
 * tree built (xrt_cloth_bvh_construct).  A pair count above the capacity is reported by
 * the next xrt_sync (pair_capacity_per_face) */
int xrt_repulsion_update_pairs(xrt_ctx* ctx, int32_t inst, float thickness);
/* PBDRepulsionSolver::SolveNoProximity(cloth, thickness, relaxation_rate, ...): every pair's
 * VF / EE proximity tests, impulses applied x relaxation_rate (PBDRepulsion.cuh:27-28).  It
 * uses the instance's Jacobi sums too: call xrt_apply_deltas after xrt_ee_dcd_solve before
 * the next repulsion solve, as the reference's order does */
int xrt_repulsion_solve_no_proximity(xrt_ctx* ctx, int32_t inst, float thickness,
                                     float relaxation_rate);

/* ---------------------------------------------------------------- state I/O */
int xrt_num_instances(xrt_ctx* ctx, int32_t* n);
/* counts[0..7] = nv, nt, n_edges, n_bends, stretch colors, bend colors, LRA slots, LRA constraints */
int xrt_get_counts(xrt_ctx* ctx, int32_t inst, int32_t counts[8]);
/* edges (2*ne), rest lengths, stretch colors -- any pointer may be NULL */
int xrt_get_edges(xrt_ctx* ctx, int32_t inst, int32_t* edges, float* rest, int32_t* colors);
/* bend elements (4*nb), rest angles, bend colors -- any pointer may be NULL */
int xrt_get_bends(xrt_ctx* ctx, int32_t inst, int32_t* bends, float* theta, int32_t* colors);
/* LRA: slot vertices (nslots), CSR over targets (nv+1), slot ordinals and rest lengths (K) */
int xrt_get_lra(xrt_ctx* ctx, int32_t inst, int32_t* slot_vert, int32_t* ptr, int32_t* slot,
                float* rest);
/* quality-mode FEM rest data of one instance (constraint_model FEM, after xrt_start):
 * InitStrain_Kernel / InitIsometricBending_Kernel (xrtailor/src/physics/dynamics/
 * FEMHelper.cu:42-94, 272-301) computed on the host in FP64 and stored in FP32.  Face colors
 * and their count, rest areas (nt), inverse rest matrices (4 per face: glm [0][0], [0][1],
 * [1][0], [1][1]), isometric weights K (4 per bend) and coefficients 3 / (A0 + A1) (nb), in
 * face and bend order -- any pointer may be NULL */
int xrt_get_fem(xrt_ctx* ctx, int32_t inst, int32_t* face_colors, int32_t* n_face_colors,
                float* area, float* inv, float* iso_k, float* iso_coeff);
int xrt_get_inv_mass(xrt_ctx* ctx, int32_t inst, float* w);
/* sets an instance's positions (x0, also the predicted x) and / or velocities; NULL keeps
 * one.  Before xrt_start it replaces the initial state.  On a started context new positions
 * also void what earlier substeps derived from the old ones: the obstacle contact records
 * are cleared (none acts before the instance's next collision pass), the cloth tree and the
 * particle neighbor lists are rebuilt from the new positions.  The frame schedule (substep
 * of the frame, collision passes) is kept. */
int xrt_set_state(xrt_ctx* ctx, int32_t inst, const float* x0, const float* v);
int xrt_get_state(xrt_ctx* ctx, int32_t inst, float* x0, float* v);
/* predicted / iterate positions x (PhysicsMesh nodes' x) */
int xrt_get_predicted(xrt_ctx* ctx, int32_t inst, float* x);

/* ------------------------------------------ reference-shaped kernels (plain) */
/* PredictPositions (xrtailor/include/xrtailor/physics/ClothSolverHelper.cuh:20-41) */
int xrt_predict_positions(xrt_ctx* ctx, int32_t inst, float dt);
/* BasicConstraint::StretchSolver::Solve(nodes, prev, color, dt, iter)
 * (xrtailor/include/xrtailor/physics/dynamics/BasicSolver.cuh:35-54) */
int xrt_stretch_solve(xrt_ctx* ctx, int32_t inst, int32_t color, float dt, int32_t iter);
/* BasicConstraint::BendSolver::Solve(nodes, color, dt, iter) (BasicSolver.cuh:15-33) */
int xrt_bend_solve(xrt_ctx* ctx, int32_t inst, int32_t color, float dt, int32_t iter);
/* FEM::StrainSolver::Solve(nodes, color) (xrtailor/include/xrtailor/physics/dynamics/
 * FEMSolver.cuh:40, FEMSolver.cu:36-38): one face color of the Green-strain constraint
 * (constraint_model FEM) */
int xrt_strain_solve(xrt_ctx* ctx, int32_t inst, int32_t color);
/* FEM::IsometricBendingSolver::Solve(nodes, color, iter, dt) (FEMSolver.cuh:22,
 * FEMSolver.cu:83-86): one bend
 * color of the isometric bending constraint, multipliers reset when iter == 0
 * (constraint_model FEM) */
int xrt_isometric_bending_solve(xrt_ctx* ctx, int32_t inst, int32_t color, int32_t iter,
                                float dt);
/* SolveGeodesicLRA (xrtailor/include/xrtailor/physics/dynamics/LongRangeAttachments.cuh:11-13) */
int xrt_solve_lra(xrt_ctx* ctx, int32_t inst);
/* ground plane projection (SDFCollider type 1, xrtailor/include/xrtailor/physics/sdf/SDFCollider.cuh:36-43) */
int xrt_solve_ground(xrt_ctx* ctx, int32_t inst);
/* ApplyDeltas (xrtailor/include/xrtailor/physics/ClothSolverHelper.cuh:37): the instance's
 * contact Jacobi sums (xrt_ee_dcd_solve) applied x relaxation_factor and cleared; no-op
 * for an instance without contact state */
int xrt_apply_deltas(xrt_ctx* ctx, int32_t inst);
/* Finalize (ClothSolverHelper.cuh:20-41) */
int xrt_finalize(xrt_ctx* ctx, int32_t inst, float dt);
/* ------------------------------- obstacle skinning and cloth-to-body bindings */
/* BindingConstraint::Add(indices, stiffnesses, distances, offset)
 * (xrtailor/include/xrtailor/physics/dynamics/BindingConstraint.cuh:27-28), before
 * xrt_start: instance-local cloth vertices bound to the instance's obstacle (set first;
 * >= 2 faces); distance < 1e-6 takes the distance at xrt_start.  A vertex is bound at most
 * once.  At xrt_start each bound instance gets QueryNearestBarycentricClothToObstacle over
 * its cloth vertices and BindingConstraint::Generate; every iteration ends with
 * BindingConstraint::Solve against the obstacle's frame positions. */
int xrt_add_bindings(xrt_ctx* ctx, int32_t inst, const int32_t* idx, const float* stiffness,
                     const float* distance, int32_t n);
/* skinning weights of the obstacle's rest vertices (the body's sparse LBS weights,
 * xrtailor/src/runtime/rag_doll/smpl/cuda/body.cu:39-58), before xrt_start: CSR over the
 * obstacle vertices (ptr nv + 1, joint / weight ptr[nv]), 1..256 joints */
int xrt_set_obstacle_skinning(xrt_ctx* ctx, int32_t inst, const int32_t* ptr, const int32_t* joint,
                              const float* weight, int32_t n_joints);
/* linear blend skinning (body.cu:39-58) of every skinned obstacle into its frame positions
 * (as xrt_update_obstacle does from host positions): transforms of every skinned instance
 * in instance order, n_joints x 12 floats each (3x4 row-major [R | t]).  The transforms are
 * copied into pinned staging before the call returns (the caller may reuse its array at
 * once); the upload and the skinning are queued on the context's stream without waiting. */
int xrt_skin_obstacles(xrt_ctx* ctx, const float* transforms);
int xrt_skin_obstacle(xrt_ctx* ctx, int32_t inst, const float* transforms);
/* QueryNearestBarycentricClothToObstacle (xrtailor/include/xrtailor/physics/
 * ClothSolverHelper.cuh:20-41) of the instance's cloth x0 against its obstacle's x0 tree
 * (bvh tolerance): per cloth vertex idx[4] = (face vertices u, v, w, face) and
 * uvw[4] = (weights u, v, w, distance) */
int xrt_query_nearest_barycentric(xrt_ctx* ctx, int32_t inst, int32_t* idx, float* uvw);
/* the generated bindings (BindingConstraint::bindings): idx[4] = (cloth vertex, body
 * vertices u, v, w), par[5] = (u, v, w, L, stiffness) */
int xrt_get_bindings(xrt_ctx* ctx, int32_t inst, int32_t* idx, float* par);
/* BindingConstraint::Solve(cloth_nodes, obstacle_nodes) (BindingConstraint.cuh:31) */
int xrt_solve_bindings(xrt_ctx* ctx, int32_t inst);
/* obstacle positions: start of the frame (x0) and the frame's target (x); either may be null */
int xrt_get_obstacle(xrt_ctx* ctx, int32_t inst, float* x0, float* x);
/* timing helper: `n` launches of the skinning of every skinned obstacle (the last uploaded
 * transforms) and of the binding solve of every instance, bracketed by HIP events;
 * ms[0] / ms[1] = average device time per launch.  The cloth state is moved by the solves.
 * No counterpart in the reference (measurement hook). */
int xrt_time_binding(xrt_ctx* ctx, int32_t n, float ms[2]);
/* device time of the FEM kernels (constraint_model FEM): `n` passes over every face color
 * of every instance (wide strain launches), then `n` over every bend color (wide isometric
 * bending launches, pass r as iteration r), each launch bracketed by its own HIP event pair
 * on the context stream; ms[0] / ms[1] = average device time per launch.  The cloth state is moved by the
 * solves.  No counterpart in the reference (measurement hook). */
int xrt_time_fem(xrt_ctx* ctx, int32_t n, float ms[2]);

/* SpatialHashGPU::Hash(nodes, n) of one instance's predicted positions x
 * (xrtailor/include/xrtailor/physics/broad_phase/spatial_hash/SpatialHashGPU.cuh:26-28):
 * hash, sort, cell table, neighbor lists (needs particle collision on) */
int xrt_particle_hash(xrt_ctx* ctx, int32_t inst);
/* CollideParticles(deltas, counts, nodes, neighbors) + ApplyDeltas x relaxation_factor
 * (xrtailor/include/xrtailor/physics/ClothSolverHelper.cuh:20-41) */
int xrt_collide_particles(xrt_ctx* ctx, int32_t inst);
/* the hash state of one instance (SpatialHashGPU's public buffers); any output may be
 * null.  neighbors: nv * max_num_neighbors, entry s of particle i at s * nv + i, a list
 * ends at 0xffffffff unless full; sorted_hash / sorted_index: nv; cell_start / cell_end:
 * 2 nv (cell_end is meaningful where cell_start != 0xffffffff).  info: particle diameter,
 * cell spacing. */
int xrt_get_particle_hash(xrt_ctx* ctx, int32_t inst, uint32_t* neighbors, uint32_t* sorted_hash,
                          int32_t* sorted_index, uint32_t* cell_start, uint32_t* cell_end,
                          float info[2]);

/* ------------------------------------------------------------------ drivers */
/* one substep of one instance through the plain per-color launches, in the order of
 * ClothSolverBase::SimulateSwift (xrtailor/src/pipeline/base/ClothSolverBase.cpp:190-240) */
int xrt_substep_plain(xrt_ctx* ctx, int32_t inst);
/* `nframes` frames of every instance, one instance after another (one cloth per launch
 * sequence) -- the plain path of the batched workload */
int xrt_frames_plain(xrt_ctx* ctx, int32_t nframes);
/* one substep of all instances: without contacts one persistent launch (CSR instance
 * table); with body or self contact every constraint color, LRA pass, contact phase and
 * LBVH operation of all instances in one launch each.  Results equal xrt_substep_plain's
 * bit for bit. */
int xrt_substep_batched(xrt_ctx* ctx);
/* `nframes` frames of all instances; the frame's substep launches are captured once in
 * a hipGraph and replayed.  The capture holds the frame schedule it was recorded at (which
 * substep starts the frame, when collision passes and particle hashes run); after single
 * substeps have shifted the schedule the next call captures again, so batched frames equal
 * plain frames after any sequence of substeps. */
int xrt_frames_batched(xrt_ctx* ctx, int32_t nframes);
/* Autotuner (no counterpart in the reference): times every candidate launch shape of the
 * path the context's batched substep takes (without contacts, particles or bindings: the
 * persistent substep's workgroup size 1024 / 768 / 512 / 256 x unroll 0 / 1 / 2, and the
 * wide substep; wide: 1 or 2 elements per thread) over `substeps` batched
 * substeps of the context's own instances, each from the current state, and keeps the
 * fastest.  The state (positions, velocities, frame position, obstacle frame-start
 * positions, hash schedule) is restored afterwards, so results are those of an untuned
 * run; the next batched frame re-captures its graph.  out (may be NULL) receives the
 * choice and the timings. */
int xrt_autotune(xrt_ctx* ctx, int32_t substeps, xrt_tuning* out);
int xrt_get_tuning(xrt_ctx* ctx, xrt_tuning* out);
/* applies persistent_block / persistent_unroll / wide_per_thread / path (validated; every
 * choice gives the same results bit for bit) */
int xrt_set_tuning(xrt_ctx* ctx, const xrt_tuning* t);
/* `nframes` batched frames (as xrt_frames_batched) with the per-frame cache of
 * ClothSolverUniversal::SaveCache (xrtailor/src/pipeline/impl/universal/ClothSolver.cpp:
 * 107-144): after every frame the positions of every instance (instance order, float
 * [nframes][N][3], N = all vertices) and, when `normals` is non-null, their vertex normals
 * (UpdateNormals' formula, same layout) are copied out; the copy of a frame overlaps the
 * next frame's computation. */
int xrt_record_frames(xrt_ctx* ctx, int32_t nframes, float* positions, float* normals);
/* SaveObj (xrtailor/src/runtime/export/ObjSequenceExporter.cpp:22-47): one mesh as OBJ text
 * ("o mesh", v / vn lines with 6 decimals, "s 1", 1-based faces); normals may be null */
int xrt_write_obj(const char* path, const float* pos, const float* normals, int32_t nv,
                  const int32_t* tris, int32_t nt);
/* ObjSequenceExporter::ExportObjSequence (ObjSequenceExporter.cpp:151-200): frames
 * first_frame .. first_frame + nframes - 1 of one mesh as dir/%06d.obj; pos / normals hold
 * nframes x nv x 3 floats */
int xrt_write_obj_sequence(const char* dir, int32_t first_frame, int32_t nframes, const float* pos,
                           const float* normals, int32_t nv, const int32_t* tris, int32_t nt);
/* timing helper: average device time (ms) of one batched substep over `n` substeps,
 * bracketed by HIP events on the context stream */
int xrt_time_batched_substep(xrt_ctx* ctx, int32_t n, float* avg_ms);
/* wall time (ms, HIP events on the context stream) of `nframes` plain or batched frames */
int xrt_time_frames(xrt_ctx* ctx, int32_t batched, int32_t nframes, float* ms);

/* ---------------------------------------------------------------- adjoint */
/* Differentiable rollout.  The reference has no gradients; these entry points are the
 * differentiation interface of its substep (ClothSolverBase::SimulateSwift,
 * xrtailor/src/pipeline/base/ClothSolverBase.cpp:190-240): reverse-mode adjoints of
 * predict, the colored stretch / bend projections, the ground plane, finalize and the
 * body-cloth VF / EE solves, the latter replaying the forward's contact records.
 * Unsupported (XRT_ERR_UNSUPPORTED): self contact, long-range attachments, an obstacle
 * moved by xrt_update_obstacle since its last frame end.
 *
 * xrt_adjoint_begin: runs `nframes` frames of every instance from the current state (a
 * frame boundary), keeping the state every `checkpoint_every` frames; the context is left
 * at the final state.  The reverse pass recomputes each segment between checkpoints once
 * with the tape on: `checkpoint_every` frames of tape when they fit (96 GiB and 60 % of the
 * free device memory), else one frame at a time. */
int xrt_adjoint_begin(xrt_ctx* ctx, int32_t nframes, int32_t checkpoint_every);
/* dL/dx and dL/dv of the final state of one instance (3*nv each, NULL = zero); seeds not
 * given stay zero */
int xrt_adjoint_seed(xrt_ctx* ctx, int32_t inst, const float* grad_x, const float* grad_v);
/* reverse pass over the recorded frames; afterwards the gradients below are available and
 * the context is at the final state again */
int xrt_adjoint_backward(xrt_ctx* ctx);
/* gradients of one instance: per-edge stretch compliance (ne, the instance's edge order)
 * and per-element bend compliance (nb, bend order); either may be NULL */
int xrt_adjoint_get_grad_compliance(xrt_ctx* ctx, int32_t inst, float* grad_stretch,
                                    float* grad_bend);
/* gradients w.r.t. the initial positions and velocities (3*nv each; either may be NULL) */
int xrt_adjoint_get_grad_state(xrt_ctx* ctx, int32_t inst, float* grad_x0, float* grad_v0);
/* reverse of ONE substep from the current state (taken as a frame start) with the current
 * adjoint seeds as the adjoints of its end state; the seeds become the start state's
 * adjoints, compliance gradients accumulate; the state is not advanced */
int xrt_adjoint_substep(xrt_ctx* ctx);
/* timing helper (HIP events, ms per substep over `n` substeps of every instance from the
 * current state, which is restored): ms[0] forward substep (wide launches), ms[1] taped
 * forward substep, ms[2] reverse substep */
int xrt_time_adjoint(xrt_ctx* ctx, int32_t n, float ms[3]);
/* Runs n batched substeps (not captured) with HIP events around every launch of the EE
 * repulsion kernel (one edge-edge test item per thread over every instance's item list)
 * and returns its average device time per launch; counts = EE items, VF items, stored
 * pairs and instances on the item path, averaged over the n substeps' pair refreshes.  The
 * state advances by n substeps.  No counterpart in the reference (measurement hook; the EE
 * tests are EvaluateEEProximity, xrtailor/src/physics/repulsion/PBDRepulsion.cu:411-468). */
int xrt_time_repulsion(xrt_ctx* ctx, int32_t n, float* avg_ms, int64_t counts[4]);
/* timing helper (particle collision on): `n` batched substeps with HIP events around each
 * particle collision kernel and each hash step; ms[0] / ms[1] = their average device time,
 * counts[0] = neighbor-list entries after the last hash, counts[1] = particles.  The state
 * advances by n substeps.  No counterpart in the reference (measurement hook). */
int xrt_time_particles(xrt_ctx* ctx, int32_t n, float ms[2], int64_t counts[2]);

/* ------------------------------------------------------------------- LBVH */
/* Standalone linear BVH over triangle faces (Karras 2012), the reference's BVH class
 * (xrtailor/include/xrtailor/physics/broad_phase/lbvh/BVH.cuh:77-125).  Leaves are the
 * faces; node arrays follow the reference layout: internal nodes [0, F-1), leaves
 * [F-1, 2F-1) in Morton order; node = (parent, left, right, object), -1 for none
 * (0xFFFFFFFF in the reference); boxes = (lower xyz, upper xyz). */
typedef struct xrt_bvh xrt_bvh;
int xrt_bvh_create(xrt_bvh** out, int device);
int xrt_bvh_destroy(xrt_bvh* bvh);
/* BVH::UpdateData(faces, tol, ccd, n) + BVH::Construct(scr) (BVH.cu:565-573,618-722): face
 * corners x0_k + tol n_k (and x_k + tol n_k when ccd); x and n may be NULL (n = 0). */
int xrt_bvh_build(xrt_bvh* bvh, const float* x0, const float* x, const float* n, int32_t nv,
                  const int32_t* tris, int32_t nt, float tol, int32_t ccd, float scr);
/* BVH::RefitMidstep(faces, tol) (BVH.cu:795-818): corners x1 + tol n1, leaf boxes grow,
 * parents are re-merged */
int xrt_bvh_refit_midstep(xrt_bvh* bvh, const float* x1, const float* n1, float tol);
/* BVH::Update(faces, ccd) (BVH.cu:839-852): leaf boxes grow by the faces' bounds of x (and
 * x0 with ccd), ancestors grow */
int xrt_bvh_update(xrt_bvh* bvh, const float* x, const float* x0, int32_t ccd);
/* BVH::NumObjects */
int xrt_bvh_num_objects(xrt_bvh* bvh, int32_t* n);
/* sorted Morton codes (F), sorted object order (F), NodesHost (4*(2F-1)), AABBsHost
 * (6*(2F-1)), InternalEscapeIndexTableHost (F), ExternalEscapeIndexTableHost (F),
 * GetRootAABB (6); any pointer may be NULL */
int xrt_bvh_get(xrt_bvh* bvh, uint32_t* morton, int32_t* order, int32_t* nodes, float* aabbs,
                int32_t* i_escape, int32_t* e_escape, float* root);
/* QueryDeviceStackless (BVH.cu:957-1021): strict point containment (p1 == NULL) or segment
 * p0->p1 overlap; hits is n*256 (LBVH_MAX_BUFFER_SIZE), counts is n */
int xrt_bvh_query(xrt_bvh* bvh, const float* p0, const float* p1, int32_t n, int32_t* hits,
                  int32_t* counts);
/* Traverse(bvh_lhs, bvh_rhs, faces_lhs, faces_rhs, thickness) (BVH.cuh:139-152): pairs of
 * (lhs face, rhs face) whose leaf boxes overlap after dilating the lhs box by thickness;
 * *n_out receives the count; XRT_ERR_CAPACITY if it exceeds cap (first cap written) */
int xrt_bvh_traverse(xrt_bvh* lhs, xrt_bvh* rhs, float thickness, int32_t* pairs, int64_t cap,
                     int64_t* n_out);
/* No counterpart in the reference: xrt_bvh_traverse for callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc).  Same arguments, results
 * and errors; the device staging of the returned pairs is a fixed array of 2,097,152 pairs
 * (16 MiB per device, held until the process ends) instead of a per-call allocation, and no
 * hipMalloc / hipFree is called.  XRT_ERR_CAPACITY (with *n_out set) when min(count, cap)
 * exceeds 2,097,152.  Every call on one device shares the array: calls that overlap on a
 * device (other threads, other handles) race. */
int xrt_bvh_traverse_noalloc(xrt_bvh* lhs, xrt_bvh* rhs, float thickness, int32_t* pairs,
                             int64_t cap, int64_t* n_out);
/* PhysicsMesh::UpdateNormals vertex normals (PhysicsMeshHelper.cu:196-210,282-295), summed
 * per vertex in ascending face order (deterministic) */
int xrt_vertex_normals(xrt_bvh* bvh, const float* x, int32_t nv, const int32_t* tris, int32_t nt,
                       float* n_out);

#ifdef __cplusplus
}
#endif
#endif
