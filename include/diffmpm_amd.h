/*
 * diffmpm_amd — C ABI of the MI355X (gfx950) explicit MPM time step.
 *
 * The reference (geoelements/diffmpm) has no FFI: its step is a set of Python
 * methods on JAX pytrees.  Each entry point below replaces one of those methods
 * (cited as reference file:line) or a fused phase of the step.  Everything is
 * plain C: raw device pointers, sizes and a hipStream_t passed as void*.
 *
 * Memory layout (owned by the caller, e.g. torch tensors):
 *   particle block  T pbuf[DMPM_P_ROWS][ld]   field row r of particle p is
 *                                             pbuf[r*ld + p]  (SoA)
 *   element ids     int32 eid[n]
 *   node block      T nbuf[DMPM_N_ROWS][nld]  field row r of node i is
 *                                             nbuf[r*nld + i]
 * T is float (DMPM_F32) or double (DMPM_F64), fixed per plan.  Vector fields
 * occupy 2 consecutive rows (x, y); 1D meshes use only the first.  Tensor
 * fields occupy 6 rows in Voigt order [xx, yy, zz, xy, yz, xz] (xy is the
 * engineering shear).
 *
 * Node numbering follows the reference (element.py:667-680): node
 * n = j*(nx+1) + i sits at (fl(i*hx), fl(j*hy)); cell e = J*nx + I has nodes
 * [ll, ll+1, ll+nx+2, ll+nx+1] with ll = J*(nx+1) + I (element.py:692-723).
 *
 * Return codes: 0 on success, DMPM_EINVAL for an invalid argument,
 * DMPM_EHIP for a HIP runtime error.  dmpm_last_error() gives the message
 * of the last failure on the calling host thread.
 *
 * Threading: a plan is not re-entrant; different plans are independent.
 *
 * Allocation and host synchronisation: dmpm_plan_create, dmpm_plan_set_* and
 * dmpm_plan_destroy allocate, upload and free (and wait for the device).
 * The step entry points (dmpm_run, dmpm_step_between, the phase and
 * per-method entry points, the adjoint entry points) enqueue work on the
 * given stream without waiting for it; each workspace they use is allocated
 * (hipMalloc) by the first call that needs it, or reallocated when a call
 * needs it larger (hipFree of the smaller one, which waits for the device),
 * and kept until dmpm_plan_destroy.  dmpm_prepare_run allocates the
 * workspaces of dmpm_run beforehand, so that later dmpm_run calls allocate
 * nothing and can be captured into a HIP graph.
 */
#ifndef DIFFMPM_AMD_H
#define DIFFMPM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DMPM_OK 0
#define DMPM_EINVAL (-1)
#define DMPM_EHIP (-2)

enum { DMPM_F32 = 0, DMPM_F64 = 1 };
enum { DMPM_LINEAR1D = 1, DMPM_QUAD4 = 2 };
enum { DMPM_LINEAR_ELASTIC = 0, DMPM_SIMPLE_MATERIAL = 1 };
enum { DMPM_USF = 0, DMPM_USL = 1 };
enum { DMPM_FN_UNIT = 0, DMPM_FN_LINEAR = 1 };

/* Particle block rows (particle.py:53-71). */
enum {
  DMPM_P_X = 0,           /* loc (2) */
  DMPM_P_V = 2,           /* velocity (2) */
  DMPM_P_MASS = 4,
  DMPM_P_VOLUME = 5,
  DMPM_P_DENSITY = 6,
  DMPM_P_SIZE = 7,        /* size (2) */
  DMPM_P_MOMENTUM = 9,    /* (2) */
  DMPM_P_ACCEL = 11,      /* (2) never written by the step (as in the reference) */
  DMPM_P_STRAIN = 13,     /* (6) */
  DMPM_P_STRESS = 19,     /* (6) */
  DMPM_P_STRAIN_RATE = 25,/* (6) */
  DMPM_P_DSTRAIN = 31,    /* (6) */
  DMPM_P_FEXT = 37,       /* (2) */
  DMPM_P_TRACTION = 39,   /* (2) */
  DMPM_P_XI = 41,         /* reference_loc (2) */
  DMPM_P_DVOLSTRAIN = 43, /* dvolumetric_strain */
  DMPM_P_VOLSTRAIN_C = 44,/* volumetric_strain_centroid */
  DMPM_P_YOUNGS = 45,     /* per-particle Young's modulus (extension) */
  DMPM_P_ROWS = 46
};

/* Node block rows (node.py:64-71). */
enum {
  DMPM_N_LOC = 0,   /* (2) */
  DMPM_N_V = 2,     /* (2) */
  DMPM_N_A = 4,     /* (2) */
  DMPM_N_MASS = 6,
  DMPM_N_P = 7,     /* momentum (2) */
  DMPM_N_FINT = 9,  /* (2) */
  DMPM_N_FEXT = 11, /* (2) */
  DMPM_N_FDAMP = 13,/* (2) */
  DMPM_N_ROWS = 15
};

typedef struct dmpm_plan dmpm_plan;

typedef struct dmpm_plan_desc {
  int32_t element;        /* DMPM_LINEAR1D or DMPM_QUAD4 */
  int32_t dtype;          /* DMPM_F32 or DMPM_F64 */
  int32_t nx, ny;         /* cells per axis (ny = 1 for 1D) */
  double hx, hy;          /* element lengths; node coords are fl(i*h) in T */
  int64_t nparticles;     /* particles in the set */
  int64_t ld;             /* row stride of the particle block, >= nparticles */
  int64_t nld;            /* row stride of the node block, >= node count */
  int32_t material;       /* DMPM_LINEAR_ELASTIC or DMPM_SIMPLE_MATERIAL */
  int32_t per_particle_E; /* 1: E_p from row DMPM_P_YOUNGS (LinearElastic only) */
  /* LinearElastic: De coefficients (a1, a2, G) computed by the caller in
   * double as material.py:60-111 does; with per_particle_E they are the
   * coefficients per unit E.  SimpleMaterial: de[0] = E (material.py:130);
   * a SimpleMaterial with per-particle E is DMPM_LINEAR_ELASTIC with
   * per_particle_E and de = (1, 0, 1) (dsigma = E_p dstrain; the plan
   * rejects DMPM_SIMPLE_MATERIAL with per_particle_E). */
  double de[3];
} dmpm_plan_desc;

/* ---- plan lifetime -------------------------------------------------------- */
/* Allocates the sort workspace for one particle set on one mesh.  No
 * reference counterpart (the reference has no persistent workspace). */
int dmpm_plan_create(const dmpm_plan_desc* desc, dmpm_plan** out);
/* A batch plan: nsims independent 2D Quadrilateral4Node simulations (own
 * mesh, particles, material coefficients de, constraints) stepped together,
 * one launch of each kernel per step for all of them.  Their particles are
 * laid end to end in one particle block (simulation s owns the columns
 * [sum of the earlier nparticles, + its nparticles)), their nodes end to end
 * in one node block, likewise; ld and nld are those blocks' row strides.
 * Node ids in the constraint and nodal-force tables are global (offset by the
 * earlier simulations' node counts); element ids stay per simulation.  All
 * simulations share dtype, material kind, per_particle_E, scheme, time step
 * and gravity.  Batch plans run dmpm_run's fused 2D path and
 * dmpm_step_adjoint; the per-method entry points reject them.  No reference
 * counterpart (the reference runs one simulation per MPMExplicit). */
int dmpm_batch_plan_create(const dmpm_plan_desc* descs, int32_t nsims, int64_t ld, int64_t nld,
                           dmpm_plan** out);
int dmpm_plan_destroy(dmpm_plan* plan);
const char* dmpm_last_error(void);
/* ABI version (major*100 + minor) and the row counts above, for bindings. */
int dmpm_abi_version(void);

/* Time functions (functions.py:16-41): kinds[k] is DMPM_FN_UNIT or
 * DMPM_FN_LINEAR; a Linear function uses xs/fs[offsets[k] .. +counts[k]). */
int dmpm_plan_set_functions(dmpm_plan* plan, int32_t nfn, const int32_t* kinds,
                            const int32_t* offsets, const int32_t* counts,
                            const double* xs, const double* fs);
/* Velocity constraints, applied in list order (element.py:331-334,
 * constraint.py:29-44): entry k sets dof dirs[k] of node node_ids[k]. */
int dmpm_plan_set_constraints(dmpm_plan* plan, int32_t n, const int32_t* node_ids,
                              const int32_t* dirs, const double* velocities);
/* Concentrated nodal forces, in list order (element.py:270-284): entry k adds
 * f_{fn[k]}(t) * forces[k] to f_ext[node_ids[k], dirs[k]]. */
int dmpm_plan_set_nodal_forces(dmpm_plan* plan, int32_t n, const int32_t* node_ids,
                               const int32_t* dirs, const int32_t* fn,
                               const double* forces);
/* Particle tractions, in list order (mesh.py:60-77, particle.py:329-343):
 * entry k adds f_{fn[k]}(t) * tractions[k] * V / size[dir] to traction[pids[k], dirs[k]]. */
int dmpm_plan_set_tractions(dmpm_plan* plan, int32_t n, const int32_t* pids,
                            const int32_t* dirs, const int32_t* fn,
                            const double* tractions);
/* Cells per workgroup of the fused 2D step's P2G cell pass (64, 128 or 256;
 * 0 = default 128).  A tuning knob with no effect on results.  No reference
 * counterpart. */
int dmpm_plan_set_block(dmpm_plan* plan, int32_t cells_per_block);
/* Execution options (no effect on results; no reference counterpart):
 * DMPM_OPT_USF_SPLIT = 1: the fused USF step updates stresses in a streamed
 * particle kernel before the force P2G cell pass (1, default) or inside that
 * pass (0).
 * DMPM_OPT_STEP_KERNEL = 2: the fused 2D step after the re-sort runs as one
 * kernel launch (1, default; workgroups own tiles of cells and recompute a
 * two-cell halo), as three launches (2: kinematic P2G and the stress and
 * force phases over tiles that own nodes and recompute one column and one row
 * of cells, then G2P over tiles of cells), or as one launch per phase (0).
 * Any other value is DMPM_EINVAL. */
enum { DMPM_OPT_USF_SPLIT = 1, DMPM_OPT_STEP_KERNEL = 2 };
int dmpm_plan_set_option(dmpm_plan* plan, int32_t option, int32_t value);
/* Declares that the element ids changed outside the library (forces a re-sort). */
int dmpm_mark_unsorted(dmpm_plan* plan, void* stream);

/* ---- per-method entry points (drop-in for the reference methods) ---------- */
/* element.py:517-541 (1D), 820-840 (Q4): element ids from loc. */
int dmpm_set_particle_element_ids(dmpm_plan* plan, void* pbuf, int32_t* eid, void* stream);
/* particle.py:180-206: reference_loc from loc and element ids. */
int dmpm_update_natural_coords(dmpm_plan* plan, void* pbuf, const int32_t* eid, void* stream);
/* element.py:543-546 (1D, intended h), 903-909 (Q4): element volume (one
 * value, in the working precision).  Host-only; needs no device. */
int dmpm_element_volume(int32_t element, int32_t dtype, double hx, double hy,
                        double* out_volume);
/* particle.py:158-178: V_p = V_e / count(e), size = V^(1/ndim), m = V * rho. */
int dmpm_compute_particle_volume(dmpm_plan* plan, void* pbuf, const int32_t* eid,
                                 double element_volume, void* stream);
/* element.py:122-152. */
int dmpm_compute_nodal_mass(dmpm_plan* plan, const void* pbuf, const int32_t* eid,
                            void* nbuf, void* stream);
/* element.py:154-189. */
int dmpm_compute_nodal_momentum(dmpm_plan* plan, const void* pbuf, const int32_t* eid,
                                void* nbuf, void* stream);
/* element.py:191-202. */
int dmpm_compute_velocity(dmpm_plan* plan, void* nbuf, void* stream);
/* element.py:331-334. */
int dmpm_apply_boundary_constraints(dmpm_plan* plan, void* nbuf, void* stream);
/* element.py:204-234. */
int dmpm_compute_external_force(dmpm_plan* plan, const void* pbuf, const int32_t* eid,
                                void* nbuf, void* stream);
/* element.py:236-268. */
int dmpm_compute_body_force(dmpm_plan* plan, const void* pbuf, const int32_t* eid,
                            void* nbuf, const double* gravity, void* stream);
/* mesh.py:60-77 first half: zero_traction + assign_traction at time t. */
int dmpm_assign_tractions(dmpm_plan* plan, void* pbuf, double t, void* stream);
/* element.py:286-303. */
int dmpm_apply_particle_traction_forces(dmpm_plan* plan, const void* pbuf,
                                        const int32_t* eid, void* nbuf, void* stream);
/* element.py:270-284. */
int dmpm_apply_concentrated_nodal_forces(dmpm_plan* plan, void* nbuf, double t, void* stream);
/* element.py:548-600 (1D), 842-901 (Q4). */
int dmpm_compute_internal_force(dmpm_plan* plan, const void* pbuf, const int32_t* eid,
                                void* nbuf, void* stream);
/* element.py:305-329. */
int dmpm_update_nodal_acceleration_velocity(dmpm_plan* plan, void* nbuf, double dt,
                                            void* stream);
/* particle.py:208-247. */
int dmpm_update_position_velocity(dmpm_plan* plan, void* pbuf, const int32_t* eid,
                                  const void* nbuf, double dt, int32_t velocity_update,
                                  void* stream);
/* particle.py:249-313. */
int dmpm_compute_strain(dmpm_plan* plan, void* pbuf, const int32_t* eid, const void* nbuf,
                        double dt, void* stream);
/* particle.py:324-327. */
int dmpm_update_volume(dmpm_plan* plan, void* pbuf, void* stream);
/* particle.py:315-322 -> material.py:113-116 / 130-131. */
int dmpm_compute_stress(dmpm_plan* plan, void* pbuf, void* stream);

/* ---- fused phases (scheme.py:22-67) and the whole step (solver.py:179-185) -- */
/* scheme.py:22-29: element ids, natural coords, nodal mass/momentum/velocity,
 * constraints. */
int dmpm_phase_nodal_kinematics(dmpm_plan* plan, void* pbuf, int32_t* eid, void* nbuf,
                                void* stream);
/* scheme.py:31-35: strain, volume, stress. */
int dmpm_phase_stress_strain(dmpm_plan* plan, void* pbuf, const int32_t* eid,
                             const void* nbuf, double dt, void* stream);
/* scheme.py:37-56: external, body, traction, nodal and internal forces at
 * t = step*dt (computed in T as the reference's traced step loop does). */
int dmpm_phase_forces(dmpm_plan* plan, void* pbuf, const int32_t* eid, void* nbuf,
                      const double* gravity, int64_t step, double dt, void* stream);
/* scheme.py:58-67: nodal acceleration/velocity update and G2P. */
int dmpm_phase_particle_kinematics(dmpm_plan* plan, void* pbuf, const int32_t* eid,
                                   void* nbuf, double dt, int32_t velocity_update,
                                   void* stream);
/* nsteps explicit steps starting at step index step0 (solver.py:179-213),
 * USF or USL.  flags bit 0: also materialise the per-step temporaries
 * (strain_rate, dstrain, dvolumetric_strain, reference_loc, momentum);
 * bit 1: particle f_ext is non-zero; bit 2: use the per-phase kernels even
 * where the fused node-tile kernels apply (2D, no particle tractions, zero
 * particle f_ext).  Both paths give bit-identical results.  Bit 3
 * (DMPM_RUN_CLOCK, fused 2D path only): see below. */
int dmpm_run(dmpm_plan* plan, void* pbuf, int32_t* eid, void* nbuf, int32_t scheme,
             int32_t velocity_update, const double* gravity, double dt, int64_t step0,
             int64_t nsteps, int32_t flags, void* stream);

/* ---- graph capture (no reference counterpart; the reference traces its
 * step loop with jax.lax, solver.py:179-213) ----------------------------------
 * dmpm_run on the fused 2D path enqueues only kernels and device-to-device
 * copies on its stream, with no host synchronisation, so a run can be
 * captured into a HIP graph (hipStreamBeginCapture ... hipGraphLaunch) once
 * dmpm_prepare_run has allocated its workspaces.  With DMPM_RUN_CLOCK the
 * steps' indices (the time t = step * dt of the time functions) are read
 * from the plan's device step counter when the kernels run, step0 is
 * ignored, and the run adds nsteps to the counter at its end: each replay
 * of a captured run continues where the previous one ended.  A graph stays
 * valid while the plan and its buffers live and its tables (constraints,
 * functions, nodal forces) are not set again; replays and other work on the
 * plan must be ordered on one stream. */
enum { DMPM_RUN_CLOCK = 8 };
/* Sets the plan's step counter (enqueued on stream). */
int dmpm_plan_set_clock(dmpm_plan* plan, int64_t step, void* stream);
/* Reads the plan's step counter (waits for stream). */
int dmpm_plan_get_clock(dmpm_plan* plan, int64_t* step, void* stream);
/* Allocates every workspace that dmpm_run with this scheme and flags uses and
 * uploads its tables, and waits for stream; enqueues no work on the particle
 * or node state. */
int dmpm_prepare_run(dmpm_plan* plan, int32_t scheme, int32_t flags, void* stream);
/* One fused 2D step (index step) from particle block pbuf_in to particle
 * block pbuf_out: pbuf_in is left unchanged, pbuf_out receives the state
 * after the step; pbuf_out's constant rows (mass, Young's modulus) must hold
 * the particle set's values.  eid and the plan's sort state follow the
 * particles as in dmpm_run.  flags bit 0: materialise the temporaries (into
 * pbuf_out); bit 1: the input is the output of this plan's previous call,
 * unchanged since (skips the check for particles moved by the caller).
 * State tensors (8 x state_ld, rows x, y, vx, vy, sigma_xx, sigma_yy,
 * sigma_xy, V; row r at base + r*state_ld): with state_out only, it also
 * receives those rows after the step; with state_in and state_out, those
 * rows are read from state_in instead of pbuf_in and written to state_out
 * instead of pbuf_out (the blocks carry the other rows).  Results equal
 * dmpm_run's on a copy of the input.  A caller that alternates two blocks
 * steps without copying the state (the reference's solver.py:179-213 step,
 * split for differentiable rollouts); no reference counterpart. */
int dmpm_step_between(dmpm_plan* plan, const void* pbuf_in, void* pbuf_out, int32_t* eid,
                      void* nbuf, int32_t scheme, int32_t velocity_update,
                      const double* gravity, double dt, int64_t step, int32_t flags,
                      const void* state_in, void* state_out, int64_t state_ld, void* stream);

/* One step of a differentiable rollout that keeps what its adjoint needs:
 * dmpm_step_between from pbuf_in to pbuf_out (same flags, except that bit 0,
 * temporaries, is rejected; same state tensors) that writes every node row of
 * nout (mass, forces, velocity, acceleration, momentum, and the damping force,
 * which the step reads from nprev; its coordinate rows are not written),
 * keeps the velocity of nprev's velocity rows at massless nodes (nprev: the
 * previous step's node block, read only), and writes the step's kinematic nodal
 * velocity to vkin_out [2][nnodes] (particle dtype) and its sort by cell to
 * perm_out [nparticles] and cs_out [ncells+2], as dmpm_plan_export_step
 * would.  Results equal dmpm_step_between on nout holding nprev's rows
 * followed by dmpm_plan_export_step.  No reference counterpart (the reference
 * differentiates solver.py:179-213 with jax.grad); for the saved forward
 * steps that dmpm_step_adjoint_saved_state reverses. */
int dmpm_step_between_saved(dmpm_plan* plan, const void* pbuf_in, void* pbuf_out, int32_t* eid,
                            const void* nprev, void* nout, int32_t scheme,
                            int32_t velocity_update, const double* gravity, double dt,
                            int64_t step, int32_t flags, const void* state_in, void* state_out,
                            int64_t state_ld, void* vkin_out, int32_t* perm_out,
                            int32_t* cs_out, void* stream);

/* dmpm_step_between for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc); no
 * counterpart in the reference.  Same arguments, restrictions and results
 * as dmpm_step_between for single-simulation plans whose one-launch step
 * applies; never calls hipMalloc or hipFree.  Its workspaces (kinematic
 * nodal velocity, previous and next nodal velocities, overflow lists, step
 * tables: 24 bytes per node in FP32, 48 in FP64, plus 8 bytes per 16x16
 * tile and 1 KiB) live in one static device array of
 * DMPM_NOALLOC_SCRATCH_BYTES per GPU, which every call on that GPU shares:
 * calls that overlap on one GPU race, and the step's kinematic nodal
 * velocity is not kept for dmpm_plan_export_step.  A plan that needs more
 * is rejected with DMPM_EINVAL (at 32 MiB: up to about 1.3 M nodes in FP32,
 * 0.65 M in FP64). */
#define DMPM_NOALLOC_SCRATCH_BYTES (32u << 20)
int dmpm_step_between_noalloc(dmpm_plan* plan, const void* pbuf_in, void* pbuf_out,
                              int32_t* eid, void* nbuf, int32_t scheme, int32_t velocity_update,
                              const double* gravity, double dt, int64_t step, int32_t flags,
                              const void* state_in, void* state_out, int64_t state_ld,
                              void* stream);

/* ---- adjoint of the fused 2D step ------------------------------------------
 * The reference differentiates a rollout with jax.grad (no hand-written
 * adjoint; diffmpm/solver.py:179-213).  This is the reverse-mode
 * step of dmpm_run's fused 2D path for the dynamic particle state, with the
 * cotangent rows below (block [DMPM_C_ROWS][ld], same element type as the
 * particle block).  Mass, mesh, constraints and nodal forces are held fixed;
 * strain, density, volumetric strain and sigma_zz do not feed back into the
 * 2D dynamics and carry no cotangent. */
enum {
  DMPM_C_X = 0,    /* 2 rows: x, y */
  DMPM_C_V = 2,    /* 2 rows: vx, vy */
  DMPM_C_SXX = 4,
  DMPM_C_SYY = 5,
  DMPM_C_SXY = 6,
  DMPM_C_VOL = 7,
  DMPM_C_ROWS = 8
};
/* pin: particle block at the start of step `step` (read only; the rows
 * loc, velocity, mass, volume, density, stress and (per-particle E)
 * youngs_modulus are used); pwork, nwork, eid_work: scratch particle block,
 * node block and element ids (the forward step is recomputed there; pwork's
 * other rows must hold finite values); cot: in = cotangent of the state after the
 * step, out = cotangent of the state before it; lebar (may be NULL): [ld],
 * += dL/dlog(E_p) (per-particle E) or its per-particle terms (scalar E: sum
 * them).  Covers the fused path's plans (2D Q4, no particle tractions or
 * particle f_ext) and single Linear1D plans (the per-phase step of
 * scheme.py:22-67, no particle tractions or particle f_ext; in 1D the y and
 * vy rows pass their cotangents through).  Uses the plan's sort state: use a
 * separate plan from the forward rollout's to keep that one's incremental
 * re-sort. */
int dmpm_step_adjoint(dmpm_plan* plan, const void* pin, void* pwork, int32_t* eid_work,
                      void* nwork, void* cot, void* lebar, int32_t scheme,
                      int32_t velocity_update, const double* gravity, double dt, int64_t step,
                      void* stream);
/* After a one-step dmpm_run of the fused path: copies the step's kinematic
 * nodal velocity (vkin_out [2][nnodes], particle dtype) and its sort by cell
 * (perm_out [nparticles], cs_out [ncells+2]) so that the adjoint of that step
 * can run later without recomputing it. */
int dmpm_plan_export_step(dmpm_plan* plan, void* vkin_out, int32_t* perm_out, int32_t* cs_out,
                          void* stream);
/* dmpm_step_adjoint from saved forward quantities instead of a recomputation:
 * pin the step's input block; pout a particle block whose stress and volume
 * rows hold the step's output values; nsaved the node block after the step;
 * vkin_saved, perm_saved, cs_saved from dmpm_plan_export_step. */
int dmpm_step_adjoint_saved(dmpm_plan* plan, const void* pin, const void* pout,
                            const void* nsaved, const void* vkin_saved,
                            const int32_t* perm_saved, const int32_t* cs_saved, void* cot,
                            void* lebar, int32_t scheme, int32_t velocity_update,
                            const double* gravity, double dt, void* stream);
/* dmpm_step_adjoint_saved with the step's input and output state given as
 * (8 x state_ld) state tensors (rows x, y, vx, vy, sigma_xx, sigma_yy,
 * sigma_xy, V; row r at state + r*state_ld) instead of particle blocks;
 * pconst: a particle block whose mass and Young's modulus rows hold the
 * particle set's values.  Same results.  No reference counterpart (the
 * reference differentiates with jax.grad). */
int dmpm_step_adjoint_saved_state(dmpm_plan* plan, const void* pconst, const void* state_in,
                                  const void* state_out, int64_t state_ld, const void* nsaved,
                                  const void* vkin_saved, const int32_t* perm_saved,
                                  const int32_t* cs_saved, void* cot, void* lebar,
                                  int32_t scheme, int32_t velocity_update,
                                  const double* gravity, double dt, void* stream);

/* ---- output (replaces the host callback of solver.py:113-114, 187-210) ------
 * Copies the output rows of the first n particles of a particle block
 * (row stride ld, element type dtype) into host_dst as 16 rows of n
 * elements: block rows 0..3 (loc, velocity; 1D uses rows 0 and 2) then rows
 * 13..24 (strain, stress).  Two hipMemcpy2DAsync on stream: no host
 * synchronisation and no allocation, so the copy can be captured into a HIP
 * graph.  host_dst should be pinned host memory; it is valid once the
 * stream has passed the copy. */
int dmpm_snapshot_rows(const void* pbuf, int64_t ld, int32_t dtype, int64_t n, void* host_dst,
                       void* stream);

/* ---- measurement (no reference counterpart) ------------------------------- */
/* Kernel classes for dmpm_timing_enable. */
enum {
  DMPM_KC_LOCATE = 0, DMPM_KC_SORT = 1, DMPM_KC_P2G_KIN = 2, DMPM_KC_STRESS = 3,
  DMPM_KC_P2G_FORCE = 4, DMPM_KC_GRID = 5, DMPM_KC_G2P = 6, DMPM_KC_SMALL = 7,
  DMPM_KC_TILE_KIN = 8, DMPM_KC_TILE_FORCE = 9, DMPM_KC_TILE_G2P = 10,
  DMPM_KC_ADJ = 11, DMPM_KC_BATCH = 12, DMPM_KC_STEP = 13, DMPM_KC_ALL = -2
};
/* Times every launch of one kernel class (DMPM_KC_ALL: of every class) with
 * HIP event pairs recorded on the launch stream; kernel_class = -1 disables.
 * A ring of 4 pairs per class is reused, so the host waits for the timed
 * launch of that class 4 launches back before recording a new pair.
 * max_launches is ignored (kept for ABI stability). */
int dmpm_timing_enable(dmpm_plan* plan, int32_t kernel_class, int32_t max_launches);
/* Waits for the recorded events; returns the summed duration and count of the
 * enabled class (of all classes after DMPM_KC_ALL). */
int dmpm_timing_read(dmpm_plan* plan, double* total_ms, int64_t* launches);
/* The same for one class. */
int dmpm_timing_read_class(dmpm_plan* plan, int32_t kernel_class, double* total_ms,
                           int64_t* launches);

#ifdef __cplusplus
}
#endif

#endif /* DIFFMPM_AMD_H */
