/*
 * flatworld_fem.h -- C ABI of the MI355X explicit-FEM substep.
 *
 * One library, libflatworld_fem.so, built from flatworld_amd/csrc/ for gfx950.
 * Plain C types only: pointers, sizes, int status codes.  Every function
 * returns FW_OK (0) or a negative FW_E* code; fw_last_error() gives the text
 * of the most recent failure on the calling thread.  No C++ exception crosses
 * this boundary.  Host pointers passed in are borrowed for the duration of the
 * call only; all device memory belongs to the arena.
 *
 * Model: an *arena* is a packed region of HBM on one GPU holding any number of
 * *scenes*.  A scene is one FEM body (linear triangles, one material) or one
 * spring-mass body (two-node springs) with up to FW_MAX_GROUNDS ground contacts.  fw_step advances every resident scene by
 * n substeps, one fused kernel launch per substep; each scene carries its own
 * dt, damping, step counter and horizon.
 *
 * Reference interfaces this replaces (FlatWorld, paths relative to its root):
 *   fw_scene_add      FemSpringManager.__init__/addFemDomain/processConditions/
 *                     prepareFEMData_ + MixedContact._build_batched_contacts
 *                     (flatworld/femspringmanager.py:509-638, 660-813, 881-897;
 *                      flatworld/mixedcontact.py:1756-2269)
 *   fw_step           ExplicitLoop.advance: BVH gate, run_batched_contacts,
 *                     FemSpringManager.substep -> femStep
 *                     (flatworld/explicitloop.py:473-555;
 *                      flatworld/femspringmanager.py:938-1033)
 *   fw_scene_read     the .numpy() reads of coords/V/U/Fext/elemStress/...
 *                     (flatworld/femspringmanager.py:566-614; femdomain.py:47-50)
 *   fw_scene_remove   no counterpart: the reference never removes a domain.
 *   fw_scene_mass_scaling  FemSpringManager.applyMassScaling
 *                     (flatworld/femspringmanager.py:141-201, 899-924)
 *   fw_flex_add       the FlexFlex part of MixedContact._build_batched_contacts
 *                     and the FEM spatial hash's allocation
 *                     (flatworld/mixedcontact.py:1857-1957;
 *                      flatworld/femspringmanager.py:556; spatialmanager.py:508-565)
 *   fw_flex_rebuild   FemSpringManager.populate_fem_spatial_hash
 *                     (flatworld/femspringmanager.py:421-487, 1041-1065;
 *                      flatworld/spatialmanager.py:567-712)
 *   fw_step (groups)  _batched_flexflex_contact_kernel_sh, launched by
 *                     run_batched_contacts (flatworld/mixedcontact.py:886-1025, 1663-1695)
 * Spring-mass scenes replace addSpringDomain and _spring_force_kernel
 * (flatworld/femspringmanager.py:814-870, 283-312; flatworld/springmass.py).
 */
#ifndef FLATWORLD_FEM_H
#define FLATWORLD_FEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FW_ABI_VERSION 8
#define FW_MAX_GROUNDS 8

/* status codes */
#define FW_OK 0
#define FW_E_INVALID (-1)   /* bad argument (maps to ValueError) */
#define FW_E_CAPACITY (-2)  /* arena full / too many scenes (maps to RuntimeError) */
#define FW_E_HIP (-3)       /* HIP runtime failure */
#define FW_E_NOT_FOUND (-4) /* unknown scene id */
#define FW_E_UNSUPPORTED (-5)

/* material codes (flatworld/definitions.py:80-83) */
#define FW_MAT_ELASTIC 1
#define FW_MAT_NEOHOOKEAN 2
#define FW_MAT_MISES 3
#define FW_MAT_SPRING 4 /* spring-mass domain: two-node springs (flatworld/springmass.py) */

/* ground kinds */
#define FW_GROUND_PLANE 1
#define FW_GROUND_HEIGHTFIELD 2
#define FW_GROUND_VOXEL 3
/* FW_GROUND_FLEX: the scene's FEM-FEM contact contributions, filled each substep by
 * its contact group (fw_flex_add).  It goes where the reference launches the
 * FEM-FEM kernel: after the plane records, before the height-field and voxel
 * records.  Of its descriptor only kind and flex_cap are read: flex_cap bounds the
 * contributions per substep (own work items + 2 x work items that target this body). */
#define FW_GROUND_FLEX 4
/* FW_GROUND_RIGID: a rigid primitive (ball, box, capsule) at a pose the caller sets
 * (fw_scene_set_rigid), pushing the scene's boundary nodes (a spring scene: every
 * node) with the reference's FEM-rigid and spring-rigid penalty contact
 * (_batched_rigid_prim_kernel, flatworld/mixedcontact.py:1025-1117; work lists at
 * mixedcontact.py:1987-2030).  It goes where the reference launches that
 * kernel: after the FEM-FEM record, before the height-field and voxel records.
 * aabb_lo / aabb_hi hold the primitive's AABB at the pose (getBallBBox /
 * getBoxBBox / getCapsuleBBox, flatworld/utils.py:235-302); the gate pads it on BVH
 * rebuild substeps like any ground AABB.  The rigid is kinematic: its reaction
 * impulses are not accumulated. */
#define FW_GROUND_RIGID 5

/* rigid types (flatworld/definitions.py:50-55) */
#define FW_RIGID_BALL 1
#define FW_RIGID_BOX 2
#define FW_RIGID_CAPSULE 8
/* A rigid triangle mesh touched through its boundary edges: the FEM side of
 * _batched_rigid_mesh_kernel_sh, its workset filter and the mesh tie kernels
 * (flatworld/mixedcontact.py:519-588, 1119-1565).  The record carries the loop's
 * mesh rigids (boundary node positions at the pose, edges, edge AABBs and margin
 * bases) and the RigidManager's spatial hash over their edges, all built on the
 * host; per work item (the scene's boundary nodes) the device keeps the cached
 * edge, the skip epoch, the friction history and the tie.  Its aabb is the mesh's
 * AABB at the pose.  Non-tied items are gated by the pair gate inside the rigid
 * pass; tied items are resolved once at local step 0 and pulled to their edge
 * point every substep whatever the gate. */
#define FW_RIGID_MESH 16

/* node flag word: bits 0..8 BC bits (flatworld/definitions.py:95-104),
 * bit 15 boundary node, bits 16..31 row of the scene's BC value table */
#define FW_NODE_BC_MASK 0x1FFu
#define FW_NODE_BOUNDARY 0x8000u
#define FW_NODE_BCV_SHIFT 16

/* fw_step flags */
#define FW_STEP_DIAG 1u      /* last substep writes element stress/strain, nodal force, flags */
#define FW_STEP_NO_CONTACT 2u /* skip ground contact (femStep alone) */
#define FW_STEP_PER_SCENE 4u  /* one launch per scene per substep (the unbatched launch pattern) */

/* fields for fw_scene_read / fw_scene_write (node fields in the caller's node order) */
#define FW_FIELD_X 1        /* float32[N][2] positions */
#define FW_FIELD_V 2        /* float32[N][2] velocities */
#define FW_FIELD_U 3        /* float32[N][2] accumulated displacement */
#define FW_FIELD_FORCE 4    /* float32[N][2] nodal force of the last DIAG substep (after BCs) */
#define FW_FIELD_MASS 5     /* float32[N] lumped mass */
#define FW_FIELD_STRESS 6   /* float32[E][3] Cauchy stress (xx, yy, xy), last DIAG substep */
#define FW_FIELD_STRAIN 7   /* float32[E][3] Green strain (xx, yy, xy), last DIAG substep */
#define FW_FIELD_EPS_P 8    /* float32[E][3] plastic strain */
#define FW_FIELD_EQPS 9     /* float32[E] equivalent plastic strain */
#define FW_FIELD_AREA 10    /* float32[E] rest area */
#define FW_FIELD_JINV 11    /* float32[E][4] inverse rest Jacobian, row-major */
#define FW_FIELD_YIELD 12   /* uint8[E] f_trial > 0 in the last DIAG substep */
#define FW_FIELD_CONTACT 13 /* uint8[N] a ground pushed the node in the last DIAG substep */
#define FW_FIELD_AABB 14    /* float32[4] boundary-node AABB lb.x lb.y ub.x ub.y at the current x */

/* One body-vs-ground contact pair (one work-list entry group of
 * flatworld/mixedcontact.py:1794-1856, 2131-2261).  A spring scene's penalty is
 * 100 x its stiffness (flatworld/femcontact.py:256-279) and every one of its nodes
 * is a work item: flag them all FW_NODE_BOUNDARY. */
typedef struct {
    int32_t kind;       /* FW_GROUND_* */
    float penalty;      /* float32 penalty k = 2K (flatworld/femcontact.py:100-119) */
    float aabb_lo[2];   /* ground AABB used by the activation gate */
    float aabb_hi[2];
    float point[2];     /* plane */
    float normal[2];
    const float* heights; /* height field samples [hf_nx] */
    int32_t hf_nx;
    float hf_lb_x;
    float hf_ub_x;
    int32_t reverse;
    const int32_t* occ; /* voxel occupancy occ[i * vox_ny + j], i along x */
    int32_t vox_nx;
    int32_t vox_ny;
    float vox_lb[2];
    float vox_dx;
    float vox_dz;
    int32_t flex_cap;   /* FW_GROUND_FLEX: contribution capacity */
    /* FW_GROUND_RIGID: FW_RIGID_* type; centre (rigidParams[i, 0]); rotation matrix,
       row-major [[cos a, -sin a], [sin a, cos a]]; primary parameters (rigidParams[i, 1]:
       box full extents, capsule axis end lc relative to the centre); radius (ball,
       capsule); linear and angular velocity (friction only); the rigid's friction
       coefficient; 1 for a tied pair */
    int32_t rigid_type;
    float rigid_center[2];
    float rigid_rot[4];
    float rigid_prim[2];
    float rigid_radius;
    float rigid_vel[2];
    float rigid_omega;
    float rigid_friction;
    int32_t rigid_tied;
    /* FW_RIGID_MESH: boundary node positions [2 x mesh_nodes], edges as node index
       pairs [2 x mesh_edges], edge AABBs [2 x mesh_edges each] and margin bases
       [mesh_edges]; the hash: registry (rigid id, edge) [2 x mesh_nreg], per cell
       (start, end) [2 x mesh_tc], entries [mesh_npairs], grid origin, cell size and
       cells per axis; the capacities reserved at admission for the registry, the
       cells and the entries; the record's rigid id, the substeps between full hash
       queries and the local step before which the skip epochs reset (-1: none).
       mesh_update (fw_scene_set_rigid) selects what is copied: 1 positions, 2 edge
       AABBs and margins, 4 hash; the header always. */
    const float* mesh_coords;
    const int32_t* mesh_conn;
    const float* mesh_elb;
    const float* mesh_eub;
    const float* mesh_margin;
    const int32_t* mesh_reg;
    const int32_t* mesh_cell;
    const int32_t* mesh_sorted;
    int32_t mesh_nodes, mesh_edges, mesh_nreg, mesh_tc, mesh_npairs;
    int32_t mesh_cap_reg, mesh_cap_cells, mesh_cap_pairs;
    float mesh_glb[2];
    float mesh_gs[2];
    int32_t mesh_cn[2];
    int32_t mesh_did, mesh_subcycle, mesh_reset_at;
    int32_t mesh_update;
} fw_ground_desc;

/* One FEM or spring-mass body; arrays in the caller's node/element order.
 * FW_MAT_SPRING: conn holds [2E] node ids, E is the stiffness (matE), rho the mass
 * of every node, spring_c the spring damping coefficient (matDamping) and rest_len
 * the [E] rest lengths (flatworld/femspringmanager.py:846-862); nu, sigma_y and H
 * are unused. */
typedef struct {
    int32_t n_nodes;
    int32_t n_elems;
    int32_t n_grounds;
    int32_t n_bc_values;
    const float* x0;        /* [2N] rest = initial positions */
    const float* v0;        /* [2N] initial velocities */
    const int32_t* conn;    /* [3E] triangle node ids; [2E] for springs */
    const uint32_t* flags;  /* [N] node flag words (FW_NODE_*) */
    const float* bc_values; /* [2 * n_bc_values] BC value table */
    int32_t mat_type;       /* FW_MAT_* */
    float E, nu, rho, sigma_y, H;
    float dt;               /* substep dt (float32), used when fw_step gets dt_override < 0 */
    float damping;
    int32_t horizon;        /* substeps to run after the warm-up; 0 = unbounded */
    int32_t track_disp;     /* 1: accumulate U every substep (the reference always does) */
    const fw_ground_desc* grounds;
    const float* rest_len;  /* FW_MAT_SPRING: [E] rest lengths; NULL otherwise */
    float spring_c;         /* FW_MAT_SPRING: damping coefficient */
} fw_scene_desc;

typedef struct {
    int32_t tile_nodes;   /* max owned nodes per tile */
    int32_t tile_slots;   /* max element slots (owned + halo) per tile: 512, 768, 1024 or 2048 */
    int32_t block;        /* threads per workgroup: 256 */
    int32_t grid_per_cu;  /* persistent workgroups per CU; 0 = one workgroup per tile */
} fw_tuning;

typedef struct {
    int64_t capacity_bytes;
    int64_t used_bytes;
    int32_t n_scenes;
    int32_t n_tiles;
    int64_t n_nodes;
    int64_t n_elems;
    int64_t n_elems_j2;
    int64_t global_step;       /* substeps launched so far on this arena */
    double algo_bytes_per_substep; /* sum of E*B_e + N*B_n over resident scenes */
    int64_t halo_slots;        /* element slots evaluated on top of owned ones */
    int32_t n_cu;              /* compute units of the device */
    int32_t tiles_per_cu;      /* substep workgroups resident per CU for the tuning (HIP occupancy) */
} fw_arena_info;

/* FEM-FEM contact group: the FEM bodies of one loop (scenes added to the same
 * arena before the same substep) and the batched work list of their FEM-FEM pairs
 * (flatworld/mixedcontact.py:1857-1911).  Node ids are body-local, in the
 * caller's order.  Segments are every body's boundary edges, body by body, each
 * body's in boundary-edge order (flatworld/femspringmanager.py:780-799). */
typedef struct {
    int32_t n_bodies;
    const int32_t* scene_ids;     /* [n_bodies] */
    const int32_t* body_seg_off;  /* [n_bodies + 1] */
    const int32_t* seg_node;      /* [2 * n_segs] */
    int32_t n_segs;
    int32_t n_items;
    const int32_t* item_body;     /* body of the contact node */
    const int32_t* item_node;     /* its body-local id (a boundary node) */
    const int32_t* item_target;   /* body whose segments it is checked against */
    const int32_t* item_pair;
    const float* item_penalty;
    const float* item_friction;
    int32_t n_pairs;
    const int32_t* pair_body;     /* [2 * n_pairs] */
    const int32_t* body_bnodes_off; /* [n_bodies + 1] */
    const int32_t* body_bnodes;   /* boundary nodes of each body (body-local ids) */
    int32_t gated;                /* 0: every pair is active (the loop skips its BVH) */
    int32_t max_cells;            /* spatial hash limits (SpatialHashManager defaults: 100000, 512) */
    int32_t max_query;
    int32_t max_pairs;
    /* [n_items] 1 where the item's pair is tied (ExplicitLoop ties=, contact.tied),
       0 elsewhere; NULL: no tied pair (flatworld/mixedcontact.py:1871-1932) */
    const int32_t* item_tied;
} fw_flex_desc;

/* fields for fw_flex_read */
#define FW_FLEX_CACHE 1      /* int32[n_items] cached segment per work item, -1 for none */
#define FW_FLEX_GRID 2       /* 8 x 4 bytes: lower corner x, y and cell size x, y (float32); cells x, y,
                                cells used, exact flag (int32) */
#define FW_FLEX_CELL_START 3 /* int32[max_cells] */
#define FW_FLEX_CELL_END 4   /* int32[max_cells] */
#define FW_FLEX_SORTED 5     /* int32[max_pairs] segment ids, ascending within each cell */
#define FW_FLEX_TIE_SEG 6    /* int32[n_items] resolved tie edge per work item, -1 for none */
#define FW_FLEX_TIE_STATE 7  /* float32[4 * n_items] weight 0, weight 1, normal gap, 0 */

typedef struct fw_arena fw_arena;
typedef struct fw_prepared fw_prepared;

const char* fw_last_error(void);
int32_t fw_abi_version(void);

/* Arena on `device` with `capacity_bytes` of HBM (0 = 80% of free memory). */
int32_t fw_arena_create(int32_t device, int64_t capacity_bytes, const fw_tuning* tuning, fw_arena** out);
int32_t fw_arena_destroy(fw_arena* a);
int32_t fw_arena_get_info(fw_arena* a, fw_arena_info* out);

/* Bytes of arena memory a scene needs (admission control). */
int32_t fw_scene_footprint(const fw_arena* a, const fw_scene_desc* desc, int64_t* bytes);
/* Copy a scene in, compute area / J^-1 / lumped mass on the device and schedule
 * its zero-dt warm-up for the next substep.  Returns FW_E_CAPACITY when the
 * arena cannot hold it (nothing is changed then). */
int32_t fw_scene_add(fw_arena* a, const fw_scene_desc* desc, int32_t* scene_id);

/* fw_scene_add in two halves, so that a server can do the host work of a scene
 * (validation, tiling, packing its block) on worker threads while the GPU runs:
 *   fw_scene_prepare      host only, thread-safe, no GPU and no arena needed; the
 *                         scene is tiled for `tuning` (NULL: the defaults) and the
 *                         description's arrays are copied, so they can be freed;
 *   fw_prepared_footprint arena bytes the prepared scene will take (admission);
 *   fw_scene_add_prepared copies it into an arena of the same kernel shape
 *                         (FW_E_CAPACITY, nothing changed, when it does not fit;
 *                         it can be retried later); a prepared scene is added once;
 *   fw_prepared_free      releases it (added or not).
 * Same reference interfaces as fw_scene_add. */
int32_t fw_scene_prepare(const fw_scene_desc* desc, const fw_tuning* tuning, fw_prepared** out);
int32_t fw_prepared_footprint(const fw_prepared* p, int64_t* bytes);
int32_t fw_scene_add_prepared(fw_arena* a, fw_prepared* p, int32_t* scene_id);
int32_t fw_prepared_free(fw_prepared* p);
/* Move a prepared scene's image into page-locked memory (on the worker thread that
 * prepared it), so that fw_scene_add_prepared issues the upload without a host
 * copy.  Needs the GPU runtime; without page-locked memory it changes nothing.  The
 * buffers are pooled per process and reused once their upload has completed.  No
 * reference counterpart. */
int32_t fw_prepared_stage(fw_prepared* p);
int32_t fw_scene_remove(fw_arena* a, int32_t scene_id);
/* fw_scene_remove with the scene's final state kept: the copies of its positions
 * (and velocities, when v_out is not NULL) are queued on the arena stream behind
 * every substep launched so far, then the scene is removed; fw_retire_wait waits
 * for the copies of every retirement queued since the last wait (not for later
 * work on the stream) and writes them, in the caller's node order, to x_out /
 * v_out (N x 2 floats each), which must stay valid until then.  A server retires
 * finished jobs and launches its next frame before waiting, so the device does not
 * idle while results come back.  No reference counterpart. */
int32_t fw_scene_retire(fw_arena* a, int32_t scene_id, float* x_out, float* v_out);
int32_t fw_retire_wait(fw_arena* a);
/* Restart the scene's BVH-rebuild schedule: the next substep counts as step 0
 * of the activation gate (ExplicitLoop.reset, flatworld/explicitloop.py:635-679). */
int32_t fw_scene_reset_gate(fw_arena* a, int32_t scene_id);
/* Selective mass scaling (FemSpringManager.applyMassScaling): enqueued behind the
 * substeps already launched, so it sees the current positions.  Every triangle
 * whose critical step 0.4 L / sqrt(E / rho) (L its shortest edge) is below
 * dt_target adds (E (dt_target / (0.4 L))^2 - rho) * area / 3 to each of its
 * nodes; every spring end node is raised to at least 4 k dt_target^2.
 * ExplicitLoop calls it right after the warm-up (flatworld/explicitloop.py:164-170). */
int32_t fw_scene_mass_scaling(fw_arena* a, int32_t scene_id, float dt_target);
/* Move a FW_GROUND_RIGID record of a scene to a new pose: of `g` the rigid_* fields
 * and aabb_lo / aabb_hi are read (kind must be FW_GROUND_RIGID, and rigid_type and
 * rigid_tied must equal the record's).  Stream-ordered: substeps launched after the
 * call see the new pose, those launched before the old one.  Replaces the pose
 * update the reference's RigidManager does between substeps
 * (flatworld/rigidmanager.py:765-793, 2113-2200), for a rigid the caller moves. */
int32_t fw_scene_set_rigid(fw_arena* a, int32_t scene_id, int32_t ground_index, const fw_ground_desc* g);
/* Substeps completed by the scene (warm-up excluded) and whether it reached its horizon. */
int32_t fw_scene_progress(fw_arena* a, int32_t scene_id, int64_t* steps_done, int32_t* finished);

/* FEM-FEM contact groups.  fw_flex_add needs every body in a pair to carry one
 * FW_GROUND_FLEX record of sufficient capacity, and all of the group's scenes
 * added before the same substep with the same horizon; they cannot be removed
 * before the group.  fw_flex_rebuild enqueues the spatial-hash build at the state
 * the next substep starts from (after the scenes' warm-up), for one group or, with
 * group_id -1, for every group in one launch; the caller keeps the
 * reference's rebuild schedule (every 1/500 s of simulated time,
 * flatworld/femspringmanager.py:1035-1039).  fw_step then runs the group's
 * contact pass before every substep.  fw_flex_schedule_rebuild puts the same
 * rebuild on the device instead, before substep `at` (0: the next one) of the
 * following fw_step / fw_step_graph / fw_step_noalloc calls, for one group or,
 * with group_id -1, every group; a group holds up to FW_FLEX_SCHED pending
 * rebuilds (FW_E_CAPACITY beyond).  A captured frame checks every group's schedule
 * before each substep, so fw_step_graph replays groups whatever their rebuild
 * schedule.  FW_STEP_PER_SCENE is refused while a group exists.
 * Tied work items never slide: fw_flex_init_ties enqueues, at the state the next
 * substep starts from and with the group's current hash, the one-time resolution
 * of every tied, unresolved item to its nearest edge
 * (_initialize_fem_fem_ties_once_kernel, flatworld/mixedcontact.py:687-770, which
 * run_batched_contacts launches once, after the hash refresh of the first
 * substep, mixedcontact.py:1668-1689); from then on every substep's contact
 * pass applies a penalty spring toward the resolved edge point at the initial
 * normal gap (_apply_fem_fem_ties_fixed_kernel, mixedcontact.py:773-811), each
 * node's tie contributions before its sliding-contact ones.  It returns
 * FW_E_INVALID for a group with tied items whose hash was never built. */
int32_t fw_flex_add(fw_arena* a, const fw_flex_desc* desc, int32_t* group_id);
#define FW_FLEX_SCHED 4
int32_t fw_flex_rebuild(fw_arena* a, int32_t group_id);
int32_t fw_flex_schedule_rebuild(fw_arena* a, int32_t group_id, int32_t at);
int32_t fw_flex_init_ties(fw_arena* a, int32_t group_id);
int32_t fw_flex_remove(fw_arena* a, int32_t group_id);
int32_t fw_flex_read(fw_arena* a, int32_t group_id, int32_t field, void* out, int64_t out_bytes);

/* Enqueue n substeps of every resident scene on the arena stream (asynchronous);
 * n == 0 only uploads the device tables that adding or removing scenes and groups
 * left pending.
 * dt_override >= 0 replaces every scene's dt and damping_override >= 0 every
 * scene's damping for these substeps (FemSpringManager.substep(dt, damping),
 * flatworld/femspringmanager.py:938-939).  With
 * FW_STEP_PER_SCENE each substep is issued as one launch per scene instead of
 * one launch for the whole arena; results are identical. */
int32_t fw_step(fw_arena* a, int32_t n, float dt_override, float damping_override, uint32_t flags);
/* fw_step for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation).  No reference counterpart: it
 * exists for those callers only.  It calls no hipMalloc / hipFree and no allocating
 * library routine, and keeps the FEM-FEM contact pass's per-substep scratch (per
 * work item results, per slot counts and cursors, the entry lists) in one fixed
 * device array of FW_NOALLOC_SCRATCH_BYTES per GPU instead of the groups' arena
 * blocks.  Results are identical to fw_step's.  Limits: the live groups' scratch
 * must fit that array (FW_E_CAPACITY otherwise), and after scenes or groups are
 * added or removed, fw_step(a, 0, ...) must run once first to upload the tables
 * (FW_E_INVALID otherwise).  Every call on one GPU shares that GPU's array: calls on
 * different arenas of one GPU that overlap in time (other host threads, other
 * streams) race, and must be serialised by the caller.  Its calls are not counted
 * by fw_arena_busy.  Error messages have static storage. */
#define FW_NOALLOC_SCRATCH_BYTES (32 << 20)
int32_t fw_step_noalloc(fw_arena* a, int32_t n, float dt_override, float damping_override, uint32_t flags);

/* n substeps captured once as a HIP graph (per n and flags) and replayed. */
int32_t fw_step_graph(fw_arena* a, int32_t n, uint32_t flags);
/* Capture and instantiate the graph fw_step_graph(a, n, flags) replays, without
 * launching it (a caller that times frames, or must not capture inside its loop,
 * prepares its frame lengths first).  No reference counterpart. */
int32_t fw_step_graph_prepare(fw_arena* a, int32_t n, uint32_t flags);
int32_t fw_sync(fw_arena* a);

/* GPU-busy accounting for a server loop.  With tracking on, every fw_step and
 * fw_step_graph call is bracketed by a pair of HIP events on the arena stream;
 * fw_arena_busy waits for the bracketed work and returns the summed GPU time of
 * those calls in milliseconds and their number, since tracking was first turned
 * on.  Host work between calls (admission, tiling) is outside every bracket.  No
 * reference counterpart: the reference has no served form. */
int32_t fw_arena_track_busy(fw_arena* a, int32_t enable);
int32_t fw_arena_busy(fw_arena* a, double* busy_ms, int64_t* calls);
/* With busy tracking on: block until at most `keep` of the tracked stepping calls
 * issued so far are unfinished on the device (keep 1: a server that has queued its
 * next frame waits for the previous one).  No reference counterpart. */
int32_t fw_wait_launches(fw_arena* a, int32_t keep);

/* Host-only self-check of the tiling fw_scene_add would build (no GPU needed):
 * stats[0] tiles, [1] halo slots, [2] largest tile slot count, [3] CSR entries.
 * Returns FW_E_INVALID with a message if any invariant fails. */
int32_t fw_tiling_check(const fw_scene_desc* desc, int32_t tile_nodes, int32_t tile_slots, int64_t* stats);

/* fw_scene_prepare keeps the tilings it builds, by connectivity (and, for springs,
 * rest lengths) and kernel shape: a later scene with the same connectivity reuses the
 * tiling, whatever its node positions, since results do not depend on the tiling.
 * Keys are compared in full; at most 512 MiB of host memory, least recently used
 * first out.  No reference counterpart (the reference builds no tiling): the
 * served form's admission cost.  Counters since the last clear; clear also empties
 * the cache. */
int32_t fw_tiling_cache_stats(int64_t* hits, int64_t* misses, int64_t* bytes, int32_t* entries);
int32_t fw_tiling_cache_clear(void);

/* Synchronous copies between a scene and host memory (caller's order). */
int32_t fw_scene_read(fw_arena* a, int32_t scene_id, int32_t field, void* out, int64_t out_bytes);
int32_t fw_scene_write(fw_arena* a, int32_t scene_id, int32_t field, const void* in, int64_t in_bytes);

/* Time n substeps of all scenes with HIP events on the arena stream.
 * mode 0: fw_step, 1: fw_step_graph, 2: fw_step with FW_STEP_PER_SCENE.
 * Returns total milliseconds and the average duration of one substep kernel. */
int32_t fw_time_steps(fw_arena* a, int32_t n, int32_t mode, double* total_ms, double* kernel_ms);

/* Calibration: a plain float4 copy of `bytes` bytes, `reps` times; reports the
 * read+write bandwidth and the average kernel time.  No reference counterpart:
 * it measures the achievable HBM bandwidth the substep is compared against. */
int32_t fw_bandwidth_probe(int32_t device, int64_t bytes, int32_t reps, double* gbs, double* kernel_ms);

/* Calibration: the fastest read+write stream of `bytes` bytes the device sustains,
 * over a sweep of float4 loads in flight per thread (1, 2, 4, 8) and workgroups per
 * CU (2 to 32), each `reps` times; reports that bandwidth, its kernel time and the
 * winning grid and loads per thread.  No reference counterpart: it is the ceiling
 * the substep's HBM rate is read against. */
int32_t fw_bandwidth_ceiling(int32_t device, int64_t bytes, int32_t reps, double* gbs, double* kernel_ms,
                             int32_t* grid_per_cu, int32_t* ilp);

#ifdef __cplusplus
}
#endif

#endif
