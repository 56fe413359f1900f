/*
 * fem_oracle.h -- CPU restatement of FlatWorld's explicit-FEM substep.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product path in flatworld_amd never calls
 * it.
 *
 * Each routine runs serially over "thread ids" in ascending order, which is how
 * Warp executes a kernel launch on its CPU device, so float accumulations
 * (Fext atomics, lumped mass) happen in the reference's order.  All arithmetic
 * is IEEE float32 without contraction (built with -ffp-contract=off);
 * log(J) is a float log built from double operations (ora_log_f32), faithful
 * to within one float ulp and pinned against numpy's double log.
 */
#ifndef FW_FEM_ORACLE_H
#define FW_FEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ORA_SPRING: a spring-mass domain (flatworld/springmass.py); its elements are
 * two-node springs, conn holds [2E] node ids, E is the stiffness, rho the mass of
 * every node and spring_c the damping coefficient (femspringmanager.py:814-870). */
enum { ORA_ELASTIC = 1, ORA_NEOHOOKEAN = 2, ORA_MISES = 3, ORA_SPRING = 4 };
/* ORA_FLEX: not a ground but the FEM-FEM contact contributions of the substep
 * (ora_flex_contacts), placed in the ground list where the reference's batched
 * FEM-FEM kernel runs: after the plane kernel, before the height-field and voxel
 * kernels (flatworld/mixedcontact.py:1654-1754).  It is never gated. */
enum { ORA_PLANE = 1, ORA_HEIGHTFIELD = 2, ORA_VOXEL = 3, ORA_FLEX = 4, ORA_RIGID = 5 };
/* rigid types (flatworld/definitions.py:50-55) */
enum { ORA_RIGID_BALL = 1, ORA_RIGID_BOX = 2, ORA_RIGID_CAPSULE = 8, ORA_RIGID_MESH = 16 };

typedef struct {
    int32_t kind;
    float penalty;
    float aabb_lo[2];
    float aabb_hi[2];
    /* plane */
    float point[2];
    float normal[2];
    /* height field */
    const float* heights;
    int32_t hf_nx;
    float hf_lb_x;
    float hf_ub_x;
    int32_t reverse;
    /* voxel grid, occ[i * ny + j] */
    const int32_t* occ;
    int32_t vox_nx;
    int32_t vox_ny;
    float vox_lb[2];
    float vox_dx;
    float vox_dz;
    /* ORA_FLEX: contributions to node n are flex_f[2j .. 2j+1], j in [flex_off[n], flex_off[n+1]) */
    const int32_t* flex_off;
    const float* flex_f;
    /* ORA_RIGID: a rigid primitive at a given pose (aabb_lo / aabb_hi hold its AABB at
       that pose): type, centre, rotation matrix (row-major), primary parameters (box:
       full extents; capsule: one axis end relative to the centre), radius, linear and
       angular velocity (friction), friction coefficient and tied flag */
    int32_t rigid_type;
    float rigid_center[2];
    float rigid_rot[4];
    float rigid_prim[2];
    float rigid_radius;
    float rigid_vel[2];
    float rigid_omega;
    float rigid_friction;
    int32_t rigid_tied;
    /* ORA_RIGID of type ORA_RIGID_MESH: the loop's mesh rigids as the RigidManager
       keeps them and its spatial hash over their boundary edges, built on the host:
       boundary node positions [2 x nodes]; edges as two node indices [2 x elems];
       each edge's AABB and margin base [2 x elems each, elems]; the hash registry
       (rigid id, edge) [2 x registered], per cell (start, end) [2 x tc], the cell
       entries (registry indices) and the grid (origin, cell size, cells per axis,
       cell count).  mesh_did is the record's rigid; full queries run on substeps
       counter % mesh_subcycle == 0 and the skip epochs reset before the substep
       counter == mesh_reset_at (the host's schedule).  Per work item (the body's boundary nodes in order) the state
       the reference keeps in arrays: cached (edge, rigid), skip epoch, friction
       history (edge, valid, weights, force, penetration), tie (resolved, weights). */
    const float* mesh_coords;
    const int32_t* mesh_conn;
    const float* mesh_elb;
    const float* mesh_eub;
    const float* mesh_margin;
    const int32_t* mesh_reg;
    const int32_t* mesh_cell;
    const int32_t* mesh_sorted;
    float mesh_glb[2];
    float mesh_gs[2];
    int32_t mesh_cn[2];
    int32_t mesh_tc;
    int32_t mesh_did;
    int32_t mesh_subcycle;
    int32_t mesh_reset_at;
    int32_t* st_cache;   /* [2 x items] (edge, rigid), -1 when empty */
    int32_t* st_skip;    /* [items] skip epoch, -1 or 1 */
    int32_t* st_pelem;   /* [items] friction: previous edge (-1) */
    int32_t* st_pvalid;  /* [items] */
    float* st_pw;        /* [2 x items] */
    float* st_pf;        /* [2 x items] */
    float* st_ppen;      /* [items] */
    int32_t* st_tres;    /* [items] tie resolved */
    float* st_tw;        /* [2 x items] tie weights */
} ora_ground;

typedef struct {
    int32_t n_nodes;
    int32_t n_elems;
    int32_t n_bnodes;
    int32_t n_grounds;
    float* x;        /* [2N] positions (x, y) */
    float* v;        /* [2N] velocities */
    float* u;        /* [2N] accumulated displacement */
    float* fext;     /* [2N] nodal force of the last substep (after BCs) */
    float* mass;     /* [N] lumped mass, filled by ora_prepare */
    const int32_t* conn;   /* [3E] triangles, [2E] springs */
    const int32_t* bc;     /* [N] BC bits */
    const float* bcv;      /* [2N] BC value vector */
    const int32_t* bnodes; /* [Nb] boundary nodes */
    float* jinv;     /* [4E] row-major inverse reference Jacobian */
    float* area;     /* [E] */
    int32_t mat_type;
    float E, nu, rho, sigma_y, H;
    float* eps_p;    /* [3E] plastic strain (xx, yy, xy) */
    float* eqps;     /* [E] equivalent plastic strain */
    float* stress;   /* [3E] Cauchy stress (xx, yy, xy) of the last substep */
    float* strain;   /* [3E] Green strain (xx, yy, xy) of the last substep */
    uint8_t* yield_flag;   /* [E] f_trial > 0 in the last substep (J2 only) */
    uint8_t* contact_flag; /* [N] a ground pushed this node in the last substep */
    float aabb[4];   /* boundary-node AABB lb.x, lb.y, ub.x, ub.y at the current x */
    const ora_ground* grounds;
    const float* rest_len; /* [E] spring rest lengths (ORA_SPRING) */
    float spring_c;        /* spring damping coefficient (ORA_SPRING) */
} ora_scene;

/* flatworld/femspringmanager.py:113-138 (area, J^-1, lumped mass by atomics);
 * a spring domain's nodes all take the domain mass (femspringmanager.py:852-853). */
void ora_prepare(ora_scene* s);

/* FemSpringManager.applyMassScaling (flatworld/femspringmanager.py:141-201, 899-924),
 * run after the warm-up as ExplicitLoop does (explicitloop.py:164-170).  Returns the
 * added mass as the reference accumulates it. */
float ora_mass_scaling(ora_scene* s, float dt_target);

/* One ExplicitLoop.advance() for a one-FEM-body loop
 * (flatworld/explicitloop.py:473-555): gate, ground forces, element forces,
 * integrate, AABB.  with_contact=0 gives FemSpringManager.femStep alone (the
 * zero-dt warm-up of flatworld/femspringmanager.py:872-879). */
void ora_substep(ora_scene* s, float dt, float damping, int32_t counter, int32_t with_contact);

/* Element-level building blocks, exported for unit tests. */
void ora_stress2d(int32_t type, float E, float nu, const float F[4], float stress[3], float strain[3]);
int32_t ora_mises2d(float E, float nu, float sy, float H, const float F[4], const float eps_p_old[3],
                    float eqps_old, float stress[3], float eps_p_new[3], float* eqps_new, float strain[3]);
void ora_element_force(const ora_scene* s, int32_t e, float f[6], float cauchy[3], float strain[3],
                       float eps_p_new[3], float* eqps_new, int32_t* yielded);
/* flatworld/femspringmanager.py:283-312: the force node 0 of spring e receives
 * with a minus sign and node 1 with a plus sign. */
void ora_spring_force(const ora_scene* s, int32_t e, float f[2]);
int32_t ora_ground_force(const ora_ground* g, const float p[2], const float vel[2], float dt, float force[2]);
/* detectPointToPrimitive (flatworld/contact_detection.py:147-191): penetration,
 * normal and contact point of point p against the rigid primitive of record g. */
float ora_rigid_detect(const ora_ground* g, const float p[2], float normal[2], float cpoint[2]);
/* AABB of a rigid primitive at its pose: getBallBBox / getBoxBBox / getCapsuleBBox
 * (flatworld/utils.py:235-302). */
void ora_rigid_aabb(int32_t type, const float center[2], const float rot[4], const float prim[2], float radius,
                    float lb[2], float ub[2]);
int32_t ora_gate(const float fem_lb[2], const float fem_ub[2], const ora_ground* g, int32_t counter);
void ora_update_bbox(ora_scene* s);
float ora_log_f32(float x);

/* ------------------------------------------------------------------ FEM-FEM contact
 * The batched FEM-FEM penalty contact with the FEM spatial hash
 * (flatworld/mixedcontact.py:886-1025, spatialmanager.py, femspringmanager.py:421-487,
 * 1035-1073), restated with the reference's kernels run serially in thread order.
 * The reference runs them in parallel: its spatial-hash scatter and its Fext
 * updates are atomics, and every thread of the contact kernel shares one query
 * buffer.  Serial thread order is the one execution of those kernels that is
 * race-free; it fixes the order of a hash cell's entries (ascending segment), of
 * each node's contributions (ascending work item) and gives every work item its
 * own query list. */
typedef struct {
    /* boundary segments of the loop's FEM bodies: body by body in loop order, each
       body's boundary edges in insertion order (femspringmanager.py:780-799) */
    int32_t n_bodies;
    int32_t n_segs;
    const int32_t* body_seg_off; /* [n_bodies + 1] */
    const int32_t* seg_body;     /* [n_segs] */
    const int32_t* seg_node;     /* [2 n_segs] body-local node ids */
    /* work items in the reference's order (mixedcontact.py:1862-1911) */
    int32_t n_items;
    const int32_t* item_body;    /* body of the contact node */
    const int32_t* item_node;    /* body-local node id */
    const int32_t* item_target;  /* body whose segments the node is checked against */
    const int32_t* item_pair;
    const float* item_penalty;
    const float* item_friction;
    int32_t* cache;              /* [n_items] cached segment (global index) or -1 */
    int32_t n_pairs;
    const int32_t* pair_body;    /* [2 n_pairs] */
    int32_t gated;               /* 0: every pair active (the loop skips its BVH) */
    /* tied pairs (mixedcontact.py:687-811): per work item the tie flag and the
       state ora_flex_ties_init resolves once (edge, weights, normal gap) */
    const int32_t* item_tied;    /* [n_items] 1: the item's pair is tied */
    int32_t* tie_elem;           /* [n_items] resolved edge or -1 */
    float* tie_w;                /* [2 n_items] */
    float* tie_gap;              /* [n_items] */
    /* spatial hash */
    int32_t max_cells, max_query, max_pairs;
    float glb[2], gub[2], gs[2];
    int32_t cn[2], total_cells, n_hash_pairs;
    int32_t* cell_start;         /* [max_cells] */
    int32_t* cell_end;           /* [max_cells] */
    int32_t* sorted_seg;         /* [max_pairs] */
    /* output of ora_flex_contacts, per body: node CSR [N_b + 1] and float2 entries */
    int32_t** c_off;
    float** c_f;
} ora_flex;

/* pointToEdgeContact (flatworld/contact_detection.py:15-61): returns is_inside. */
int32_t ora_point_edge(const float p[2], const float a[2], const float b[2], float* pen, float normal[2], float w[2]);
/* FemSpringManager.populate_fem_spatial_hash + SpatialHashManager.build at the
 * bodies' current positions and AABBs (velocity buffer 0). */
void ora_flex_hash_build(ora_flex* f, ora_scene* const* bodies);
/* query_point_with_buffer (spatialmanager.py:131-199): segments of body `target`
 * in the cells the buffered point overlaps, deduplicated, at most max_query. */
int32_t ora_flex_query(const ora_flex* f, const float p[2], float buffer, int32_t target, int32_t* out);
/* _initialize_fem_fem_ties_once_kernel (mixedcontact.py:687-770) at the bodies'
 * current state: every tied, unresolved work item takes its nearest edge. */
void ora_flex_ties_init(ora_flex* f, ora_scene* const* bodies);
/* One run of the FEM-FEM kernels at the bodies' current state: the tie springs of
 * resolved tied items (mixedcontact.py:773-811), then the sliding contact of the
 * others (886-1025); updates the cache and writes each body's contribution lists
 * (c_off / c_f): per node the tie contributions in ascending work item, then the
 * contact contributions in ascending work item, the order of the two launches. */
void ora_flex_contacts(ora_flex* f, ora_scene* const* bodies, int32_t counter);
int32_t ora_gate_pair(const float alb[2], const float aub[2], const float blb[2], const float bub[2],
                      int32_t counter);

#ifdef __cplusplus
}
#endif

#endif
