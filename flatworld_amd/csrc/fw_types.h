// fw_types.h -- device-resident tables of the FEM arena (shared by host runtime and kernels).
//
// HBM layout per scene (all SoA, internal node/element order, 256-byte aligned blocks):
//   nodes    pos[2] float2 (ping-pong: substep k reads pos[(k+1)&1], writes pos[k&1]),
//            rest float2 (reference positions), vel float2, mass float, flag uint32,
//            disp float2 (optional), force float2 (diag), cflag uint8 (diag)
//   halo     hpos[2] float2 per (tile, halo node): copies of neighbour-owned positions,
//            written by the owner tile in the same substep as pos (same ping-pong);
//            hrest float2 per (tile, halo node): their rest positions (triangles);
//            hvel[2] float2 per (tile, halo node): velocity copies (springs, ping-pong)
//   elements jinv float4 + area float (written once, for readback only: the substep
//            recomputes both from the rest positions), j2[2] float4 (eps_xx, eps_yy,
//            eps_xy, eqps; ping-pong, J2 scenes only), stress / strain float4 (diag),
//            yflag uint8 (diag)
//            spring scenes hold no element arrays: a spring's rest length rides in its
//            slot record
//   tiles    rec uint2 per element slot, hel int32 per halo slot, csr_ptr uint16 per
//            owned node (+1), push int2 per halo copy, grouped by the tile owning the
//            source node
//   ground   height samples float / voxel cell extents uint16 x4
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define FW_MAX_GROUNDS_DEV 8
#define FW_MAT_SPRING_DEV 4  // FW_MAT_SPRING of the public header
#define FW_GROUND_FLEX_DEV 4 // FW_GROUND_FLEX: the scene's FEM-FEM contact contributions
#define FW_GROUND_RIGID_DEV 5 // FW_GROUND_RIGID: a rigid primitive at the pose the host set
#define FW_GROUND_RIGID_MESH_DEV 6 // FW_GROUND_RIGID of type FW_RIGID_MESH: never gated by the
                                   // substep (the rigid pass gates its items itself)
#define FW_RIGID_MESH_TYPE 16  // FW_RIGID_MESH

// A mesh rigid record's header, at the start of its data block after the per-node
// force array and the boundary-node ids: the loop's mesh rigids and the spatial
// hash over their boundary edges (built on the host), and the offsets, from the
// block start, of the arrays the rigid pass reads and writes.
struct MeshDev {
    float glb[2], gs[2];  // grid origin and cell size
    int32_t cn[2], tc;    // cells per axis, cell count
    int32_t did;          // the record's rigid id
    int32_t subcycle;     // full hash queries on local steps % subcycle == 0
    int32_t reset_at;     // local step before which the skip epochs reset (-1: none)
    int32_t n_items;      // work items (the scene's boundary nodes)
    int32_t pad;
    int64_t o_coords;     // float2 [mesh nodes] boundary node positions
    int64_t o_conn;       // int2 [edges] node indices
    int64_t o_elb, o_eub; // float2 [edges] edge AABBs
    int64_t o_margin;     // float [edges] margin bases
    int64_t o_reg;        // int2 [registered] (rigid id, edge)
    int64_t o_cell;       // int2 [tc] (start, end)
    int64_t o_sorted;     // int32 [pairs] registry indices
    int64_t o_cache;      // int2 [items] cached (edge, rigid id)
    int64_t o_skip;       // int32 [items] skip epoch
    int64_t o_fric_i;     // int2 [items] friction history: (edge, valid)
    int64_t o_fric_f;     // float4 [items] (weight 0, weight 1, force x, force y)
    int64_t o_fric_pen;   // float [items] penetration
    int64_t o_tie;        // float4 [items] (resolved, weight 0, weight 1, 0)
};

struct GroundDev {
    int32_t kind;
    float penalty;
    float lo[2], hi[2];    // raw ground AABB
    float plo[2], phi[2];  // padded by 0.1 * extent (BVH rebuild substeps)
    float a[4];            // plane: p.x p.y n.x n.y | HF: lb_x ub_x | voxel: lb.x lb.y dx dz
    int32_t nx, ny, reverse;
    float hf_dxw;          // HF: sample spacing (ub_x - lb_x) / (nx - 1), rounded once in float32 on the host
    float b[8];            // HF: b[0] the skip level (no node above it, below it when reversed, can
                           // touch the field);
                           // rigid: rotation (row-major, 4), velocity (2), angular velocity, friction;
                           // with a[0..1] the centre, a[2..3] the primary parameters, nx the
                           // type, ny the tied flag, hf_dxw the radius, reverse the scene's
                           // boundary-node count, lo / hi the rigid's AABB; data: float4 per node
                           // (force x, y, 1 when a force was applied, 0) written by the rigid
                           // pass before each substep, then the boundary nodes' internal ids
    const void* data;      // HF heights (float) | voxel cell extents (uint16 x4 per cell: scan end points
                           // along -x, +x, -z, +z; first = 0xFFFF when the cell is not occupied) |
                           // FEM-FEM: int2 (first entry, count) per node, then the float2 entries
                           // (nx = the scene's node count); never gated
};

// FEM-FEM contact group: the FEM bodies of one loop with the work list of their
// FEM-FEM pairs and the spatial hash over their boundary segments.  Node ids are
// scene-internal; a "slot" is a boundary node of a body in some pair, the unit
// of the per-substep contribution lists.
struct FlexGrid {
    float glb[2], gs[2];
    int32_t cn[2], tc;
    int32_t exact;  // every segment sits in every cell of its box (no cell or pair cap hit)
};

// Pending scheduled hash rebuilds a group holds (the public header's figure).
#ifndef FW_FLEX_SCHED
#define FW_FLEX_SCHED 4
#endif

// A body's arrays the contact pass reads, copied from its scene entry when the
// group is added (a scene's block does not move while it is resident), so that a
// work item reaches positions and AABBs without going through the scene table.
struct FlexBody {
    const float2* pos[2];
    const float2* vel;
    const int32_t* aabb;  // the scene's boundary-AABB sub-slots
    int32_t aabb_k;
    int32_t pad_;
};

struct FlexDev {
    int32_t n_bodies, n_segs, n_items, n_pairs, n_slots;
    int32_t gated, max_cells, max_query, max_pairs;
    int32_t n_tied;                // tied work items (0: item_tied is never read)
    int32_t s0, horizon;           // the group's scenes' warm-up step and horizon (all equal)
    const int32_t* sched;          // [FW_FLEX_SCHED] global steps before which the hash is rebuilt
                                   // (INT32_MIN: free), written by fw_flex_schedule_rebuild
    const FlexBody* bodies;        // [B]
    const int32_t* body_scene;     // [B] arena scene id
    const int32_t* body_seg_off;   // [B + 1]
    const int32_t* body_slot_off;  // [B + 1]
    char* const* body_flex;        // [B] the scene's FEM-FEM block (nullptr: body in no pair)
    const int32_t* body_nn;        // [B] node count
    const int32_t* slot_node;      // [n_slots] internal node id
    const int32_t* slot_body;      // [n_slots]
    const int32_t* seg_body;       // [S]
    const int2* seg_node;          // [S] internal node ids
    const int2* seg_slot;          // [S] slots of its nodes (-1: body in no pair)
    const int32_t* item_body;      // [W]
    const int32_t* item_node;      // [W] internal node id
    const int32_t* item_slot;      // [W]
    const int32_t* item_target;    // [W]
    const int32_t* item_pair;      // [W]
    const float* item_penalty;     // [W]
    const float* item_friction;    // [W]
    int32_t* cache;                // [W] cached segment or -1
    const int2* pair_body;         // [P]
    // tied pairs: per work item the tie flag, the edge flex_ties_kernel resolved
    // (-1: none yet) and (weight 0, weight 1, normal gap, 0)
    const int32_t* item_tied;      // [W]
    int32_t* tie_seg;              // [W]
    float4* tie_wg;                // [W]
    // spatial hash, rebuilt by flex_hash_kernel
    FlexGrid* grid;
    float4* seg_box;               // [S] segment boxes of the last build
    int4* seg_cells;               // [S] their clamped cell ranges (lx, ly, ux, uy)
    int4* body_cells;              // [B] union of the body's segment cell ranges (empty: lx > ux)
    int32_t* seg_pairs;            // [S] scratch: pairs per segment, then the serial pair base
    int32_t* cell_start;           // [max_cells]
    int32_t* cell_end;             // [max_cells]
    int2* cell_rng;                // [max_cells] (start, end) together: one load per cell in a query
    int32_t* sorted;               // [max_pairs] segment ids, ascending within a cell
    int32_t* cell_ent;             // [max_pairs] the same entries packed for the query walk:
                                   // segment << 2 | (cell is the segment's first column) << 1
                                   // | (first row); a body's segments are one id range
                                   // (body_seg_off), so the walk tests the body by range
    // per-substep scratch of flex_contact_kernel
    int2* res_i;                   // [W] (hit, segment)
    float4* res_f;                 // [W] (force x, y, w0, w1)
    int32_t* slot_start;           // [n_slots]
    int32_t* slot_cur;             // [n_slots]
    int32_t* ent_key;              // [3W] work item of each entry
    float2* ent_f;                 // [3W]
    int32_t* list_total;           // [1] entries placed by the previous substep's lists pass
    // offset of this group's per-substep scratch in the compact layout fw_step_noalloc
    // uses (the live groups' scratch one after another, flex_scratch_bytes each)
    int64_t scr_off;
};

// The per-substep scratch of one group's contact pass, laid out compactly for a
// caller-independent region: (force, weights) and (hit, segment) per work item,
// up to three (work item, force) entries per work item, a count / start and a
// cursor per slot, and the previous substep's entry total.  Offsets in bytes.
struct FlexScratchLayout {
    int64_t res_f, res_i, ent_f, ent_key, slot_start, slot_cur, list_total, bytes;
};
__host__ __device__ inline int64_t fw_align16(int64_t b) { return (b + 15) & ~(int64_t)15; }
__host__ __device__ inline FlexScratchLayout flex_scratch_layout(int64_t n_items, int64_t n_slots) {
    FlexScratchLayout L;
    int64_t o = 0;
    L.res_f = o;
    o += fw_align16(16 * n_items);
    L.res_i = o;
    o += fw_align16(8 * n_items);
    L.ent_f = o;
    o += fw_align16(24 * n_items);
    L.ent_key = o;
    o += fw_align16(12 * n_items);
    L.slot_start = o;
    o += fw_align16(4 * n_slots);
    L.slot_cur = o;
    o += fw_align16(4 * n_slots);
    L.list_total = o;
    o += 16;
    L.bytes = fw_align16(o);
    return L;
}

// The scene fields every tile reads (material, dt, gate state), kept together at
// the front of the scene entry.  Spring scenes keep the stiffness in E and the
// spring damping coefficient in nu.
struct SceneHot {
    int32_t mat_type, n_grounds;
    float E, nu, sigma_y, H;
    float dt, damping;
    int32_t gate_base;  // local step that counts as 0 for the gate's rebuild schedule
    // Boundary AABB (ordered-int encoded), triple buffered by local step and split in
    // aabb_k sub-slots (aabb_k = min(tiles, FW_AABB_MAX_SUB)): tile t of the scene
    // writes sub-slot t % aabb_k, so no more than ceil(tiles / FW_AABB_MAX_SUB) tiles
    // update one word with atomics, and a reader takes the min / max over the
    // sub-slots.  Buffer b, sub-slot j: aabb + (b * aabb_k + j) * FW_AABB_SUB_INTS.
    int32_t aabb_k;
    int32_t* aabb;
};

// Sub-slots of a scene's boundary AABB: each 64 bytes apart, so that sub-slots do
// not share a cache line.
#define FW_AABB_MAX_SUB 64
#define FW_AABB_SUB_INTS 16

struct SceneDev {
    SceneHot hot;
    // node arrays
    float2* pos[2];
    const float2* rest;
    float2* vel;
    float* mass;
    const uint32_t* flag;
    const float2* bcv;
    float2* disp;   // nullptr when displacement tracking is off
    float2* force;  // diag
    uint8_t* cflag; // diag
    float2* hpos[2];       // halo copies (host bookkeeping; the substep uses the tile's pointers)
    float2* hvel[2];       // halo velocity copies, spring scenes (host bookkeeping)
    // element arrays
    float4* jinv;
    float* area;
    float4* j2[2];
    float4* stress;  // diag (xx, yy, xy, 0)
    float4* strain;  // diag
    uint8_t* yflag;  // diag
    // scalars
    int32_t n_nodes, n_elems;
    float rho;        // density; spring scenes: the mass of every node
    int32_t s0;       // global step at which the scene runs its warm-up (local step -1 is s0 - 1)
    int32_t horizon;  // substeps after the warm-up; INT32_MAX when unbounded
    int32_t live;     // 0 once removed (host bookkeeping: the tile table holds live scenes only)
    GroundDev ground[FW_MAX_GROUNDS_DEV];
};

// One tile of one scene.  It locates every array the substep's first loads touch
// (one base pointer + 32-bit offsets, to keep the record in few scalar registers),
// so those loads depend on this record alone; the scene entry is read in parallel,
// for material, ground and diagnostics.
struct TileDev {
    char* base;              // the scene's block; array = base + (offset << 8) (blocks are 256-byte aligned)
    uint32_t o_rec;          // uint2 per element slot: .x = w0 | w1 << 16, .y = w2 (bits 16-31 zero)
    //   for a triangle, .y = the rest length's float bits for a spring;
    //   corner word w = l | r << FW_REC_ID_BITS: l the local id of the corner's node (owned
    //   nodes first, then the tile's halo nodes), r the rank of this (element, corner)
    //   in the node's incidence list.  An owned node's corner force goes to LDS entry
    //   cptr[l] + r; a halo node's (l >= n_nodes) is its owner tile's business.
    uint32_t o_hel;          // int32 per halo slot: storage index of the element (J2 state)
    uint32_t o_bcv;          // float2 BC value table of the scene (indexed by the node flag's row)
    uint32_t o_pos[2];       // float2 node positions (ping-pong)
    uint32_t o_hpos[2];      // float2 halo copies (ping-pong)
    uint32_t o_aux;          // float2 second per-node stage array: rest positions (triangles) or
                             // velocities (springs, whose force needs both ends' velocities)
    uint32_t o_haux[2];      // its halo copies: rest positions (both entries) or velocity copies
                             // (ping-pong like o_hpos)
    uint32_t o_vel;          // float2
    uint32_t o_mass;         // float
    uint32_t o_flag;         // uint32
    uint32_t o_disp;         // float2; 0 when displacement tracking is off
    uint32_t o_cptr;         // uint16 node CSR: n_nodes + 1 offsets per tile into its node-major LDS entries
    uint32_t o_push;         // int2 push list: (local owned node, halo copy index in the scene)
    uint32_t o_j2[2];        // float4 J2 state (ping-pong); 0 unless J2
    int32_t scene, s0, horizon;         // s0 / horizon as in SceneDev
    int32_t aabb_slot;                  // AABB sub-slot this tile writes (its index in the scene % aabb_k)
    int32_t aabb_mode;                  // 0: the sub-slot is this tile's alone (plain stores); 1: shared
                                        // (atomics); 2: shared, and this tile resets it for two steps on
    int32_t node0, n_nodes;             // owned nodes [node0, node0 + n_nodes)
    int32_t elem0, n_own;               // owned elements [elem0, elem0 + n_own)
    int32_t n_halo;                     // halo element slots after the owned ones
    int32_t rec_off;                    // rec[rec_off + slot]
    int32_t hnode_off, n_hnodes;        // halo copies hpos[.][hnode_off .. +n_hnodes)
    int32_t ptr_off;                    // cptr[ptr_off .. ptr_off + n_nodes]
    int32_t push_off, n_push;           // push[push_off .. +n_push)
    int32_t hel_off;                    // hel[hel_off + slot - n_own] for halo slots
};

// Slot record corner word: local node id in the low bits, incidence rank above.  A
// triangle's record keeps five more rank bits per corner in the free upper half of
// .y (corner c at bit 16 + 5c), so a triangle corner's rank reaches 1023; a spring's
// .y holds its rest length, so a spring corner's rank stops at 31.
#define FW_REC_ID_BITS 11
#define FW_REC_ID_MASK 0x7FFu
#define FW_REC_MAX_RANK 31
#define FW_REC_MAX_RANK_TRI 1023

// Compiled substep kernel shapes: (threads, element slots, LDS nodes, owned nodes per
// thread, LDS nodes of a spring tile).  Triangle tiles: 16 B per LDS node (positions,
// rest positions) + 24 B per slot (three corner forces) + 2 B per CSR offset (owned
// nodes + 1), sized so that 4 (1024 slots), 5 (768), 8 (512) or 2 (2048) workgroups
// share a CU's 160 KiB.  Spring tiles: 16 B per LDS node (positions, velocities) +
// 16 B per slot (two corner forces); a spring tile fills its slots at about 6 springs
// per owned node, so its node capacity is smaller and 6 (1024 slots), 8 (768, 512)
// or 3 (2048) of its workgroups share a CU.
#define FW_SHAPES(X)              \
    X(256, 1024, 952, 2, 448)     \
    X(256, 768, 704, 2, 384)      \
    X(512, 1024, 952, 1, 448)     \
    X(256, 512, 472, 1, 256)      \
    X(512, 2048, 1912, 2, 768)    \
    X(512, 2048, 1912, 1, 768)    \
    X(1024, 2048, 1912, 1, 768)

// HBM accesses of the kernels go through fw_ld / fw_st: they address global
// memory explicitly.  A plain access through a pointer the compiler cannot prove
// global (one read from a table or computed from a char*) is emitted as a flat
// access, which counts against the LDS counter too and so makes every later LDS
// wait also wait for the outstanding HBM traffic.
#define FW_GLOBAL __attribute__((address_space(1)))
typedef float fw_v2f __attribute__((ext_vector_type(2)));
typedef float fw_v4f __attribute__((ext_vector_type(4)));
typedef uint32_t fw_v2u __attribute__((ext_vector_type(2)));
typedef uint32_t fw_v4u __attribute__((ext_vector_type(4)));
typedef int32_t fw_v2i __attribute__((ext_vector_type(2)));
typedef int32_t fw_v4i __attribute__((ext_vector_type(4)));

__device__ inline float2 fw_ld(const float2* p) { const fw_v2f v = *(const FW_GLOBAL fw_v2f*)p; return make_float2(v.x, v.y); }
__device__ inline float4 fw_ld(const float4* p) {
    const fw_v4f v = *(const FW_GLOBAL fw_v4f*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ inline uint4 fw_ld(const uint4* p) {
    const fw_v4u v = *(const FW_GLOBAL fw_v4u*)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ inline int4 fw_ld(const int4* p) {
    const fw_v4i v = *(const FW_GLOBAL fw_v4i*)p;
    return make_int4(v.x, v.y, v.z, v.w);
}
__device__ inline uint2 fw_ld(const uint2* p) { const fw_v2u v = *(const FW_GLOBAL fw_v2u*)p; return make_uint2(v.x, v.y); }
__device__ inline int2 fw_ld(const int2* p) { const fw_v2i v = *(const FW_GLOBAL fw_v2i*)p; return make_int2(v.x, v.y); }
__device__ inline float fw_ld(const float* p) { return *(const FW_GLOBAL float*)p; }
__device__ inline uint32_t fw_ld(const uint32_t* p) { return *(const FW_GLOBAL uint32_t*)p; }
__device__ inline int32_t fw_ld(const int32_t* p) { return *(const FW_GLOBAL int32_t*)p; }
__device__ inline uint16_t fw_ld(const uint16_t* p) { return *(const FW_GLOBAL uint16_t*)p; }

__device__ inline void fw_st(float2* p, float2 a) { *(FW_GLOBAL fw_v2f*)p = fw_v2f{a.x, a.y}; }
__device__ inline void fw_st(float4* p, float4 a) { *(FW_GLOBAL fw_v4f*)p = fw_v4f{a.x, a.y, a.z, a.w}; }
__device__ inline void fw_st(float* p, float a) { *(FW_GLOBAL float*)p = a; }
__device__ inline void fw_st(int32_t* p, int32_t a) { *(FW_GLOBAL int32_t*)p = a; }
__device__ inline void fw_st(int4* p, int4 a) { *(FW_GLOBAL fw_v4i*)p = fw_v4i{a.x, a.y, a.z, a.w}; }
__device__ inline void fw_st(int2* p, int2 a) { *(FW_GLOBAL fw_v2i*)p = fw_v2i{a.x, a.y}; }
__device__ inline void fw_st(uint8_t* p, uint8_t a) { *(FW_GLOBAL uint8_t*)p = a; }

// Tile and scene tables are never written while a kernel runs (the host edits
// them between launches, and a kernel start invalidates the scalar cache), so
// the kernels read them through the constant address space: with a uniform
// address that is one batch of scalar loads instead of vector loads followed by
// v_readfirstlane.  fw_ldc copies a whole record, fw_ldc1 one field.
#define FW_CONST __attribute__((address_space(4)))
template <typename T>
__device__ inline T fw_ldc(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "record of whole dwords");
    T out;
    const FW_CONST uint32_t* s = (const FW_CONST uint32_t*)p;
    uint32_t* d = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) d[i] = s[i];
    return out;
}
template <typename T>
__device__ inline T fw_ldc1(const T* p) {
    return *(const FW_CONST T*)p;
}

template <typename P>
__device__ inline P* fw_at(char* base, uint32_t off) {
    return reinterpret_cast<P*>(base + ((size_t)off << 8));
}

// Arena control block in device memory.
struct ArenaCtl {
    int32_t clock;        // global substep counter; substep kernel uses clock + k
    int32_t n_tiles;      // tiles in the table
    int32_t n_tri_tiles;  // the first n_tri_tiles are triangle tiles, the rest spring tiles
};

// Ordered-int encoding of floats so that integer min/max follow float order.
__host__ __device__ inline int32_t fw_f2o(float f) {
    int32_t i;
    __builtin_memcpy(&i, &f, 4);
    return i >= 0 ? i : (i ^ 0x7FFFFFFF);
}
__host__ __device__ inline float fw_o2f(int32_t i) {
    int32_t j = i >= 0 ? i : (i ^ 0x7FFFFFFF);
    float f;
    __builtin_memcpy(&f, &j, 4);
    return f;
}
