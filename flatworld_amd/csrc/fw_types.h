// fw_types.h -- device-resident tables of the FEM arena (shared by host runtime and kernels).
//
// HBM layout per scene (all SoA, internal node/element order, 256-byte aligned blocks):
//   nodes    pos[2] float2 (ping-pong: substep k reads pos[(k+1)&1], writes pos[k&1]),
//            rest float2 (reference positions), vel float2, mass float, flag uint32,
//            disp float2 (optional), force float2 (diag), cflag uint8 (diag)
//   elements jinv float4 + area float (written once, for readback only: the substep
//            recomputes both from the rest positions), j2[2] float4 (eps_xx, eps_yy,
//            eps_xy, eqps; ping-pong, J2 scenes only), stress / strain float4 (diag),
//            yflag uint8 (diag)
//   tiles    rec uint4 per element slot, hnode int32 per halo node, csr_ptr uint16 per
//            owned node (+1)
//   ground   height samples float / occupancy int8
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define FW_MAX_GROUNDS_DEV 4

struct GroundDev {
    int32_t kind;
    float penalty;
    float lo[2], hi[2];    // raw ground AABB
    float plo[2], phi[2];  // padded by 0.1 * extent (BVH rebuild substeps)
    float a[4];            // plane: p.x p.y n.x n.y | HF: lb_x ub_x | voxel: lb.x lb.y dx dz
    int32_t nx, ny, reverse, pad;
    const void* data;      // HF heights (float) | voxel occupancy (int8: 0 empty, 1 occupied, 2 other)
};

struct SceneDev {
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
    // element arrays
    float4* jinv;
    float* area;
    float4* j2[2];
    float4* stress;  // diag (xx, yy, xy, 0)
    float4* strain;  // diag
    uint8_t* yflag;  // diag
    // tile data
    // rec per element slot: .x = l0 | l1 << 16, .y = l2 | d0 << 16, .z = d1 | d2 << 16, .w = element
    //   l* local node ids of the corners (owned nodes first, then the tile's halo nodes),
    //   d* LDS entry receiving that corner's force (0xFFFF: node not owned by the tile)
    const uint4* rec;
    const int32_t* hnode;     // per tile: storage ids of the halo nodes
    const uint16_t* csr_ptr;  // per tile: n_nodes + 1 offsets into the tile's node-major LDS entries
    // scalars
    int32_t n_nodes, n_elems;
    int32_t mat_type, n_grounds;
    float E, nu, rho, sigma_y, H;
    float dt, damping;
    int32_t s0;       // global step at which the scene runs its warm-up (local step -1 is s0 - 1)
    int32_t horizon;  // substeps after the warm-up; INT32_MAX when unbounded
    int32_t live;     // 0 once removed
    int32_t gate_base; // local step that counts as 0 for the gate's rebuild schedule
    int32_t aabb[3][4];  // boundary AABB (ordered-int encoded), triple buffered by local step
    GroundDev ground[FW_MAX_GROUNDS_DEV];
};

struct TileDev {
    int32_t scene;
    int32_t node0, n_nodes;    // owned nodes [node0, node0 + n_nodes)
    int32_t elem0, n_own;      // owned elements [elem0, elem0 + n_own)
    int32_t n_halo;            // halo element slots after the owned ones
    int32_t rec_off;           // rec[rec_off + slot]
    int32_t hnode_off, n_hnodes;  // hnode[hnode_off .. +n_hnodes)
    int32_t ptr_off;           // csr_ptr[ptr_off .. ptr_off + n_nodes]
    int32_t multi;             // scene has more than one tile (AABB via atomics)
    int32_t pad;
};

// Arena control block in device memory.
struct ArenaCtl {
    int32_t clock;   // global substep counter; substep kernel uses clock + k
    int32_t n_tiles; // tiles in the table
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
