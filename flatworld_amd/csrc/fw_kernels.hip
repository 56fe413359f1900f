// fw_kernels.hip -- gfx950 kernels of the explicit-FEM substep.
//
// One fused kernel per substep.  A workgroup takes one tile: a contiguous range
// of owned nodes of one scene plus every element touching them.
//   stage    every independent load of the tile goes out at once: the owned
//            nodes' positions and rest positions (coalesced) and those of the
//            tile's few halo nodes, one 16-byte record per element slot
//            (local corner ids, LDS destinations, element id), J2 state, and
//            the node data phase B needs.  Positions land in LDS.
//   phase A  element slots: rest frame (J^-1, area) recomputed from the rest
//            positions -- bit-identical to storing it, 16 B less per element --
//            F = Jac(x) J^-1, PK2 stress (Elastic / Neo-Hookean / J2 return
//            map), corner forces area * B^T S written straight to the LDS
//            entry of the owning node's incidence list.  The owner tile writes
//            the J2 state to the other buffer (ping-pong), so halo tiles never
//            see a half-updated state.  Spring tiles stage velocities where
//            triangle tiles stage rest positions and evaluate one spring per
//            slot (stiffness and damping along the spring).
//   phase B  owned nodes: ground penalty forces (plane, height field, voxel)
//            behind the AABB activation gate, then the element forces from a
//            contiguous LDS run in ascending caller element id -- the order of
//            Warp-CPU's serial atomic adds -- then BCs and symplectic Euler,
//            then the boundary AABB for the next substep's gate.
// No float atomics: every nodal sum has a fixed order, so a scene's result does
// not depend on what else is resident in the arena.  Positions are
// double-buffered because neighbouring tiles read them; velocities, mass and
// flags are touched by the owner tile only.
//
// Arithmetic follows flatworld/femspringmanager.py:59-138, 141-201, 217-418,
// flatworld/materials/materialfunctions.py:10-101 and
// flatworld/mixedcontact.py:166-302, 660-684, 1566-1643 operation by
// operation in float32; the build uses -ffp-contract=off so no FMA changes a
// rounding.  log(J) is fw_logf, a double-operation series shared with the
// CPU restatement.
#include <hip/hip_runtime.h>

#include "fw_device.h"
#include "fw_types.h"

namespace fw {

// Inverse rest Jacobian and area (flatworld/femspringmanager.py:59-68, 113-133):
// Jac = [[x1-x3, x2-x3], [y1-y3, y2-y3]], area = |det| / 2, J^-1 = adj(Jac) / det.
__device__ __forceinline__ void rest_frame(float2 X1, float2 X2, float2 X3, float4& ji, float& area) {
    const float J0 = X1.x - X3.x, J1 = X2.x - X3.x, J2 = X1.y - X3.y, J3 = X2.y - X3.y;
    const float det = J0 * J3 - J1 * J2;
    area = fabsf(det) * 0.5f;
    const float inv_det = 1.0f / det;
    ji = make_float4(J3 * inv_det, -J1 * inv_det, -J2 * inv_det, J0 * inv_det);
}

// Corner words of a slot record: local node ids, and for corners at owned nodes
// the LDS entry of the corner force, cptr[l] + rank (0xFFFF for halo nodes).
__device__ __forceinline__ void unpack_rec(uint2 r, int n_own_nodes, const uint16_t* lds_cp, uint32_t l[3],
                                           uint32_t q[3]) {
    const uint32_t w[3] = {r.x & 0xFFFFu, r.x >> 16, r.y & 0xFFFFu};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        l[c] = w[c] & FW_REC_ID_MASK;
        // branch-free: halo corners read a valid entry and discard it; the rank is five
        // bits in the corner word and five more in the upper half of .y
        const uint32_t base = lds_cp[min((int)l[c], n_own_nodes)];
        q[c] = (int)l[c] < n_own_nodes ? base + ((w[c] >> FW_REC_ID_BITS) | (((r.y >> (16 + 5 * c)) & 31u) << 5))
                                       : 0xFFFFu;
    }
}

// One element slot: F = Jac(x) J^-1, stress, corner forces area * B^T S
// (flatworld/femspringmanager.py:71-105, 240-276), stored at the LDS entries of
// the owning tile's node-major incidence lists.
__device__ __forceinline__ void slot_force(const Mat& m, float2 p1, float2 p2, float2 p3, float2 X1, float2 X2,
                                           float2 X3, float4 st, const uint32_t q[3], bool own, int e, bool diag,
                                           float4* __restrict__ j2_wr, float4* __restrict__ stress,
                                           float4* __restrict__ strain, uint8_t* __restrict__ yflag,
                                           float2* lds_f) {
    float4 ji;
    float ar;
    rest_frame(X1, X2, X3, ji, ar);
    const float J0 = p1.x - p3.x, J1 = p2.x - p3.x, J2 = p1.y - p3.y, J3 = p2.y - p3.y;
    float F[4];
    F[0] = J0 * ji.x + J1 * ji.z;
    F[1] = J0 * ji.y + J1 * ji.w;
    F[2] = J2 * ji.x + J3 * ji.z;
    F[3] = J2 * ji.y + J3 * ji.w;
    float s[3], g[3];
    bool yielded = false;
    if (m.type == 3) {
        float4 stn;
        yielded = stress_j2(m, F, st, s, g, stn);
        if (own) fw_st(j2_wr + e, stn);
    } else {
        stress_hyper(m, F, s, g);
    }
    const float d00 = ji.x, d01 = ji.y, d10 = ji.z, d11 = ji.w;
    const float d20 = -ji.x - ji.z;
    const float d21 = -ji.y - ji.w;
    const float f0 = ((d00 * F[0]) * s[0] + (d01 * F[1]) * s[1] + (d01 * F[0] + d00 * F[1]) * s[2]) * ar;
    const float f1 = ((d00 * F[2]) * s[0] + (d01 * F[3]) * s[1] + (d00 * F[3] + d01 * F[2]) * s[2]) * ar;
    const float f2 = ((d10 * F[0]) * s[0] + (d11 * F[1]) * s[1] + (d11 * F[0] + d10 * F[1]) * s[2]) * ar;
    const float f3 = ((d10 * F[2]) * s[0] + (d11 * F[3]) * s[1] + (d10 * F[3] + d11 * F[2]) * s[2]) * ar;
    const float f4 = ((d20 * F[0]) * s[0] + (d21 * F[1]) * s[1] + (d21 * F[0] + d20 * F[1]) * s[2]) * ar;
    const float f5 = ((d20 * F[2]) * s[0] + (d21 * F[3]) * s[1] + (d20 * F[3] + d21 * F[2]) * s[2]) * ar;
    if (q[0] != 0xFFFFu) lds_f[q[0]] = make_float2(f0, f1);
    if (q[1] != 0xFFFFu) lds_f[q[1]] = make_float2(f2, f3);
    if (q[2] != 0xFFFFu) lds_f[q[2]] = make_float2(f4, f5);
    if (diag && own) {
        // PK2 -> Cauchy F S F^T / max(J, 1e-12) (flatworld/femspringmanager.py:267-272)
        const float FS0 = F[0] * s[0] + F[1] * s[2];
        const float FS1 = F[0] * s[2] + F[1] * s[1];
        const float FS2 = F[2] * s[0] + F[3] * s[2];
        const float FS3 = F[2] * s[2] + F[3] * s[1];
        const float c00 = FS0 * F[0] + FS1 * F[1];
        const float c01 = FS0 * F[2] + FS1 * F[3];
        const float c11 = FS2 * F[2] + FS3 * F[3];
        const float J = F[0] * F[3] - F[1] * F[2];
        const float jd = J > 1e-12f ? J : 1e-12f;
        fw_st(stress + e, make_float4(c00 / jd, c11 / jd, c01 / jd, 0.0f));
        fw_st(strain + e, make_float4(g[0], g[1], g[2], 0.0f));
        fw_st(yflag + e, (uint8_t)(yielded ? 1 : 0));
    }
}

// sign bit of a spring incidence entry: the entry's node takes +f (node b of the spring)
constexpr uint32_t SPRING_NEG = 0x8000u;

// One spring slot (flatworld/femspringmanager.py:283-312).  Its record: .x the two
// corner words (local node id, incidence rank), .y the rest length's bits.
// xab = x_a - x_b, lnew = |xab| (Warp's dot: a zero-initialised sum), dir = xab / lnew,
// f = k (lnew / l0 - 1) dir + c (v_ab . dir) dir.  Node a receives -f and node b
// +f.  The force is stored once, at the slot; each corner's incidence entry holds the
// slot with a sign bit (a: +, b: -), and the node sums subtract the signed force, so
// a's term is -f and b's is +f, in ascending spring id.  The incidence offsets are
// read unconditionally (clamped id) and a halo corner writes the spare entry
// `spare`, so that the LDS reads of a slot issue together.
__device__ __forceinline__ void spring_slot(float k, float c, uint2 r, int n_own, const float2* lds_x,
                                            const float2* lds_v, const uint16_t* lds_cp, int sl, int spare,
                                            float2* lds_f, uint16_t* lds_inc) {
    const uint32_t w0 = r.x & 0xFFFFu, w1 = r.x >> 16;
    const int l0 = (int)(w0 & FW_REC_ID_MASK), l1 = (int)(w1 & FW_REC_ID_MASK);
    const float2 pa = lds_x[l0], pb = lds_x[l1], va = lds_v[l0], vb = lds_v[l1];
    const uint32_t c0 = lds_cp[min(l0, n_own)], c1 = lds_cp[min(l1, n_own)];
    const float l0r = __uint_as_float(r.y);
    const float xab0 = pa.x - pb.x, xab1 = pa.y - pb.y;
    const float lnew = sqrtf((0.0f + xab0 * xab0) + xab1 * xab1);
    const float d0 = xab0 / lnew, d1 = xab1 / lnew;
    const float a = k * (lnew / l0r - 1.0f);
    float f0 = a * d0, f1 = a * d1;
    const float vab0 = va.x - vb.x, vab1 = va.y - vb.y;
    const float b = c * ((0.0f + vab0 * d0) + vab1 * d1);
    f0 = f0 + b * d0;
    f1 = f1 + b * d1;
    lds_f[sl] = make_float2(f0, f1);
    lds_inc[l0 < n_own ? (int)(c0 + (w0 >> FW_REC_ID_BITS)) : spare] = (uint16_t)sl;
    lds_inc[l1 < n_own ? (int)(c1 + (w1 >> FW_REC_ID_BITS)) : spare] = (uint16_t)(sl | SPRING_NEG);
}

// ------------------------------------------------------------------ substep kernel

struct StepArgs {
    const SceneDev* scenes;
    const TileDev* tiles;
    const ArenaCtl* ctl;
    int k;           // substep index within this launch group (global step = ctl->clock + k)
    int tile_base;   // first tile of the range
    int tile_count;  // tiles in the range; < 0: every tile of this kernel's element kind (from ctl)
    float dt_override;
    float damping_override;
    uint32_t flags;
};

// Element kinds of the substep kernel: the arena keeps triangle tiles first and
// spring tiles after them, and launches one instantiation per kind present, so
// neither kind's registers pay for the other's code.
constexpr int KIND_TRI = 0, KIND_SPRING = 1;

// Node data phase B needs, loaded before the tile's barrier.
struct NodeIn {
    float2 v;
    float2 u;
    float m;
    uint32_t fl;
};

// What a tile's stage loads into registers before its data goes to LDS: the
// element slot records, the positions and rest positions (springs: velocities)
// of every LDS node (owned nodes first, then the tile's halo copies), the owned
// nodes' CSR offsets, and for J2 tiles the state of slot 0 and the element ids
// of halo slots.
template <int SPT, int LPT, int CPT>
struct Stage {
    uint2 rc[SPT];
    float2 x[LPT];
    float2 X[LPT];
    uint32_t cp[CPT];
    float4 j2s;       // J2 state of slot 0 when it is an owned slot
    int32_t he[SPT];  // J2 tiles: element of each halo slot
};

__device__ __forceinline__ bool tile_live(const TileDev& T, int G) {
    const int local = G - T.s0;
    return local >= -1 && local < T.horizon;
}

// Issue every stage load of tile T at global step G; none depends on another, so
// they are in flight together (one HBM round trip for the whole stage).  Lanes
// past the end re-read the last valid entry instead of branching around the
// load: the loads stay unconditional, which lets the compiler track each one
// precisely instead of waiting for all of them at the first control-flow join.
template <int BLOCK, int SPT, int LPT, int CPT>
__device__ __forceinline__ void stage_load(const TileDev& T, int G, Stage<SPT, LPT, CPT>& st) {
    const int tid = threadIdx.x;
    const int rd = (G - T.s0 + 1) & 1;
    const int nslots = T.n_own + T.n_halo;
    const uint2* __restrict__ rec = fw_at<const uint2>(T.base, T.o_rec) + T.rec_off;
    if (nslots > 0) {
#pragma unroll
        for (int k = 0; k < SPT; ++k) st.rc[k] = fw_ld(rec + min(tid + k * BLOCK, nslots - 1));
    } else {
#pragma unroll
        for (int k = 0; k < SPT; ++k) st.rc[k] = make_uint2(0u, 0u);
    }
    const int nn = T.n_nodes;
    const int nl = nn + T.n_hnodes;  // >= 1: a tile owns at least one node
    const float2* __restrict__ pos_rd = fw_at<const float2>(T.base, (rd ? T.o_pos[1] : T.o_pos[0])) + T.node0;
    const float2* __restrict__ aux = fw_at<const float2>(T.base, T.o_aux) + T.node0;
    const float2* __restrict__ hpos_rd = fw_at<const float2>(T.base, (rd ? T.o_hpos[1] : T.o_hpos[0])) + T.hnode_off - nn;
    const float2* __restrict__ haux_rd = fw_at<const float2>(T.base, (rd ? T.o_haux[1] : T.o_haux[0])) + T.hnode_off - nn;
#pragma unroll
    for (int k = 0; k < LPT; ++k) {
        const int j = min(tid + k * BLOCK, nl - 1);
        const bool own = j < nn;
        st.x[k] = fw_ld((own ? pos_rd : hpos_rd) + j);
        st.X[k] = fw_ld((own ? aux : haux_rd) + j);
    }
    const uint16_t* __restrict__ cptr = fw_at<const uint16_t>(T.base, T.o_cptr) + T.ptr_off;
#pragma unroll
    for (int k = 0; k < CPT; ++k) st.cp[k] = fw_ld(cptr + min(tid + k * BLOCK, nn));
    st.j2s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int k = 0; k < SPT; ++k) st.he[k] = 0;
    if (T.o_j2[0] != 0) {
        if (T.n_own > 0)
            st.j2s = fw_ld(fw_at<const float4>(T.base, rd ? T.o_j2[1] : T.o_j2[0]) + T.elem0 + min(tid, T.n_own - 1));
        if (T.n_halo > 0) {
            const int32_t* __restrict__ hel = fw_at<const int32_t>(T.base, T.o_hel) + T.hel_off;
#pragma unroll
            for (int k = 0; k < SPT; ++k) st.he[k] = fw_ld(hel + min(max(tid + k * BLOCK - T.n_own, 0), T.n_halo - 1));
        }
    }
}

// BLOCK threads; at most SLOTS element slots, NODES_TRI (triangle tiles) or
// NODES_SPR (spring tiles) LDS nodes (owned + halo) and NPT * BLOCK owned nodes per
// tile.  Triangle tiles: 16 B per node + 24 B per slot, nothing else (the AABB
// reduction reuses the rest-position array once phase A is done), so 1024 slots
// take exactly 40 KiB and four tiles fit one CU's 160 KiB.  Spring tiles: 16 B per
// node + 16 B per slot and a small reduction array; their staged velocities are
// dead after phase A, so phase B overwrites them with the owned nodes' new
// velocities, the source of the halo velocity copies.
//
// A workgroup loops over tiles (grid-stride, XCD-aware).  Each tile starts with
// one batch of independent loads (records, positions, rest positions) whose
// addresses come from the tile record alone, read with scalar loads.
template <int BLOCK, int SLOTS, int NODES_TRI, int NPT, int NODES_SPR, int KIND>
__global__ __launch_bounds__(BLOCK) void substep_kernel(StepArgs A) {
    constexpr bool spring = KIND == KIND_SPRING;
    constexpr int NODES = spring ? NODES_SPR : NODES_TRI;
    constexpr int CORNERS = spring ? 2 : 3;
    constexpr int SPT = (SLOTS + BLOCK - 1) / BLOCK;
    constexpr int LPT = (NODES + BLOCK - 1) / BLOCK;
    constexpr int CPT = NPT + 1;  // CSR offsets per thread: owned nodes + 1
    constexpr int NRED = (BLOCK / 64) * 4;
    __shared__ float2 lds_x[NODES];
    __shared__ float2 lds_X[NODES];
    // triangles: one force per (slot, corner) at the corner's incidence entry; springs: one
    // force per slot and a 16-bit (slot, sign) per incidence entry (half the LDS)
    __shared__ float2 lds_f[SLOTS * (spring ? 1 : CORNERS)];
    __shared__ uint16_t lds_inc[spring ? SLOTS * 2 + 2 : 1];  // + a spare entry for halo corners
    __shared__ uint16_t lds_cp[NPT * BLOCK + 1];
    __shared__ int32_t lds_red_spr[spring ? NRED : 1];
    static_assert(NRED * sizeof(int32_t) <= NODES_TRI * sizeof(float2), "reduction scratch");
    static_assert(NODES <= (int)FW_REC_ID_MASK + 1, "local node ids fit the record");
    static_assert(!spring || SLOTS <= (int)SPRING_NEG, "spring slot ids fit an incidence entry");
    int32_t* lds_red = spring ? lds_red_spr : reinterpret_cast<int32_t*>(lds_X);
    float2* const lds_vn = lds_X;  // spring tiles, phase B on: new velocities of the owned nodes

    const int G = fw_ldc1(&A.ctl->clock) + A.k;
    int tbase = A.tile_base, ntiles = A.tile_count;
    if (ntiles < 0) {
        const int n_tri = fw_ldc1(&A.ctl->n_tri_tiles);
        tbase = KIND == KIND_TRI ? 0 : n_tri;
        ntiles = KIND == KIND_TRI ? n_tri : fw_ldc1(&A.ctl->n_tiles) - n_tri;
    }
    const int tid = threadIdx.x;
    const bool diag = (A.flags & STEP_DIAG) != 0;

    for (int round = 0;; ++round) {
        const int ti = tile_of_block(round, ntiles);
        if (ti >= ntiles) break;
        const TileDev T = fw_ldc(A.tiles + tbase + ti);  // scalar loads; stores below cannot alias it
        if (!tile_live(T, G)) continue;  // uniform per block
        Stage<SPT, LPT, CPT> st;
        stage_load<BLOCK>(T, G, st);
        const int local = G - T.s0;
        const int wr = local & 1;
        const int nslots = T.n_own + T.n_halo;
        const int n_nodes = T.n_nodes;
        const bool j2 = !(KIND == KIND_SPRING) && T.o_j2[0] != 0;
        const float4* __restrict__ j2_rd = fw_at<const float4>(T.base, (local + 1) & 1 ? T.o_j2[1] : T.o_j2[0]);

        // ---------------- stage: registers -> LDS
#pragma unroll
        for (int k = 0; k < LPT; ++k) {
            const int j = tid + k * BLOCK;
            if (j < n_nodes + T.n_hnodes) {
                lds_x[j] = st.x[k];
                lds_X[j] = st.X[k];
            }
        }
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            const int j = tid + k * BLOCK;
            if (j <= n_nodes) lds_cp[j] = (uint16_t)st.cp[k];
        }
        // J2 state of slot 0: owned slots loaded it with the stage batch, a halo
        // slot (tiles with fewer owned elements than threads) loads it by the
        // element id the stage brought.  The other slots' states are fetched one
        // slot ahead inside phase A.
        float4 st_cur = st.j2s;  // lanes past the owned slots hold a copy; unused there
        if (j2 && tid >= T.n_own && tid < nslots) st_cur = fw_ld(j2_rd + st.he[0]);

        // scene scalars (needed from phase A on), scalar loads of single fields: a
        // whole-record copy would hold every field in SGPRs at once and spill
        const SceneDev* Sp = A.scenes + T.scene;
        const SceneHot* sh = &Sp->hot;
        const bool warm = local == -1;
        const float dt = warm ? 0.0f : (A.dt_override >= 0.0f ? A.dt_override : fw_ldc1(&sh->dt));
        const float damping = warm ? 0.0f : (A.damping_override >= 0.0f ? A.damping_override : fw_ldc1(&sh->damping));
        const int n_grounds = fw_ldc1(&sh->n_grounds);
        const bool contact = !warm && !(A.flags & STEP_NO_CONTACT) && n_grounds > 0;
        Mat m;  // spring tiles: lam = stiffness, mu = damping coefficient
        {
            const float E = fw_ldc1(&sh->E), nu = fw_ldc1(&sh->nu);
            m.type = spring ? FW_MAT_SPRING_DEV : fw_ldc1(&sh->mat_type);
            m.lam = spring ? E : E * nu / ((1.0f + nu) * (1.0f - 2.0f * nu));
            m.mu = spring ? nu : E / (2.0f * (1.0f + nu));
            m.sy = spring ? 0.0f : fw_ldc1(&sh->sigma_y);
            m.H = spring ? 0.0f : fw_ldc1(&sh->H);
        }
        // the gate's AABB: every lane loads one sub-slot of this step's buffer (lanes past
        // the sub-slot count repeat the last), reduced over the wave after phase A
        int32_t* const aabb_sub = fw_ldc1(&sh->aabb);
        const int aabb_k = fw_ldc1(&sh->aabb_k);
        int4 gate_bb = make_int4(0, 0, 0, 0);
        if (contact)
            gate_bb = fw_ld(reinterpret_cast<const int4*>(
                aabb_sub + (mod3(local) * aabb_k + min(tid & 63, aabb_k - 1)) * FW_AABB_SUB_INTS));
        float4* const st_stress = diag ? fw_ldc1(&Sp->stress) : nullptr;
        float4* const st_strain = diag ? fw_ldc1(&Sp->strain) : nullptr;
        uint8_t* const st_yflag = diag ? fw_ldc1(&Sp->yflag) : nullptr;
        lds_barrier();

        // ---------------- phase A: element slots from LDS
        // Registers are the occupancy limit here: the next slot's J2 state is
        // fetched one slot ahead, and the nodes' inputs for phase B are fetched two
        // slots before the end, so that they land during the last slots and the gate.
        constexpr int ND_AT = SPT >= 2 ? SPT - 2 : 0;
        NodeIn nd[NPT];
        // BC value rows, fetched in the last slot iteration once the flags (which hold the
        // row) have landed, so that their round trip overlaps the last slot and the gate
        const float2* __restrict__ bcv = fw_at<const float2>(T.base, T.o_bcv);
        float2 bv[NPT];
#pragma unroll
        for (int kn = 0; kn < NPT; ++kn) bv[kn] = make_float2(0.0f, 0.0f);
#pragma unroll
        for (int k = 0; k < SPT; ++k) {
            const int sl = tid + k * BLOCK;
            float4 st_next = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (k + 1 < SPT && j2 && sl + BLOCK < nslots)
                st_next = fw_ld(j2_rd + (sl + BLOCK < T.n_own ? T.elem0 + sl + BLOCK : st.he[k + 1 < SPT ? k + 1 : k]));
            if (k == ND_AT) {
#pragma unroll
                for (int kn = 0; kn < NPT; ++kn) {
                    const int i = tid + kn * BLOCK;
                    if (i < n_nodes) {
                        const int n = T.node0 + i;
                        if constexpr (spring)
                            nd[kn].v = lds_X[i];  // the stage brought it: spring tiles stage velocities
                        else
                            nd[kn].v = fw_ld(fw_at<const float2>(T.base, T.o_vel) + n);
                        nd[kn].u = make_float2(0.0f, 0.0f);
                        if (T.o_disp) nd[kn].u = fw_ld(fw_at<const float2>(T.base, T.o_disp) + n);
                        nd[kn].m = fw_ld(fw_at<const float>(T.base, T.o_mass) + n);
                        nd[kn].fl = fw_ld(fw_at<const uint32_t>(T.base, T.o_flag) + n);
                    }
                }
            }
            if (SPT >= 2 && k == SPT - 1) {
#pragma unroll
                for (int kn = 0; kn < NPT; ++kn)
                    if (tid + kn * BLOCK < n_nodes) bv[kn] = fw_ld(bcv + (nd[kn].fl >> NODE_BCV_SHIFT));
            }
            if (sl < nslots) {
                if constexpr (spring) {
                    spring_slot(m.lam, m.mu, st.rc[k], n_nodes, lds_x, lds_X, lds_cp, sl, SLOTS * 2, lds_f, lds_inc);
                } else {
                    uint32_t l[3], q[3];
                    unpack_rec(st.rc[k], n_nodes, lds_cp, l, q);
                    slot_force(m, lds_x[l[0]], lds_x[l[1]], lds_x[l[2]], lds_X[l[0]], lds_X[l[1]], lds_X[l[2]], st_cur, q,
                               sl < T.n_own, T.elem0 + sl, diag, fw_at<float4>(T.base, (wr ? T.o_j2[1] : T.o_j2[0])),
                               st_stress, st_strain, st_yflag, lds_f);
                }
            }
            st_cur = st_next;
        }

        // ---------------- activation gate (flatworld/explicitloop.py:477-487, bvh.py:39-46, 321-342, 393-416)
        uint32_t active = 0;
        if (contact) {
            int32_t g4[4] = {gate_bb.x, gate_bb.y, gate_bb.z, gate_bb.w};
            wave_minmax(g4);
            float lbx = fw_o2f(g4[0]), lby = fw_o2f(g4[1]);
            float ubx = fw_o2f(g4[2]), uby = fw_o2f(g4[3]);
            const bool pad = ((local - fw_ldc1(&sh->gate_base)) % 100) == 0;
            if (pad) {
                const float bx = 0.1f * (ubx - lbx), by = 0.1f * (uby - lby);
                lbx = lbx - bx;
                ubx = ubx + bx;
                lby = lby - by;
                uby = uby + by;
            }
            for (int gi = 0; gi < n_grounds; ++gi) {
                const GroundDev gd = fw_ldc(&Sp->ground[gi]);
                const float* lo = pad ? gd.plo : gd.lo;
                const float* hi = pad ? gd.phi : gd.hi;
                // the FEM-FEM record is gated per pair by the contact pass itself, a mesh
                // rigid record per work item by the rigid pass
                if (gd.kind == FW_GROUND_FLEX_DEV || gd.kind == FW_GROUND_RIGID_MESH_DEV ||
                    (lbx <= hi[0] && ubx >= lo[0] && lby <= hi[1] && uby >= lo[1]))
                    active |= 1u << gi;
            }
        }
        lds_barrier();  // lds_f complete; lds_X free from here on

        // ---------------- phase B: owned nodes
        // first push entry per thread, loaded now so that it lands during phase B
        int2 pe = make_int2(0, 0);
        if (tid < T.n_push) pe = fw_ld(fw_at<const int2>(T.base, T.o_push) + T.push_off + tid);
        if (SPT < 2) {
#pragma unroll
            for (int k = 0; k < NPT; ++k)
                if (tid + k * BLOCK < n_nodes) bv[k] = fw_ld(bcv + (nd[k].fl >> NODE_BCV_SHIFT));
        }
        float2* const st_force = diag ? fw_ldc1(&Sp->force) : nullptr;
        uint8_t* const st_cflag = diag ? fw_ldc1(&Sp->cflag) : nullptr;
        int32_t bb[4] = {fw_f2o(1e30f), fw_f2o(1e30f), fw_f2o(-1e30f), fw_f2o(-1e30f)};
        float2* __restrict__ pos_wr = fw_at<float2>(T.base, (wr ? T.o_pos[1] : T.o_pos[0]));
        float2* __restrict__ vel = fw_at<float2>(T.base, T.o_vel);
        float2* __restrict__ disp = T.o_disp ? fw_at<float2>(T.base, T.o_disp) : nullptr;
#pragma unroll
        for (int k = 0; k < NPT; ++k) {
            const int i = tid + k * BLOCK;
            if (i >= n_nodes) continue;
            const int n = T.node0 + i;
            const float2 p = lds_x[i];
            const float2 v = nd[k].v;
            const uint32_t fl = nd[k].fl;
            const float mass = nd[k].m;
            float fx = 0.0f, fy = 0.0f;
            bool hit = false;
            if (active && (fl & NODE_BOUNDARY)) {
                for (int gi = 0; gi < n_grounds; ++gi) {
                    if (!(active & (1u << gi))) continue;
                    const GroundDev gd = fw_ldc(&Sp->ground[gi]);
                    if (gd.kind == FW_GROUND_FLEX_DEV) {
                        // FEM-FEM contributions of this substep, in ascending work item
                        const int2* rng = static_cast<const int2*>(gd.data);
                        const float2* ent = reinterpret_cast<const float2*>(rng + gd.nx);
                        const int2 r = fw_ld(rng + n);
                        for (int j = 0; j < r.y; ++j) {
                            const float2 c = fw_ld(ent + r.x + j);
                            fx = fx + c.x;
                            fy = fy + c.y;
                        }
                    } else {
                        if (gd.kind >= FW_GROUND_RIGID_DEV) {
                            // the rigid pass evaluated this node's force against the primitive
                            const float4 rf = fw_ld(static_cast<const float4*>(gd.data) + n);
                            if (rf.z != 0.0f) {
                                fx = fx + rf.x;
                                fy = fy + rf.y;
                                hit = true;
                            }
                        } else {
                            hit |= ground_force(gd, p.x, p.y, v.x, v.y, dt, fx, fy);
                        }
                    }
                }
            }
            // element forces in ascending element id: a contiguous run of LDS entries
            // (two entries per iteration, so that their LDS reads overlap; same order)
            int j = lds_cp[i];
            const int j_end = lds_cp[i + 1];
            if constexpr (spring) {
                // the entry's sign bit moved to bit 31 flips the force's sign bits: the
                // float negation, exactly
                for (; j + 1 < j_end; j += 2) {
                    const uint32_t e0 = lds_inc[j], e1 = lds_inc[j + 1];
                    const float2 f0 = lds_f[e0 & ~SPRING_NEG], f1 = lds_f[e1 & ~SPRING_NEG];
                    const uint32_t s0 = (e0 & SPRING_NEG) << 16, s1 = (e1 & SPRING_NEG) << 16;
                    fx = fx + (-__uint_as_float(__float_as_uint(f0.x) ^ s0));
                    fy = fy + (-__uint_as_float(__float_as_uint(f0.y) ^ s0));
                    fx = fx + (-__uint_as_float(__float_as_uint(f1.x) ^ s1));
                    fy = fy + (-__uint_as_float(__float_as_uint(f1.y) ^ s1));
                }
                if (j < j_end) {
                    const uint32_t e = lds_inc[j];
                    const float2 f = lds_f[e & ~SPRING_NEG];
                    const uint32_t sg = (e & SPRING_NEG) << 16;
                    fx = fx + (-__uint_as_float(__float_as_uint(f.x) ^ sg));
                    fy = fy + (-__uint_as_float(__float_as_uint(f.y) ^ sg));
                }
            } else {
                for (; j + 1 < j_end; j += 2) {
                    const float2 f0 = lds_f[j], f1 = lds_f[j + 1];
                    fx = fx + (-f0.x);
                    fy = fy + (-f0.y);
                    fx = fx + (-f1.x);
                    fy = fy + (-f1.y);
                }
                if (j < j_end) {
                    const float2 f = lds_f[j];
                    fx = fx + (-f.x);
                    fy = fy + (-f.y);
                }
            }
            // symplectic Euler with BC precedence (flatworld/femspringmanager.py:315-358)
            if (fl & BC_U) {
                fx = 0.0f;
                fy = 0.0f;
            } else if (fl & BC_A) {
                fx = mass * bv[k].x;
                fy = mass * bv[k].y;
            } else if (fl & BC_G) {
                fx = fx + mass * bv[k].x;
                fy = fy + mass * bv[k].y;
            } else if (fl & BC_F) {
                fx = fx + bv[k].x;
                fy = fy + bv[k].y;
            }
            const float rm = 1.0f / mass;
            const float ax = 0.0f + (fx - damping * v.x) * rm;
            const float ay = 0.0f + (fy - damping * v.y) * rm;
            float vx = v.x + ax * dt;
            float vy = v.y + ay * dt;
            if (fl & BC_V) {
                vx = bv[k].x;
                vy = bv[k].y;
            } else if (fl & BC_U) {
                vx = 0.0f;
                vy = 0.0f;
            }
            fw_st(vel + n, make_float2(vx, vy));
            if (spring) lds_vn[i] = make_float2(vx, vy);  // source of the halo velocity copies
            const float dux = vx * dt, duy = vy * dt;
            if (disp) fw_st(disp + n, make_float2(nd[k].u.x + dux, nd[k].u.y + duy));
            const float nx_ = p.x + dux, ny_ = p.y + duy;
            fw_st(pos_wr + n, make_float2(nx_, ny_));
            lds_x[i] = make_float2(nx_, ny_);  // source of the halo copies pushed below
            if (diag) {
                fw_st(st_force + n, make_float2(fx, fy));
                fw_st(st_cflag + n, (uint8_t)(hit ? 1 : 0));
            }
            if (fl & NODE_BOUNDARY) {
                const int32_t ox = fw_f2o(nx_), oy = fw_f2o(ny_);
                bb[0] = min(bb[0], ox);
                bb[1] = min(bb[1], oy);
                bb[2] = max(bb[2], ox);
                bb[3] = max(bb[3], oy);
            }
        }

        // ---------------- boundary AABB at x_{n+1} (flatworld/femspringmanager.py:361-408)
        block_minmax<BLOCK>(bb, lds_red);  // its barrier also publishes the lds_x updates
        if (tid == 0) {
            // this tile's AABB sub-slot of the next step's buffer; a shared sub-slot takes
            // atomics, and its resetting tile empties it in the buffer of two steps on (read
            // at step local + 2, written at local + 1, so never touched during this step)
            int32_t* aw = aabb_sub + (mod3(local + 1) * aabb_k + T.aabb_slot) * FW_AABB_SUB_INTS;
            if (T.aabb_mode == 0) {
                fw_st(reinterpret_cast<int4*>(aw), make_int4(bb[0], bb[1], bb[2], bb[3]));
            } else {
                atomicMin(aw + 0, bb[0]);
                atomicMin(aw + 1, bb[1]);
                atomicMax(aw + 2, bb[2]);
                atomicMax(aw + 3, bb[3]);
                if (T.aabb_mode == 2) {
                    int32_t* az = aabb_sub + (mod3(local + 2) * aabb_k + T.aabb_slot) * FW_AABB_SUB_INTS;
                    fw_st(reinterpret_cast<int4*>(az),
                          make_int4(fw_f2o(1e30f), fw_f2o(1e30f), fw_f2o(-1e30f), fw_f2o(-1e30f)));
                }
            }
        }

        // ---------------- halo copies of this tile's nodes for the neighbouring tiles
        {
            float2* __restrict__ hpos_wr = fw_at<float2>(T.base, (wr ? T.o_hpos[1] : T.o_hpos[0]));
            float2* __restrict__ hvel_wr = fw_at<float2>(T.base, (wr ? T.o_haux[1] : T.o_haux[0]));
            if (tid < T.n_push) {
                fw_st(hpos_wr + pe.y, lds_x[pe.x]);
                if (spring) fw_st(hvel_wr + pe.y, lds_vn[pe.x]);
            }
            for (int q = tid + BLOCK; q < T.n_push; q += BLOCK) {
                const int2 e = fw_ld(fw_at<const int2>(T.base, T.o_push) + T.push_off + q);
                fw_st(hpos_wr + e.y, lds_x[e.x]);
                if (spring) fw_st(hvel_wr + e.y, lds_vn[e.x]);
            }
        }
        lds_barrier();  // LDS reuse by the next tile of this block
    }
}

// ------------------------------------------------------------------ rigid pass

// FEM-rigid contact forces of one substep (_batched_rigid_prim_kernel,
// flatworld/mixedcontact.py:1025-1117), one workgroup per (scene, rigid record), ahead
// of the substep launch: every boundary node's force against the primitive at x_n,
// v_n, into the record's per-node slots, which the substep adds at the record's place
// among the node's ground forces when the pair gate is active.  Kept out of the
// substep kernel so that its registers do not pay for the primitive queries.
// ---------------------------------------------------------------- mesh rigids
// The FEM side of _batched_rigid_mesh_kernel_sh, its workset filter and the mesh tie
// kernels (flatworld/mixedcontact.py:519-588, 1119-1565): one lane per work item.

__device__ __forceinline__ float mdot(float ax, float ay, float bx, float by) { return (0.0f + ax * bx) + ay * by; }
__device__ __forceinline__ float mlen(float x, float y) { return sqrtf(mdot(x, y, x, y)); }
__device__ __forceinline__ float mmin(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float mmax(float a, float b) { return a > b ? a : b; }

// _point_outside_aabb (mixedcontact.py:113-119)
__device__ __forceinline__ bool m_outside(float px, float py, float l0, float l1, float u0, float u1) {
    return mmin(px - l0, py - l1) < 0.0f || mmin(u0 - px, u1 - py) < 0.0f;
}

// _point_aabb_distance_sq (mixedcontact.py:122-132)
__device__ __forceinline__ float m_dist_sq(float px, float py, float l0, float l1, float u0, float u1) {
    float d2 = 0.0f, v = 0.0f;
    if (px < l0) v = l0 - px;
    else if (px > u0) v = px - u0;
    d2 = d2 + v * v;
    v = 0.0f;
    if (py < l1) v = l1 - py;
    else if (py > u1) v = py - u1;
    return d2 + v * v;
}

// detectPointToMeshBoundary over pointToEdgeContact (flatworld/contact_detection.py:15-90):
// the penetration when p projects inside the edge within `limit`, else 1e9; weights either way.
__device__ __forceinline__ float m_detect(float px, float py, float2 a, float2 b, float limit, float& nx, float& ny,
                                          float& w0, float& w1) {
    const float ex = b.x - a.x, ey = b.y - a.y;
    const float len = mlen(ex, ey);
    float pen = 1e9f, n0 = 0.0f, n1 = 0.0f;
    bool inside = false;
    w0 = w1 = 0.0f;
    if (len > 1e-9f) {
        const float dx = ex / len, dy = ey / len;
        n0 = dy;
        n1 = -dx;
        const float tx = px - a.x, ty = py - a.y;
        float t = mdot(tx, ty, dx, dy);
        t = t / len;
        if (0.0f <= t && t <= 1.0f) {
            inside = true;
            pen = mdot(tx, ty, n0, n1);
            w0 = 1.0f - t;
            w1 = t;
        }
    }
    if (fabsf(pen) < limit && inside) {
        nx = n0;
        ny = n1;
        return pen;
    }
    nx = ny = 0.0f;
    return 1e9f;
}

struct MeshView {
    const float2* coords;
    const int2* conn;
    const float2 *elb, *eub;
    const float* margin;
    const int2* reg;
    const int2* cell;
    const int32_t* sorted;
    int2* cache;
    int32_t* skip;
    int2* fric_i;
    float4* fric_f;
    float* fric_pen;
    float4* tie;
};

struct MeshBest {
    float pen, nx, ny, w0, w1;
    int elem, did;
    bool found, has_cand;
};

__device__ __forceinline__ void m_take(MeshBest& b, float pen, float nx, float ny, float w0, float w1, int e, int did) {
    b.pen = pen;
    b.nx = nx;
    b.ny = ny;
    b.w0 = w0;
    b.w1 = w1;
    b.elem = e;
    b.did = did;
    b.found = true;
}

// _process_rigid_sh_cell (mixedcontact.py:382-467): the edges of rigid `target` in
// cell cid; every improvement is written to the item's cache.
__device__ void m_cell(const MeshDev& H, const MeshView& V, int cid, float px, float py, int target, int item,
                       MeshBest& b) {
    if (!(cid >= 0 && cid < H.tc)) return;
    const float best_abs = fabsf(b.pen);
    if (b.found && best_abs < 1e8f) {
        const int nxc = H.cn[0], ix = cid % nxc, iy = cid / nxc;
        const float l0 = H.glb[0] + (float)ix * H.gs[0], l1 = H.glb[1] + (float)iy * H.gs[1];
        const float u0 = l0 + H.gs[0], u1 = l1 + H.gs[1];
        if (m_dist_sq(px, py, l0, l1, u0, u1) > best_abs * best_abs) return;
    }
    const int2 r = fw_ld(V.cell + cid);
    for (int q = r.x; q < r.y; ++q) {
        const int ei = fw_ld(V.sorted + q);
        if (ei < 0) continue;
        const int2 re = fw_ld(V.reg + ei);
        if (re.x != target) continue;
        b.has_cand = true;
        const int e = re.y;
        const int2 cn = fw_ld(V.conn + e);
        const float2 a = fw_ld(V.coords + cn.x), c = fw_ld(V.coords + cn.y);
        float l0 = 1e30f, l1 = 1e30f, u0 = -1e30f, u1 = -1e30f;
        l0 = mmin(l0, a.x);
        l1 = mmin(l1, a.y);
        u0 = mmax(u0, a.x);
        u1 = mmax(u1, a.y);
        l0 = mmin(l0, c.x);
        l1 = mmin(l1, c.y);
        u0 = mmax(u0, c.x);
        u1 = mmax(u1, c.y);
        const float margin = mmax(1e-4f, fw_ld(V.margin + e));
        l0 = l0 - margin;
        l1 = l1 - margin;
        u0 = u0 + margin;
        u1 = u1 + margin;
        const float best_abs2 = fabsf(b.pen);
        if (b.found && best_abs2 < 1e8f && m_dist_sq(px, py, l0, l1, u0, u1) > best_abs2 * best_abs2) continue;
        if (!m_outside(px, py, l0, l1, u0, u1)) {
            float nx, ny, w0, w1;
            const float pen = m_detect(px, py, a, c, margin, nx, ny, w0, w1);
            if (fabsf(pen) < fabsf(b.pen)) {
                m_take(b, pen, nx, ny, w0, w1, e, re.x);
                fw_st(V.cache + item, make_int2(e, re.x));
            }
        }
    }
}

// query cells of (px, py) widened by qb, clamped (mixedcontact.py:1226-1243)
__device__ __forceinline__ int4 m_cells(const MeshDev& H, float px, float py, float qb) {
    const float d0 = (px - qb) - H.glb[0], d1 = (py - qb) - H.glb[1];
    const float e0 = (px + qb) - H.glb[0], e1 = (py + qb) - H.glb[1];
    const int cx = H.cn[0] - 1, cy = H.cn[1] - 1;
    int lx = (int)floorf(d0 / H.gs[0]), ly = (int)floorf(d1 / H.gs[1]);
    int ux = (int)floorf(e0 / H.gs[0]), uy = (int)floorf(e1 / H.gs[1]);
    lx = min(max(lx, 0), cx);
    ly = min(max(ly, 0), cy);
    ux = min(max(ux, 0), cx);
    uy = min(max(uy, 0), cy);
    return make_int4(lx, ly, ux, uy);
}

// _clear_rigid_mesh_friction_state (mixedcontact.py:347-361)
__device__ __forceinline__ void m_clear(const MeshView& V, int k) {
    fw_st(V.fric_i + k, make_int2(-1, 0));
    fw_st(V.fric_f + k, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    fw_st(V.fric_pen + k, 0.0f);
}

// Force of a mesh record on work item k at p; false when it applies none.
__device__ bool mesh_item(const GroundDev& g, const MeshDev& H, const MeshView& V, int k, float px, float py, int local,
                          bool active, float& fx, float& fy) {
    const float penalty = g.penalty;
    fx = fy = 0.0f;
    if (g.ny == 1) {
        // _initialize_fem_rigid_mesh_ties_once_kernel at the first contact substep, then
        // _apply_fem_rigid_mesh_ties_fixed_kernel every substep, whatever the gate
        float4 tie = fw_ld(V.tie + k);
        if (local == 0 && tie.x == 0.0f && H.tc > 0) {
            const int4 r = m_cells(H, px, py, mmax(1e-4f, mmax(H.gs[0], H.gs[1])));
            MeshBest b{1e9f, 0.0f, 0.0f, 0.0f, 0.0f, -1, -1, false, false};
            for (int ix = r.x; ix <= r.z; ++ix)
                for (int iy = r.y; iy <= r.w; ++iy) m_cell(H, V, ix + iy * H.cn[0], px, py, H.did, k, b);
            if (b.found) {
                fw_st(V.cache + k, make_int2(b.elem, b.did));
                tie = make_float4(1.0f, b.w0, b.w1, 0.0f);
                fw_st(V.tie + k, tie);
            }
        }
        if (tie.x == 0.0f) return false;
        const int e = fw_ld(V.cache + k).x;
        const int2 cn = fw_ld(V.conn + e);
        const float2 a = fw_ld(V.coords + cn.x), c = fw_ld(V.coords + cn.y);
        // _eval_fixed_tie_point: (0, 0) + a w0 + c w1
        const float cp0 = (0.0f + a.x * tie.y) + c.x * tie.z;
        const float cp1 = (0.0f + a.y * tie.y) + c.y * tie.z;
        fx = penalty * (cp0 - px);
        fy = penalty * (cp1 - py);
        return true;
    }
    if (local == H.reset_at) fw_st(V.skip + k, -1);  // flatworld/explicitloop.py:543-553
    const int skip = fw_ld(V.skip + k);
    // _build_active_rigid_mesh_workset_kernel (mixedcontact.py:519-588)
    if (!active || skip == 1 || m_outside(px, py, g.lo[0] - 1e-4f, g.lo[1] - 1e-4f, g.hi[0] + 1e-4f, g.hi[1] + 1e-4f)) {
        m_clear(V, k);
        return false;
    }
    const int target = H.did;
    MeshBest b{1e9f, 0.0f, 0.0f, 0.0f, 0.0f, -1, 0, false, false};
    const int2 cached = fw_ld(V.cache + k);
    if (cached.x >= 0 && cached.y == target) {
        const float margin = mmax(1e-4f, fw_ld(V.margin + cached.x));
        const float2 lb = fw_ld(V.elb + cached.x), ub = fw_ld(V.eub + cached.x);
        if (!m_outside(px, py, lb.x - margin, lb.y - margin, ub.x + margin, ub.y + margin)) {
            const int2 cn = fw_ld(V.conn + cached.x);
            float nx, ny, w0, w1;
            const float pen = m_detect(px, py, fw_ld(V.coords + cn.x), fw_ld(V.coords + cn.y), margin, nx, ny, w0, w1);
            if (fabsf(pen) < fabsf(b.pen)) m_take(b, pen, nx, ny, w0, w1, cached.x, cached.y);
        }
    }
    const bool run_full = !(H.subcycle > 1 && local % H.subcycle != 0);
    if (!b.found && (skip == 1 || !run_full)) {
        m_clear(V, k);
        return false;
    }
    if (!b.found) {
        if (H.tc > 0) {
            const int4 r = m_cells(H, px, py, mmax(1e-4f, mmax(H.gs[0], H.gs[1])));
            bool local_first = false;
            int4 h = make_int4(0, 0, 0, 0);
            if (cached.x >= 0 && cached.y == target) {
                // _point_to_cell_coord of the cached edge's box centre (mixedcontact.py:322-344)
                local_first = true;
                const float2 lb = fw_ld(V.elb + cached.x), ub = fw_ld(V.eub + cached.x);
                const float c0 = 0.5f * (lb.x + ub.x), c1 = 0.5f * (lb.y + ub.y);
                int rx = (int)floorf((c0 - H.glb[0]) / H.gs[0]), ry = (int)floorf((c1 - H.glb[1]) / H.gs[1]);
                rx = min(max(rx, 0), H.cn[0] - 1);
                ry = min(max(ry, 0), H.cn[1] - 1);
                h = make_int4(max(rx, r.x), max(ry, r.y), min(rx, r.z), min(ry, r.w));
                for (int ix = h.x; ix <= h.z; ++ix)
                    for (int iy = h.y; iy <= h.w; ++iy) m_cell(H, V, ix + iy * H.cn[0], px, py, target, k, b);
            }
            for (int ix = r.x; ix <= r.z; ++ix)
                for (int iy = r.y; iy <= r.w; ++iy) {
                    if (local_first && ix >= h.x && ix <= h.z && iy >= h.y && iy <= h.w) continue;
                    m_cell(H, V, ix + iy * H.cn[0], px, py, target, k, b);
                }
        }
        if (!b.has_cand) fw_st(V.skip + k, 1);
    }
    if (!(b.found && b.pen < 0.0f)) {
        m_clear(V, k);
        return false;
    }
    const float nf0 = ((-b.nx) * penalty) * b.pen, nf1 = ((-b.ny) * penalty) * b.pen;
    float t0 = nf0, t1 = nf1;
    const float mu = g.b[7];
    if (mu > 1e-9f) {
        const float Fy = mu * mlen(nf0, nf1);
        const float kt = penalty * 1.0f;
        float pf0 = 0.0f, pf1 = 0.0f, s0 = 0.0f, s1 = 0.0f;
        const int2 fi = fw_ld(V.fric_i + k);
        if (fi.y == 1) {
            const float4 ff = fw_ld(V.fric_f + k);
            const int2 cc = fw_ld(V.conn + b.elem), pc = fw_ld(V.conn + fi.x);
            const float2 ca = fw_ld(V.coords + cc.x), cb = fw_ld(V.coords + cc.y);
            const float2 pa = fw_ld(V.coords + pc.x), pb = fw_ld(V.coords + pc.y);
            s0 = (ca.x * b.w0 + cb.x * b.w1) - (pa.x * ff.x + pb.x * ff.y);
            s1 = (ca.y * b.w0 + cb.y * b.w1) - (pa.y * ff.x + pb.y * ff.y);
            const float sd = mdot(s0, s1, b.nx, b.ny);
            s0 = s0 - sd * b.nx;
            s1 = s1 - sd * b.ny;
            pf0 = ff.z;
            pf1 = ff.w;
            const float pd = mdot(pf0, pf1, b.nx, b.ny);
            pf0 = pf0 - pd * b.nx;
            pf1 = pf1 - pd * b.ny;
        }
        const float ft0 = pf0 + kt * s0, ft1 = pf1 + kt * s1;
        const float fn = mlen(ft0, ft1);
        float fr0 = ft0, fr1 = ft1;
        if (!(fn <= Fy)) {
            fr0 = (ft0 / fn) * Fy;
            fr1 = (ft1 / fn) * Fy;
        }
        fw_st(V.fric_i + k, make_int2(b.elem, 1));
        fw_st(V.fric_f + k, make_float4(b.w0, b.w1, fr0, fr1));
        fw_st(V.fric_pen + k, b.pen);
        t0 = t0 + fr0;
        t1 = t1 + fr1;
    } else {
        m_clear(V, k);
    }
    fx = t0;
    fy = t1;
    return true;
}

// The pair gate of a record at `local` (the substep kernel's, flatworld/explicitloop.py:477-487,
// bvh.py:39-46): the body's boundary AABB of the previous substep, padded with the
// ground's on BVH rebuild steps, against the record's AABB.
__device__ bool record_gate(const SceneDev* S, const GroundDev& g, int local) {
    const int32_t* sub = S->hot.aabb;
    const int kk = S->hot.aabb_k;
    const int32_t* e = sub + mod3(local) * kk * FW_AABB_SUB_INTS;
    int32_t r[4] = {fw_ld(e + 0), fw_ld(e + 1), fw_ld(e + 2), fw_ld(e + 3)};
    for (int j = 1; j < kk; ++j) {
        const int32_t* f = e + j * FW_AABB_SUB_INTS;
        r[0] = min(r[0], fw_ld(f + 0));
        r[1] = min(r[1], fw_ld(f + 1));
        r[2] = max(r[2], fw_ld(f + 2));
        r[3] = max(r[3], fw_ld(f + 3));
    }
    float lbx = fw_o2f(r[0]), lby = fw_o2f(r[1]), ubx = fw_o2f(r[2]), uby = fw_o2f(r[3]);
    const bool pad = ((local - S->hot.gate_base) % 100) == 0;
    if (pad) {
        const float bx = 0.1f * (ubx - lbx), by = 0.1f * (uby - lby);
        lbx = lbx - bx;
        ubx = ubx + bx;
        lby = lby - by;
        uby = uby + by;
    }
    const float* lo = pad ? g.plo : g.lo;
    const float* hi = pad ? g.phi : g.hi;
    return lbx <= hi[0] && ubx >= lo[0] && lby <= hi[1] && uby >= lo[1];
}

__global__ __launch_bounds__(256) void rigid_pass_kernel(const SceneDev* scenes, const int2* rows, const ArenaCtl* ctl,
                                                         int k, uint32_t flags) {
    const int2 row = rows[blockIdx.x];
    const SceneDev* S = scenes + row.x;
    const int local = fw_ldc1(&ctl->clock) + k - S->s0;
    if (local < 0 || local >= S->horizon || (flags & STEP_NO_CONTACT)) return;  // warm-up: no contact
    const GroundDev g = S->ground[row.y];
    const float2* __restrict__ pos = S->pos[(local + 1) & 1];
    const float2* __restrict__ vel = S->vel;
    float4* __restrict__ out = static_cast<float4*>(const_cast<void*>(g.data));
    const int32_t* __restrict__ bn = reinterpret_cast<const int32_t*>(out + S->n_nodes);
    if (g.kind == FW_GROUND_RIGID_MESH_DEV) {
        char* blk = static_cast<char*>(const_cast<void*>(g.data));
        const int64_t hoff = ((int64_t)20 * S->n_nodes + 15) / 16 * 16;
        const MeshDev H = *reinterpret_cast<const MeshDev*>(blk + hoff);
        MeshView V;
        V.coords = reinterpret_cast<const float2*>(blk + H.o_coords);
        V.conn = reinterpret_cast<const int2*>(blk + H.o_conn);
        V.elb = reinterpret_cast<const float2*>(blk + H.o_elb);
        V.eub = reinterpret_cast<const float2*>(blk + H.o_eub);
        V.margin = reinterpret_cast<const float*>(blk + H.o_margin);
        V.reg = reinterpret_cast<const int2*>(blk + H.o_reg);
        V.cell = reinterpret_cast<const int2*>(blk + H.o_cell);
        V.sorted = reinterpret_cast<const int32_t*>(blk + H.o_sorted);
        V.cache = reinterpret_cast<int2*>(blk + H.o_cache);
        V.skip = reinterpret_cast<int32_t*>(blk + H.o_skip);
        V.fric_i = reinterpret_cast<int2*>(blk + H.o_fric_i);
        V.fric_f = reinterpret_cast<float4*>(blk + H.o_fric_f);
        V.fric_pen = reinterpret_cast<float*>(blk + H.o_fric_pen);
        V.tie = reinterpret_cast<float4*>(blk + H.o_tie);
        const bool active = record_gate(S, g, local);
        for (int j = threadIdx.x; j < H.n_items; j += blockDim.x) {
            const int n = fw_ld(bn + j);
            const float2 p = fw_ld(pos + n);
            float fx = 0.0f, fy = 0.0f;
            const bool hit = mesh_item(g, H, V, j, p.x, p.y, local, active, fx, fy);
            fw_st(out + n, make_float4(fx, fy, hit ? 1.0f : 0.0f, 0.0f));
        }
        return;
    }
    for (int j = threadIdx.x; j < g.reverse; j += blockDim.x) {
        const int n = fw_ld(bn + j);
        const float2 p = fw_ld(pos + n), v = fw_ld(vel + n);
        float fx = 0.0f, fy = 0.0f;
        const bool hit = rigid_force(g, p.x, p.y, v.x, v.y, fx, fy);
        fw_st(out + n, make_float4(fx, fy, hit ? 1.0f : 0.0f, 0.0f));
    }
}

// ------------------------------------------------------------------ prepare kernel

// flatworld/femspringmanager.py:113-138: area = |det Jac| / 2, J^-1 = inverse(Jac),
// lumped mass = sum over incident elements (ascending id) of area * rho / 3.  A
// spring scene's nodes all take the domain mass (femspringmanager.py:852-853).
// One workgroup per tile; MAXN = most owned nodes a tile of this shape holds.
template <int BLOCK, int SLOTS, int NODES, int MAXN>
__global__ __launch_bounds__(BLOCK) void prepare_kernel(const SceneDev* scenes, const TileDev* tiles, int tile_base,
                                                        int tile_count) {
    __shared__ float2 lds_X[NODES];
    __shared__ float lds_m[SLOTS * 3];
    __shared__ uint16_t lds_cp[MAXN + 1];
    const int ti = blockIdx.x;
    if (ti >= tile_count) return;
    const TileDev T = tiles[tile_base + ti];
    const SceneDev& S = scenes[T.scene];
    if (S.hot.mat_type == FW_MAT_SPRING_DEV) {
        float* mass = fw_at<float>(T.base, T.o_mass);
        for (int i = threadIdx.x; i < T.n_nodes; i += BLOCK) fw_st(mass + T.node0 + i, S.rho);
        return;
    }
    const uint2* rec = fw_at<const uint2>(T.base, T.o_rec) + T.rec_off;
    const float2* rest = fw_at<const float2>(T.base, T.o_aux);
    const float2* hrest = fw_at<const float2>(T.base, T.o_haux[0]);
    const uint16_t* cptr = fw_at<const uint16_t>(T.base, T.o_cptr) + T.ptr_off;
    for (int i = threadIdx.x; i < T.n_nodes; i += BLOCK) lds_X[i] = fw_ld(rest + T.node0 + i);
    for (int h = threadIdx.x; h < T.n_hnodes; h += BLOCK) lds_X[T.n_nodes + h] = fw_ld(hrest + T.hnode_off + h);
    for (int i = threadIdx.x; i <= T.n_nodes; i += BLOCK) lds_cp[i] = fw_ld(cptr + i);
    __syncthreads();
    const int nslots = T.n_own + T.n_halo;
    for (int sl = threadIdx.x; sl < nslots; sl += BLOCK) {
        uint32_t l[3], q[3];
        unpack_rec(fw_ld(rec + sl), T.n_nodes, lds_cp, l, q);
        float4 ji;
        float vol;
        rest_frame(lds_X[l[0]], lds_X[l[1]], lds_X[l[2]], ji, vol);
        if (sl < T.n_own) {
            fw_st(S.area + T.elem0 + sl, vol);
            fw_st(S.jinv + T.elem0 + sl, ji);
        }
        const float nd = vol * S.rho / 3.0f;
        if (q[0] != 0xFFFFu) lds_m[q[0]] = nd;
        if (q[1] != 0xFFFFu) lds_m[q[1]] = nd;
        if (q[2] != 0xFFFFu) lds_m[q[2]] = nd;
    }
    __syncthreads();
    float* mass = fw_at<float>(T.base, T.o_mass);
    for (int i = threadIdx.x; i < T.n_nodes; i += BLOCK) {
        float m = 0.0f;
        for (int j = lds_cp[i]; j < lds_cp[i + 1]; ++j) m += lds_m[j];
        fw_st(mass + T.node0 + i, m);
    }
}

// Selective mass scaling, flatworld/femspringmanager.py:141-201.  The reference runs
// it on one thread: for each triangle in ascending id, L = the shortest edge at the
// current positions, dt_e = 0.4 L / sqrt(E / rho); when dt_e < dt_target and
// L > 1e-15 each corner node gains dm = (E (dt / (0.4 L))^2 - rho) * area / 3; then
// for each spring, both end nodes are raised to at least 4 k dt^2.  Per node that is
// its incidence list in ascending element id, the order of the CSR entries, so one
// workgroup per tile writes the dm of each (element, corner) to LDS and each owned
// node adds its run.  An element that gains nothing contributes +0, which leaves a
// (non-negative) mass unchanged.  `buf` selects the position buffer holding the
// scene's current positions.
template <int BLOCK, int SLOTS, int NODES, int MAXN>
__global__ __launch_bounds__(BLOCK) void mass_scale_kernel(const SceneDev* scenes, const TileDev* tiles, int tile_count,
                                                           int buf, float dt_target) {
    __shared__ float2 lds_x[NODES];
    __shared__ float2 lds_X[NODES];
    __shared__ float lds_m[SLOTS * 3];
    __shared__ uint16_t lds_cp[MAXN + 1];
    const int ti = blockIdx.x;
    if (ti >= tile_count) return;
    const TileDev T = tiles[ti];
    const SceneDev& S = scenes[T.scene];
    const float E = S.hot.E;
    float* mass = fw_at<float>(T.base, T.o_mass) + T.node0;
    const uint16_t* cptr = fw_at<const uint16_t>(T.base, T.o_cptr) + T.ptr_off;
    if (S.hot.mat_type == FW_MAT_SPRING_DEV) {
        const float m_min = 4.0f * E * dt_target * dt_target;
        for (int i = threadIdx.x; i < T.n_nodes; i += BLOCK) {
            if (fw_ld(cptr + i + 1) == fw_ld(cptr + i)) continue;  // no spring ends at this node
            const float m = fw_ld(mass + i);
            if (m < m_min) fw_st(mass + i, m_min);
        }
        return;
    }
    const float2* pos = fw_at<const float2>(T.base, buf ? T.o_pos[1] : T.o_pos[0]) + T.node0;
    const float2* hpos = fw_at<const float2>(T.base, buf ? T.o_hpos[1] : T.o_hpos[0]) + T.hnode_off;
    const float2* rest = fw_at<const float2>(T.base, T.o_aux) + T.node0;
    const float2* hrest = fw_at<const float2>(T.base, T.o_haux[0]) + T.hnode_off;
    for (int i = threadIdx.x; i < T.n_nodes; i += BLOCK) {
        lds_x[i] = fw_ld(pos + i);
        lds_X[i] = fw_ld(rest + i);
    }
    for (int h = threadIdx.x; h < T.n_hnodes; h += BLOCK) {
        lds_x[T.n_nodes + h] = fw_ld(hpos + h);
        lds_X[T.n_nodes + h] = fw_ld(hrest + h);
    }
    for (int i = threadIdx.x; i <= T.n_nodes; i += BLOCK) lds_cp[i] = fw_ld(cptr + i);
    __syncthreads();
    const uint2* rec = fw_at<const uint2>(T.base, T.o_rec) + T.rec_off;
    const float rho = S.rho;
    const float safety = 0.4f;
    const float c_e = sqrtf(E / rho);
    const int nslots = T.n_own + T.n_halo;
    for (int sl = threadIdx.x; sl < nslots; sl += BLOCK) {
        uint32_t l[3], q[3];
        unpack_rec(fw_ld(rec + sl), T.n_nodes, lds_cp, l, q);
        const float2 p1 = lds_x[l[0]], p2 = lds_x[l[1]], p3 = lds_x[l[2]];
        const float a0 = p2.x - p1.x, a1 = p2.y - p1.y;
        const float b0 = p3.x - p2.x, b1 = p3.y - p2.y;
        const float c0 = p1.x - p3.x, c1 = p1.y - p3.y;
        const float e01 = sqrtf((0.0f + a0 * a0) + a1 * a1);
        const float e12 = sqrtf((0.0f + b0 * b0) + b1 * b1);
        const float e20 = sqrtf((0.0f + c0 * c0) + c1 * c1);
        const float m12 = e12 < e20 ? e12 : e20;  // Warp's min
        const float L = e01 < m12 ? e01 : m12;
        const float dt_elem = safety * L / c_e;
        float dm = 0.0f;
        if (dt_elem < dt_target && L > 1e-15f) {
            float4 ji;
            float area;
            rest_frame(lds_X[l[0]], lds_X[l[1]], lds_X[l[2]], ji, area);  // the prepare kernel's area, bit for bit
            const float r = dt_target / (safety * L);
            const float rho_new = E * r * r;
            dm = (rho_new - rho) * area / 3.0f;
        }
        if (q[0] != 0xFFFFu) lds_m[q[0]] = dm;
        if (q[1] != 0xFFFFu) lds_m[q[1]] = dm;
        if (q[2] != 0xFFFFu) lds_m[q[2]] = dm;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < T.n_nodes; i += BLOCK) {
        float m = fw_ld(mass + i);
        for (int j = lds_cp[i]; j < lds_cp[i + 1]; ++j) m = m + lds_m[j];
        fw_st(mass + i, m);
    }
}

// Plain streaming copy used to calibrate HBM counters and achievable bandwidth.
__global__ __launch_bounds__(256) void copy_probe_kernel(const float4* __restrict__ src, float4* __restrict__ dst,
                                                         int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// The same copy with ILP float4 loads per thread issued before their stores, so a
// wave keeps ILP x 1 KiB in flight; fw_bandwidth_ceiling sweeps ILP and the grid
// for the fastest read + write stream the device sustains.
template <int ILP>
__global__ __launch_bounds__(256) void stream_copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst,
                                                          int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + (ILP - 1) * stride < n; i += ILP * stride) {
        float4 v[ILP];
#pragma unroll
        for (int u = 0; u < ILP; ++u) v[u] = src[i + u * stride];
#pragma unroll
        for (int u = 0; u < ILP; ++u) dst[i + u * stride] = v[u];
    }
    for (; i < n; i += stride) dst[i] = src[i];
}

__global__ void advance_clock_kernel(ArenaCtl* ctl, int n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) ctl->clock += n;
}

}  // namespace fw

// ------------------------------------------------------------------ host launchers

// Kernel shapes (FW_SHAPES): (threads, element slots, LDS nodes, owned nodes per
// thread).  The host keeps tile_nodes <= NPT * threads and the LDS node count of a
// tile <= the shape's LDS nodes; a shape is named by (threads, slots, NPT).
// kind: 0 triangle tiles, 1 spring tiles.
extern "C" hipError_t fw_launch_substep(const SceneDev* scenes, const TileDev* tiles, const ArenaCtl* ctl, int k,
                                        int tile_base, int tile_count, float dt_override, float damping_override,
                                        uint32_t flags, int grid, int block, int slots, int npt, int kind,
                                        hipStream_t stream) {
    fw::StepArgs A{scenes, tiles, ctl, k, tile_base, tile_count, dt_override, damping_override, flags};
#define FW_LAUNCH(B, S, N, P, NS)                                                                                    \
    if (block == B && slots == S && npt == P) {                                                                 \
        if (kind == fw::KIND_SPRING)                                                                            \
            hipLaunchKernelGGL((fw::substep_kernel<B, S, N, P, NS, fw::KIND_SPRING>), dim3(grid), dim3(B), 0, stream, A); \
        else                                                                                                    \
            hipLaunchKernelGGL((fw::substep_kernel<B, S, N, P, NS, fw::KIND_TRI>), dim3(grid), dim3(B), 0, stream, A);    \
        return hipGetLastError();                                                                               \
    }
    FW_SHAPES(FW_LAUNCH)
#undef FW_LAUNCH
    return hipErrorInvalidValue;
}

// Workgroups of the substep kernel resident per CU for a shape (HIP's occupancy
// calculator: LDS, registers and wave slots).
extern "C" hipError_t fw_substep_occupancy(int block, int slots, int npt, int* per_cu) {
    *per_cu = 0;
#define FW_OCC(B, S, N, P, NS) \
    if (block == B && slots == S && npt == P)  \
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, fw::substep_kernel<B, S, N, P, NS, fw::KIND_TRI>, B, 0);
    FW_SHAPES(FW_OCC)
#undef FW_OCC
    return hipErrorInvalidValue;
}

extern "C" hipError_t fw_launch_prepare(const SceneDev* scenes, const TileDev* tiles, int tile_base, int tile_count,
                                        int block, int slots, hipStream_t stream) {
    if (tile_count <= 0) return hipSuccess;
#define FW_PREP(B, S, N, P, NS)                                                                                \
    if (block == B && slots == S) {                                                                       \
        hipLaunchKernelGGL((fw::prepare_kernel<256, S, N, B * P>), dim3(tile_count), dim3(256), 0, stream, \
                           scenes, tiles, tile_base, tile_count);                                         \
        return hipGetLastError();                                                                         \
    }
    FW_SHAPES(FW_PREP)
#undef FW_PREP
    return hipErrorInvalidValue;
}

extern "C" hipError_t fw_launch_mass_scaling(const SceneDev* scenes, const TileDev* tiles, int tile_count, int buf,
                                             float dt_target, int block, int slots, hipStream_t stream) {
    if (tile_count <= 0) return hipSuccess;
#define FW_MS(B, S, N, P, NS)                                                                                  \
    if (block == B && slots == S) {                                                                       \
        hipLaunchKernelGGL((fw::mass_scale_kernel<256, S, N, B * P>), dim3(tile_count), dim3(256), 0, stream, \
                           scenes, tiles, tile_count, buf, dt_target);                                    \
        return hipGetLastError();                                                                         \
    }
    FW_SHAPES(FW_MS)
#undef FW_MS
    return hipErrorInvalidValue;
}

extern "C" hipError_t fw_launch_copy_probe(const void* src, void* dst, int64_t n_float4, int grid, hipStream_t stream) {
    hipLaunchKernelGGL(fw::copy_probe_kernel, dim3(grid), dim3(256), 0, stream, (const float4*)src, (float4*)dst,
                       n_float4);
    return hipGetLastError();
}

extern "C" hipError_t fw_launch_stream_copy(const void* src, void* dst, int64_t n_float4, int grid, int ilp,
                                            hipStream_t stream) {
#define FW_SC(U)                                                                                              \
    if (ilp == U) {                                                                                           \
        hipLaunchKernelGGL(fw::stream_copy_kernel<U>, dim3(grid), dim3(256), 0, stream, (const float4*)src,    \
                           (float4*)dst, n_float4);                                                            \
        return hipGetLastError();                                                                             \
    }
    FW_SC(1)
    FW_SC(2)
    FW_SC(4)
    FW_SC(8)
#undef FW_SC
    return hipErrorInvalidValue;
}

extern "C" hipError_t fw_launch_rigid_pass(const SceneDev* scenes, const int2* rows, int n_rows, const ArenaCtl* ctl,
                                           int k, uint32_t flags, hipStream_t stream) {
    if (n_rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(fw::rigid_pass_kernel, dim3(n_rows), dim3(256), 0, stream, scenes, rows, ctl, k, flags);
    return hipGetLastError();
}

extern "C" hipError_t fw_launch_advance_clock(ArenaCtl* ctl, int n, hipStream_t stream) {
    hipLaunchKernelGGL(fw::advance_clock_kernel, dim3(1), dim3(64), 0, stream, ctl, n);
    return hipGetLastError();
}
