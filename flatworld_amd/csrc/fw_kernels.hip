// fw_kernels.hip -- gfx950 kernels of the explicit-FEM substep.
//
// One fused kernel per substep.  A workgroup takes one tile: a contiguous range
// of owned nodes of one scene plus every element touching them (owned elements
// are a contiguous range, halo elements come from a short list).
//   phase A  one element slot per lane: F = Jac(x) J^-1, PK2 stress
//            (Elastic / Neo-Hookean / J2 return map), corner forces
//            area * B^T S -> LDS.  J2 state is read from one buffer and the
//            owner tile writes the other (ping-pong), so halo tiles never see a
//            half-updated state.
//   phase B  one owned node per lane: ground penalty forces (plane, height
//            field, voxel grid) behind the AABB activation gate, then the
//            element forces gathered from LDS in ascending element id through
//            the tile's node CSR -- the order in which Warp's CPU backend
//            applies its atomic adds -- then BCs and symplectic Euler, then the
//            boundary AABB for the next substep's gate.
// No float atomics: every nodal sum has a fixed order, so a scene's result does
// not depend on what else is resident in the arena.  Positions are
// double-buffered because neighbouring tiles read them; velocities, mass and
// flags are touched by the owner tile only.
//
// Arithmetic follows flatworld/femspringmanager.py:59-138, 217-280, 315-418,
// flatworld/materials/materialfunctions.py:10-101 and
// flatworld/mixedcontact.py:166-302, 660-684, 1566-1643 operation by
// operation in float32; the build uses -ffp-contract=off so no FMA changes a
// rounding.  log(J) is log((double)J) rounded to float.
#include <hip/hip_runtime.h>

#include "fw_types.h"

namespace fw {

constexpr uint32_t BC_U = 1u, BC_V = 2u, BC_A = 4u, BC_F = 8u, BC_G = 16u;
constexpr uint32_t NODE_BOUNDARY = 0x8000u;
constexpr int NODE_BCV_SHIFT = 16;
constexpr uint32_t STEP_DIAG = 1u, STEP_NO_CONTACT = 2u;

__device__ __forceinline__ int mod3(int k) {
    int r = k % 3;
    return r < 0 ? r + 3 : r;
}

// ------------------------------------------------------------------ materials

struct Mat {
    int type;
    float lam, mu, sy, H;
};

// flatworld/materials/materialfunctions.py:10-29 (Elastic / Neo-Hookean PK2).
__device__ __forceinline__ void stress_hyper(const Mat& m, const float F[4], float s[3], float g[3]) {
    float C0 = F[0] * F[0] + F[2] * F[2];
    float C1 = F[0] * F[1] + F[2] * F[3];
    float C2 = F[1] * F[0] + F[3] * F[2];
    float C3 = F[1] * F[1] + F[3] * F[3];
    float e00 = 0.5f * (C0 - 1.0f);
    float e01 = 0.5f * (C1 - 0.0f);
    float e11 = 0.5f * (C3 - 1.0f);
    g[0] = e00;
    g[1] = e11;
    g[2] = e01;
    if (m.type == 1) {
        float ltr = m.lam * (e00 + e11);
        float tmu = 2.0f * m.mu;
        s[0] = ltr * 1.0f + tmu * e00;
        s[1] = ltr * 1.0f + tmu * e11;
        s[2] = ltr * 0.0f + tmu * e01;
    } else {
        float inv_det = 1.0f / (C0 * C3 - C1 * C2);
        float Ci0 = C3 * inv_det;
        float Ci1 = -C1 * inv_det;
        float Ci3 = C0 * inv_det;
        float J = F[0] * F[3] - F[1] * F[2];
        float c = m.lam * (float)log((double)J) - m.mu;
        s[0] = c * Ci0 + m.mu * 1.0f;
        s[1] = c * Ci3 + m.mu * 1.0f;
        s[2] = c * Ci1 + m.mu * 0.0f;
    }
}

// flatworld/materials/materialfunctions.py:35-101 (radial return, taken verbatim
// including the plane-strain mean stress p = (s_xx + s_yy) / 3).
__device__ __forceinline__ bool stress_j2(const Mat& m, const float F[4], float4 st, float s[3], float g[3],
                                          float4& st_new) {
    float C0 = F[0] * F[0] + F[2] * F[2];
    float C1 = F[0] * F[1] + F[2] * F[3];
    float C3 = F[1] * F[1] + F[3] * F[3];
    float et0 = 0.5f * (C0 - 1.0f);
    float et1 = 0.5f * (C3 - 1.0f);
    float et2 = 0.5f * (C1 - 0.0f);
    g[0] = et0;
    g[1] = et1;
    g[2] = et2;
    float ee0 = et0 - st.x;
    float ee1 = et1 - st.y;
    float ee2 = et2 - st.z;
    float tr = ee0 + ee1;
    float tmu = 2.0f * m.mu;
    float s0 = m.lam * tr + tmu * ee0;
    float s1 = m.lam * tr + tmu * ee1;
    float s2 = tmu * ee2;
    float p = (s0 + s1) / 3.0f;
    float d0 = s0 - p;
    float d1 = s1 - p;
    float d2 = s2;
    float vm = sqrtf(d0 * d0 + d1 * d1 - d0 * d1 + 3.0f * d2 * d2);
    float f_trial = vm - (m.sy + m.H * st.w);
    st_new = st;
    s[0] = s0;
    s[1] = s1;
    s[2] = s2;
    if (f_trial > 0.0f) {
        float denom = 3.0f * m.mu + m.H;
        float dg = f_trial / denom;
        st_new.w = st.w + dg;
        float scale = 1.0f - 3.0f * m.mu * dg / vm;
        s[0] = d0 * scale + p;
        s[1] = d1 * scale + p;
        s[2] = d2 * scale;
        st_new.x = st.x + dg * (d0 / vm);
        st_new.y = st.y + dg * (d1 / vm);
        st_new.z = st.z + dg * (d2 / vm);
        return true;
    }
    return false;
}

// ------------------------------------------------------------------ grounds

// flatworld/mixedcontact.py:166-217 (_nearest_on_curve_2d); returns the signed distance.
__device__ __forceinline__ float hf_signed(const GroundDev& g, float x, float z, float& n0o, float& n1o) {
    const float* hgt = (const float*)g.data;
    const int nx = g.nx;
    const float lbx = g.a[0], ubx = g.a[1];
    float span = ubx - lbx;
    float u = 0.0f;
    if (span > 1e-12f) u = (x - lbx) / span;
    u = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
    float sg = u * (float)(nx - 1);
    int i0 = (int)floorf(sg);
    int i1 = i0 + 1;
    if (i1 >= nx) i1 = nx - 1;
    if (i0 < 0) i0 = 0;
    float t = sg - (float)i0;
    float h0 = hgt[i0], h1 = hgt[i1];
    float h = h0 + (h1 - h0) * t;
    int ir = (int)roundf(sg);
    if (ir < 0) ir = 0;
    if (ir >= nx) ir = nx - 1;
    int j0 = ir - 1, j1 = ir + 1;
    if (j0 < 0) j0 = 0;
    if (j1 >= nx) j1 = nx - 1;
    float dxw = span / (float)(nx - 1);
    float denom = dxw * (float)(j1 - j0);
    float dhdx = 0.0f;
    if (denom > 1e-6f) dhdx = (hgt[j1] - hgt[j0]) / denom;
    float n0 = -dhdx, n1 = 1.0f;
    if (g.reverse == 1) {
        n0 = -n0;
        n1 = -n1;
    }
    float nlen = sqrtf(n0 * n0 + n1 * n1);
    if (nlen > 1e-12f) {
        n0 = n0 / nlen;
        n1 = n1 / nlen;
    }
    n0o = n0;
    n1o = n1;
    return (x - x) * n0 + (z - h) * n1;
}

// flatworld/mixedcontact.py:220-302 (_voxel_signed_distance_2d, interior cells only).
__device__ __forceinline__ float voxel_signed(const GroundDev& g, float px, float py, float& n0o, float& n1o) {
    const int8_t* occ = (const int8_t*)g.data;
    const int nx = g.nx, ny = g.ny;
    const float lb0 = g.a[0], lb1 = g.a[1], dx = g.a[2], dz = g.a[3];
    float best = 1e9f;
    n0o = 0.0f;
    n1o = 1.0f;
    int i = (int)floorf((px - lb0) / dx);
    int j = (int)floorf((py - lb1) / dz);
    if (0 <= i && i < nx && 0 <= j && j < ny && occ[(int64_t)i * ny + j] == 1) {
        float dnx = 1e9f, dpx = 1e9f, dny = 1e9f, dpy = 1e9f;
        int idx = i - 1;
        while (idx >= 0 && occ[(int64_t)idx * ny + j] != 0) --idx;
        dnx = px - (lb0 + (float)(idx + 1) * dx);
        idx = i + 1;
        while (idx < nx && occ[(int64_t)idx * ny + j] != 0) ++idx;
        dpx = (lb0 + (float)idx * dx) - px;
        idx = j - 1;
        while (idx >= 0 && occ[(int64_t)i * ny + idx] != 0) --idx;
        dny = py - (lb1 + (float)(idx + 1) * dz);
        idx = j + 1;
        while (idx < ny && occ[(int64_t)i * ny + idx] != 0) ++idx;
        dpy = (lb1 + (float)idx * dz) - py;
        float mv = dnx;
        int face = 0;
        if (dpx < mv) { mv = dpx; face = 1; }
        if (dny < mv) { mv = dny; face = 2; }
        if (dpy < mv) { mv = dpy; face = 3; }
        if (face == 0) { best = -dnx; n0o = -1.0f; n1o = 0.0f; }
        else if (face == 1) { best = -dpx; n0o = 1.0f; n1o = 0.0f; }
        else if (face == 2) { best = -dny; n0o = 0.0f; n1o = -1.0f; }
        else { best = -dpy; n0o = 0.0f; n1o = 1.0f; }
    }
    return best;
}

// One ground on one boundary node; adds into (fx, fy) only when it pushes.
// Plane flatworld/mixedcontact.py:660-684, height field 1566-1602, voxel 1605-1643.
__device__ __forceinline__ bool ground_force(const GroundDev& g, float px, float py, float vx, float vy, float dt,
                                             float& fx, float& fy) {
    const float k = g.penalty;
    if (g.kind == 1) {
        float pen = (px - g.a[0]) * g.a[2] + (py - g.a[1]) * g.a[3];
        if (pen < 0.0f) {
            fx = fx + (-((g.a[2] * k) * pen));
            fy = fy + (-((g.a[3] * k) * pen));
            return true;
        }
        return false;
    }
    float n0, n1, sd;
    if (g.kind == 2) {
        sd = hf_signed(g, px, py, n0, n1);
        if (!(sd < 0.0f)) return false;
        float c0 = ((-n0) * k) * sd;
        float c1 = ((-n1) * k) * sd;
        float vn = vx * n0 + vy * n1;
        if (vn < 0.0f) {
            c0 = c0 + (((((-n0) * 0.5f) * k) * vn) * dt);
            c1 = c1 + (((((-n1) * 0.5f) * k) * vn) * dt);
        }
        fx = fx + c0;
        fy = fy + c1;
        return true;
    }
    sd = voxel_signed(g, px, py, n0, n1);
    if (!(sd < 0.0f)) return false;
    float pen = -sd;
    float c0 = (n0 * k) * pen;
    float c1 = (n1 * k) * pen;
    float vn = vx * n0 + vy * n1;
    if (vn < 0.0f) {
        c0 = c0 + (((((-n0) * 0.5f) * k) * vn) * dt);
        c1 = c1 + (((((-n1) * 0.5f) * k) * vn) * dt);
    }
    fx = fx + c0;
    fy = fy + c1;
    return true;
}

// ------------------------------------------------------------------ tile loop helpers

__device__ __forceinline__ int tile_of_block(int round, int ntiles_total_grid) {
    // XCD-aware mapping: hardware places block b on XCD b % 8; give each XCD a
    // contiguous run of tiles so halo elements and nodes shared by neighbouring
    // tiles are reused from that XCD's L2.
    const int G = gridDim.x;
    const int b = blockIdx.x;
    int v = b;
    if ((G & 7) == 0) v = (b & 7) * (G >> 3) + (b >> 3);
    return v + round * G;
}

template <int BLOCK>
__device__ __forceinline__ void block_minmax(int32_t v[4], int32_t* red) {
    // v: lb.x, lb.y (min), ub.x, ub.y (max) as ordered ints
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        int32_t a0 = __shfl_xor(v[0], off, 64);
        int32_t a1 = __shfl_xor(v[1], off, 64);
        int32_t a2 = __shfl_xor(v[2], off, 64);
        int32_t a3 = __shfl_xor(v[3], off, 64);
        v[0] = min(v[0], a0);
        v[1] = min(v[1], a1);
        v[2] = max(v[2], a2);
        v[3] = max(v[3], a3);
    }
    if (lane == 0) {
        red[wid * 4 + 0] = v[0];
        red[wid * 4 + 1] = v[1];
        red[wid * 4 + 2] = v[2];
        red[wid * 4 + 3] = v[3];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < BLOCK / 64; ++w) {
            v[0] = min(v[0], red[w * 4 + 0]);
            v[1] = min(v[1], red[w * 4 + 1]);
            v[2] = max(v[2], red[w * 4 + 2]);
            v[3] = max(v[3], red[w * 4 + 3]);
        }
    }
}

// One element slot: F = Jac(x) J^-1, stress, corner forces area * B^T S
// (flatworld/femspringmanager.py:71-105, 240-276), stored at the LDS entries
// of the owning tile's node-major incidence lists.
__device__ __forceinline__ void slot_force(const Mat& m, float2 p1, float2 p2, float2 p3, float4 ji, float ar, float4 st,
                                           uint2 dd, bool own, int e, bool diag, float4* __restrict__ j2_wr,
                                           float4* __restrict__ stress, float4* __restrict__ strain,
                                           uint8_t* __restrict__ yflag, float2* lds_f) {
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
        if (own) j2_wr[e] = stn;
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
    const uint32_t q0 = dd.x & 0xFFFFu, q1 = dd.x >> 16, q2 = dd.y & 0xFFFFu;
    if (q0 != 0xFFFFu) lds_f[q0] = make_float2(f0, f1);
    if (q1 != 0xFFFFu) lds_f[q1] = make_float2(f2, f3);
    if (q2 != 0xFFFFu) lds_f[q2] = make_float2(f4, f5);
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
        stress[e] = make_float4(c00 / jd, c11 / jd, c01 / jd, 0.0f);
        strain[e] = make_float4(g[0], g[1], g[2], 0.0f);
        yflag[e] = yielded ? 1 : 0;
    }
}

// ------------------------------------------------------------------ substep kernel

struct StepArgs {
    const SceneDev* scenes;
    const TileDev* tiles;
    const ArenaCtl* ctl;
    int k;           // substep index within this launch group (global step = ctl->clock + k)
    int tile_base;   // first tile of the range
    int tile_count;  // tiles in the range; < 0: up to ctl->n_tiles
    float dt_override;
    float damping_override;
    uint32_t flags;
};

template <int BLOCK, int SLOTS>
__global__ __launch_bounds__(BLOCK) void substep_kernel(StepArgs A) {
    __shared__ float2 lds_f[SLOTS * 3];
    __shared__ int32_t lds_red[(BLOCK / 64) * 4];

    const int G = A.ctl->clock + A.k;
    const int ntiles = A.tile_count >= 0 ? A.tile_count : (A.ctl->n_tiles - A.tile_base);

    for (int round = 0;; ++round) {
        const int ti = tile_of_block(round, ntiles);
        if (ti >= ntiles) break;
        const TileDev T = A.tiles[A.tile_base + ti];
        const SceneDev& S = A.scenes[T.scene];
        const int local = G - S.s0;
        if (!S.live || local < -1 || local >= S.horizon) continue;  // uniform per block

        const bool warm = local == -1;
        const float dt = warm ? 0.0f : (A.dt_override >= 0.0f ? A.dt_override : S.dt);
        const float damping = warm ? 0.0f : (A.damping_override >= 0.0f ? A.damping_override : S.damping);
        const bool diag = (A.flags & STEP_DIAG) != 0;
        const bool contact = !warm && !(A.flags & STEP_NO_CONTACT) && S.n_grounds > 0;
        const int rd = (local + 1) & 1;
        const int wr = local & 1;
        const float2* __restrict__ pos_rd = S.pos[rd];
        float2* __restrict__ pos_wr = S.pos[wr];

        Mat m;
        m.type = S.mat_type;
        m.lam = S.E * S.nu / ((1.0f + S.nu) * (1.0f - 2.0f * S.nu));
        m.mu = S.E / (2.0f * (1.0f + S.nu));
        m.sy = S.sigma_y;
        m.H = S.H;

        // ---------------- phase A: element slots, two per lane per iteration with
        // every independent load of both issued before the dependent gathers
        const int nslots = T.n_own + T.n_halo;
        const uint2* __restrict__ dest = S.dest + T.dest_off;
        const bool j2 = m.type == 3;
        for (int slA = threadIdx.x; slA < nslots; slA += 2 * BLOCK) {
            const int slB = slA + BLOCK;
            const bool vB = slB < nslots;
            const bool ownA = slA < T.n_own, ownB = slB < T.n_own;
            const int eA = ownA ? T.elem0 + slA : S.halo[T.halo_off + slA - T.n_own];
            const int eB = !vB ? eA : (ownB ? T.elem0 + slB : S.halo[T.halo_off + slB - T.n_own]);
            const int a1 = S.c0[eA], a2 = S.c1[eA], a3 = S.c2[eA];
            const int b1 = S.c0[eB], b2 = S.c1[eB], b3 = S.c2[eB];
            const float4 jiA = S.jinv[eA], jiB = S.jinv[eB];
            const float arA = S.area[eA], arB = S.area[eB];
            const uint2 ddA = dest[slA];
            const uint2 ddB = vB ? dest[slB] : make_uint2(0xFFFFFFFFu, 0xFFFFFFFFu);
            float4 stA = make_float4(0.0f, 0.0f, 0.0f, 0.0f), stB = stA;
            if (j2) {
                stA = S.j2[rd][eA];
                stB = S.j2[rd][eB];
            }
            const float2 pA1 = pos_rd[a1], pA2 = pos_rd[a2], pA3 = pos_rd[a3];
            const float2 pB1 = pos_rd[b1], pB2 = pos_rd[b2], pB3 = pos_rd[b3];
            slot_force(m, pA1, pA2, pA3, jiA, arA, stA, ddA, ownA, eA, diag, S.j2[wr], S.stress, S.strain, S.yflag, lds_f);
            if (vB)
                slot_force(m, pB1, pB2, pB3, jiB, arB, stB, ddB, ownB, eB, diag, S.j2[wr], S.stress, S.strain, S.yflag,
                           lds_f);
        }

        // ---------------- activation gate (flatworld/explicitloop.py:477-487, bvh.py:39-46, 321-342, 393-416)
        uint32_t active = 0;
        if (contact) {
            const int ar3 = mod3(local);
            float lbx = fw_o2f(S.aabb[ar3][0]), lby = fw_o2f(S.aabb[ar3][1]);
            float ubx = fw_o2f(S.aabb[ar3][2]), uby = fw_o2f(S.aabb[ar3][3]);
            const bool pad = ((local - S.gate_base) % 100) == 0;
            if (pad) {
                const float bx = 0.1f * (ubx - lbx), by = 0.1f * (uby - lby);
                lbx = lbx - bx;
                ubx = ubx + bx;
                lby = lby - by;
                uby = uby + by;
            }
            for (int gi = 0; gi < S.n_grounds; ++gi) {
                const GroundDev& gd = S.ground[gi];
                const float* lo = pad ? gd.plo : gd.lo;
                const float* hi = pad ? gd.phi : gd.hi;
                if (lbx <= hi[0] && ubx >= lo[0] && lby <= hi[1] && uby >= lo[1]) active |= 1u << gi;
            }
        }
        __syncthreads();

        // ---------------- phase B: owned nodes
        int32_t bb[4] = {fw_f2o(1e30f), fw_f2o(1e30f), fw_f2o(-1e30f), fw_f2o(-1e30f)};
        const uint16_t* __restrict__ cptr = S.csr_ptr + T.ptr_off;
        for (int i = threadIdx.x; i < T.n_nodes; i += BLOCK) {
            const int n = T.node0 + i;
            const float2 p = pos_rd[n];
            float2 v = S.vel[n];
            const uint32_t fl = S.flag[n];
            const float mass = S.mass[n];
            const int jb = cptr[i], je = cptr[i + 1];
            float fx = 0.0f, fy = 0.0f;
            bool hit = false;
            if (active && (fl & NODE_BOUNDARY)) {
                for (int gi = 0; gi < S.n_grounds; ++gi)
                    if (active & (1u << gi)) hit |= ground_force(S.ground[gi], p.x, p.y, v.x, v.y, dt, fx, fy);
            }
            // element forces in ascending element id: a contiguous run of LDS entries
            for (int j = jb; j < je; ++j) {
                const float2 f = lds_f[j];
                fx = fx + (-f.x);
                fy = fy + (-f.y);
            }
            // symplectic Euler with BC precedence (flatworld/femspringmanager.py:315-358)
            const float2 bv = S.bcv[fl >> NODE_BCV_SHIFT];
            if (fl & BC_U) {
                fx = 0.0f;
                fy = 0.0f;
            } else if (fl & BC_A) {
                fx = mass * bv.x;
                fy = mass * bv.y;
            } else if (fl & BC_G) {
                fx = fx + mass * bv.x;
                fy = fy + mass * bv.y;
            } else if (fl & BC_F) {
                fx = fx + bv.x;
                fy = fy + bv.y;
            }
            const float rm = 1.0f / mass;
            const float ax = 0.0f + (fx - damping * v.x) * rm;
            const float ay = 0.0f + (fy - damping * v.y) * rm;
            float vx = v.x + ax * dt;
            float vy = v.y + ay * dt;
            if (fl & BC_V) {
                vx = bv.x;
                vy = bv.y;
            } else if (fl & BC_U) {
                vx = 0.0f;
                vy = 0.0f;
            }
            S.vel[n] = make_float2(vx, vy);
            const float dux = vx * dt, duy = vy * dt;
            if (S.disp) {
                const float2 u = S.disp[n];
                S.disp[n] = make_float2(u.x + dux, u.y + duy);
            }
            const float nx_ = p.x + dux, ny_ = p.y + duy;
            pos_wr[n] = make_float2(nx_, ny_);
            if (diag) {
                S.force[n] = make_float2(fx, fy);
                S.cflag[n] = hit ? 1 : 0;
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
        block_minmax<BLOCK>(bb, lds_red);
        if (threadIdx.x == 0) {
            int32_t* aw = const_cast<int32_t*>(S.aabb[mod3(local + 1)]);
            int32_t* az = const_cast<int32_t*>(S.aabb[mod3(local + 2)]);
            if (T.multi) {
                atomicMin(aw + 0, bb[0]);
                atomicMin(aw + 1, bb[1]);
                atomicMax(aw + 2, bb[2]);
                atomicMax(aw + 3, bb[3]);
                az[0] = fw_f2o(1e30f);
                az[1] = fw_f2o(1e30f);
                az[2] = fw_f2o(-1e30f);
                az[3] = fw_f2o(-1e30f);
            } else {
                aw[0] = bb[0];
                aw[1] = bb[1];
                aw[2] = bb[2];
                aw[3] = bb[3];
            }
        }
        __syncthreads();  // LDS reuse by the next tile of this block
    }
}

// ------------------------------------------------------------------ prepare kernel

// flatworld/femspringmanager.py:113-138: area = |det Jac| / 2, J^-1 = inverse(Jac),
// lumped mass = sum over incident elements (ascending id) of area * rho / 3.
template <int BLOCK, int SLOTS>
__global__ __launch_bounds__(BLOCK) void prepare_kernel(const SceneDev* scenes, const TileDev* tiles, int tile_base,
                                                        int tile_count) {
    __shared__ float lds_m[SLOTS * 3];
    const int ti = blockIdx.x;
    if (ti >= tile_count) return;
    const TileDev T = tiles[tile_base + ti];
    const SceneDev& S = scenes[T.scene];
    const float2* __restrict__ x0 = S.pos[0];
    const int nslots = T.n_own + T.n_halo;
    for (int sl = threadIdx.x; sl < nslots; sl += BLOCK) {
        const bool own = sl < T.n_own;
        const int e = own ? T.elem0 + sl : S.halo[T.halo_off + sl - T.n_own];
        const float2 p1 = x0[S.c0[e]], p2 = x0[S.c1[e]], p3 = x0[S.c2[e]];
        const float J0 = p1.x - p3.x, J1 = p2.x - p3.x, J2 = p1.y - p3.y, J3 = p2.y - p3.y;
        const float det = J0 * J3 - J1 * J2;
        const float vol = fabsf(det) * 0.5f;
        if (own) {
            const float inv_det = 1.0f / det;
            S.area[e] = vol;
            S.jinv[e] = make_float4(J3 * inv_det, -J1 * inv_det, -J2 * inv_det, J0 * inv_det);
        }
        const float nd = vol * S.rho / 3.0f;
        const uint2 dd = S.dest[T.dest_off + sl];
        const uint32_t q0 = dd.x & 0xFFFFu, q1 = dd.x >> 16, q2 = dd.y & 0xFFFFu;
        if (q0 != 0xFFFFu) lds_m[q0] = nd;
        if (q1 != 0xFFFFu) lds_m[q1] = nd;
        if (q2 != 0xFFFFu) lds_m[q2] = nd;
    }
    __syncthreads();
    const uint16_t* cptr = S.csr_ptr + T.ptr_off;
    for (int i = threadIdx.x; i < T.n_nodes; i += BLOCK) {
        float m = 0.0f;
        for (int j = cptr[i]; j < cptr[i + 1]; ++j) m += lds_m[j];
        S.mass[T.node0 + i] = m;
    }
}

// Plain streaming copy used to calibrate HBM counters and achievable bandwidth.
__global__ __launch_bounds__(256) void copy_probe_kernel(const float4* __restrict__ src, float4* __restrict__ dst,
                                                         int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

__global__ void advance_clock_kernel(ArenaCtl* ctl, int n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) ctl->clock += n;
}

}  // namespace fw

// ------------------------------------------------------------------ host launchers

extern "C" hipError_t fw_launch_substep(const SceneDev* scenes, const TileDev* tiles, const ArenaCtl* ctl, int k,
                                        int tile_base, int tile_count, float dt_override, float damping_override,
                                        uint32_t flags, int grid, int block, int slots, hipStream_t stream) {
    fw::StepArgs A{scenes, tiles, ctl, k, tile_base, tile_count, dt_override, damping_override, flags};
    if (block == 256 && slots == 1024)
        hipLaunchKernelGGL((fw::substep_kernel<256, 1024>), dim3(grid), dim3(256), 0, stream, A);
    else if (block == 256 && slots == 2048)
        hipLaunchKernelGGL((fw::substep_kernel<256, 2048>), dim3(grid), dim3(256), 0, stream, A);
    else if (block == 512 && slots == 2048)
        hipLaunchKernelGGL((fw::substep_kernel<512, 2048>), dim3(grid), dim3(512), 0, stream, A);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

extern "C" hipError_t fw_launch_prepare(const SceneDev* scenes, const TileDev* tiles, int tile_base, int tile_count,
                                        int slots, hipStream_t stream) {
    if (tile_count <= 0) return hipSuccess;
    if (slots == 1024)
        hipLaunchKernelGGL((fw::prepare_kernel<256, 1024>), dim3(tile_count), dim3(256), 0, stream, scenes, tiles,
                           tile_base, tile_count);
    else if (slots == 2048)
        hipLaunchKernelGGL((fw::prepare_kernel<256, 2048>), dim3(tile_count), dim3(256), 0, stream, scenes, tiles,
                           tile_base, tile_count);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

extern "C" hipError_t fw_launch_copy_probe(const void* src, void* dst, int64_t n_float4, int grid, hipStream_t stream) {
    hipLaunchKernelGGL(fw::copy_probe_kernel, dim3(grid), dim3(256), 0, stream, (const float4*)src, (float4*)dst,
                       n_float4);
    return hipGetLastError();
}

extern "C" hipError_t fw_launch_advance_clock(ArenaCtl* ctl, int n, hipStream_t stream) {
    hipLaunchKernelGGL(fw::advance_clock_kernel, dim3(1), dim3(64), 0, stream, ctl, n);
    return hipGetLastError();
}
