// fw_device.h -- device helpers of the substep kernels: material laws, ground contact
// laws, tile scheduling and block reductions.  Included by fw_kernels.hip only.
#pragma once
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

// Natural log of a float, evaluated with IEEE double operations only and rounded
// to float once: x = m * 2^e with m in [sqrt(1/2), sqrt(2)), then
// log m = 2 atanh(s), s = (m - 1) / (m + 1), |s| <= 0.172, summed to s^13
// (truncation < 2e-12 relative).  Every operation is correctly rounded and none
// is contracted, so the CPU oracle's copy of this routine gives the same bits.
__device__ __forceinline__ float fw_logf(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -__builtin_inff() : __builtin_nanf("");
    if (x == __builtin_inff()) return x;
    int e = 0;
    if (x < 1.17549435e-38f) {  // subnormal: scale into the normal range (exact)
        x = x * 8388608.0f;
        e = -23;
    }
    uint32_t b = __float_as_uint(x);
    e += (int)((b >> 23) & 0xFFu) - 127;
    float m = __uint_as_float((b & 0x007FFFFFu) | 0x3F800000u);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e += 1;
    }
    const double f = (double)m - 1.0;
    const double s = f / (2.0 + f);
    const double s2 = s * s;
    double p = 1.0 / 13.0;
    p = p * s2 + 1.0 / 11.0;
    p = p * s2 + 1.0 / 9.0;
    p = p * s2 + 1.0 / 7.0;
    p = p * s2 + 1.0 / 5.0;
    p = p * s2 + 1.0 / 3.0;
    p = p * s2 + 1.0;
    const double r = (double)e * 0.69314718055994530942 + (2.0 * s) * p;
    return (float)r;
}

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
        float c = m.lam * fw_logf(J) - m.mu;
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
// The unit normal at the rounded sample index comes from the per-sample table the
// host evaluated with the same float32 operations (nx float2 after the heights).
__device__ __forceinline__ float hf_signed(const GroundDev& g, float x, float z, float& n0o, float& n1o) {
    const float* hgt = (const float*)g.data;
    const int nx = g.nx;
    const float2* nrm = (const float2*)(hgt + ((nx + 1) & ~1));
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
    int ir = (int)roundf(sg);
    if (ir < 0) ir = 0;
    if (ir >= nx) ir = nx - 1;
    // the three samples in one round trip
    const float h0 = fw_ld(hgt + i0), h1 = fw_ld(hgt + i1);
    const float2 n = fw_ld(nrm + ir);
    float t = sg - (float)i0;
    float h = h0 + (h1 - h0) * t;
    n0o = n.x;
    n1o = n.y;
    return (x - x) * n.x + (z - h) * n.y;
}

// flatworld/mixedcontact.py:220-302 (_voxel_signed_distance_2d, interior cells only).
// The reference scans from an occupied cell along -x, +x, -z, +z for the first empty
// cell; the scan end points come precomputed per cell (host build_voxel_extents), so
// one 8-byte load replaces the walk and the distances below are the same float
// expressions on the same integers.
__device__ __forceinline__ float voxel_signed(const GroundDev& g, float px, float py, float& n0o, float& n1o) {
    const float lb0 = g.a[0], lb1 = g.a[1], dx = g.a[2], dz = g.a[3];
    float best = 1e9f;
    n0o = 0.0f;
    n1o = 1.0f;
    const int i = (int)floorf((px - lb0) / dx);
    const int j = (int)floorf((py - lb1) / dz);
    if (0 <= i && i < g.nx && 0 <= j && j < g.ny) {
        const uint2 e = fw_ld((const uint2*)g.data + (int64_t)i * g.ny + j);
        if ((e.x & 0xFFFFu) != 0xFFFFu) {
            const float dnx = px - (lb0 + (float)(int)(e.x & 0xFFFFu) * dx);
            const float dpx = (lb0 + (float)(int)(e.x >> 16) * dx) - px;
            const float dny = py - (lb1 + (float)(int)(e.y & 0xFFFFu) * dz);
            const float dpy = (lb1 + (float)(int)(e.y >> 16) * dz) - py;
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
    }
    return best;
}

// One ground on one boundary node; adds into (fx, fy) only when it pushes.
// Plane flatworld/mixedcontact.py:660-684, height field 1566-1602, voxel 1605-1643.
// Warp's vec2 helpers as the rigid kernels use them: dot and length as a
// zero-initialised sum of products.
__device__ __forceinline__ float wdot(float ax, float ay, float bx, float by) { return (0.0f + ax * bx) + ay * by; }
__device__ __forceinline__ float wlen(float x, float y) { return sqrtf(wdot(x, y, x, y)); }
__device__ __forceinline__ float wmax(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float wmin(float a, float b) { return a < b ? a : b; }

// detectPointToPrimitive (flatworld/contact_detection.py:147-191) with sphereSDFQuery,
// obbSDFQuery and capsuleSDFQuery / calMinDisNode2Segment (flatworld/utils.py:34-145):
// penetration, normal and (when penetrating) the contact point.
__device__ __forceinline__ float rigid_detect(const GroundDev& g, float px, float py, float& n0, float& n1,
                                              float& cx, float& cy) {
    const float* r = g.b;  // rotation, row-major
    float pen = 1e9f;
    n0 = 0.0f;
    n1 = 0.0f;
    cx = px;
    cy = py;
    if (g.nx == 1) {  // ball
        const float dx = px - g.a[0], dy = py - g.a[1];
        const float nl = wlen(dx, dy);
        pen = nl - g.hf_dxw;
        const float d = wmax(nl, 1e-9f);
        n0 = dx / d;
        n1 = dy / d;
    } else if (g.nx == 2) {  // box, full extents a[2..3]
        float best = -1e9f, m0 = 0.0f, m1 = 0.0f;
        const float h0 = 0.5f * g.a[2], h1 = 0.5f * g.a[3];
        const float q0 = px - g.a[0], q1 = py - g.a[1];
        const float l0 = r[0] * q0 + r[2] * q1;  // transpose(rot) @ q
        const float l1 = r[1] * q0 + r[3] * q1;
        const float sd0 = l0 - h0, sd1 = -l0 - h0, sd2 = l1 - h1, sd3 = -l1 - h1;
        if (sd0 > best) { best = sd0; m0 = 1.0f; m1 = 0.0f; }
        if (sd1 > best) { best = sd1; m0 = -1.0f; m1 = 0.0f; }
        if (sd2 > best) { best = sd2; m0 = 0.0f; m1 = 1.0f; }
        if (sd3 > best) { best = sd3; m0 = 0.0f; m1 = -1.0f; }
        if (sd0 > 0.0f || sd1 > 0.0f || sd2 > 0.0f || sd3 > 0.0f) {
            float kx = l0, ky = l1;
            if (l0 < -h0) kx = -h0;
            else if (l0 > h0) kx = h0;
            if (l1 < -h1) ky = -h1;
            else if (l1 > h1) ky = h1;
            const float dx = l0 - kx, dy = l1 - ky;
            const float dist = wlen(dx, dy);
            if (dist > 1e-9f) {
                m0 = dx / dist;
                m1 = dy / dist;
            }
            best = dist;
        }
        n0 = r[0] * m0 + r[1] * m1;
        n1 = r[2] * m0 + r[3] * m1;
        pen = best;
    } else if (g.nx == 8) {  // capsule: axis from lc = centre + rot @ a[2..3] to 2 centre - lc
        const float lcx = g.a[0] + (r[0] * g.a[2] + r[1] * g.a[3]);
        const float lcy = g.a[1] + (r[2] * g.a[2] + r[3] * g.a[3]);
        const float qx = 2.0f * g.a[0] - lcx, qy = 2.0f * g.a[1] - lcy;
        const float apx = px - lcx, apy = py - lcy;
        const float abx = qx - lcx, aby = qy - lcy;
        const float ab_sq = wdot(abx, aby, abx, aby);
        float t = wdot(apx, apy, abx, aby) / ab_sq;
        t = wmax(0.0f, wmin(1.0f, t));
        const float clx = lcx + t * abx, cly = lcy + t * aby;
        pen = wlen(px - clx, py - cly) - g.hf_dxw;
        const float vx = px - clx, vy = py - cly;
        const float d = wmax(wlen(vx, vy), 1e-9f);
        n0 = vx / d;
        n1 = vy / d;
    } else {
        return pen;
    }
    if (pen < 0.0f) {
        cx = px - n0 * pen;
        cy = py - n1 * pen;
    }
    return pen;
}

// _batched_rigid_prim_kernel (flatworld/mixedcontact.py:1025-1117), the node's part:
// nothing outside the rigid's AABB; a tied node pulled to its contact point within the
// margin (half the AABB diagonal); otherwise the penalty force k (1e-4 - pen) n where
// pen < 1e-4, with Coulomb friction against the rigid's surface velocity.
// Evaluated by the rigid pass ahead of the substep (fw_kernels.hip rigid_pass_kernel),
// whose force the substep adds at the record's place among the node's ground forces.
__device__ __forceinline__ bool rigid_force(const GroundDev& g, float px, float py, float vx, float vy, float& fx,
                                            float& fy) {
    if (wmin(px - g.lo[0], py - g.lo[1]) < 0.0f || wmin(g.hi[0] - px, g.hi[1] - py) < 0.0f) return false;
    float n0, n1, cx, cy;
    const float pen = rigid_detect(g, px, py, n0, n1, cx, cy);
    const float margin = wlen(g.hi[0] - g.lo[0], g.hi[1] - g.lo[1]) * 0.5f;
    const float k = g.penalty;
    if (g.ny == 1 && pen < margin) {
        fx = fx + k * (cx - px);
        fy = fy + k * (cy - py);
        return true;
    }
    if (!(pen < margin)) return false;
    const float penetration = pen - 1e-4f;
    if (!(penetration < 0.0f)) return false;
    const float nf0 = ((-n0) * k) * penetration, nf1 = ((-n1) * k) * penetration;
    float t0 = nf0, t1 = nf1;
    const float mu = g.b[7];
    if (mu > 1e-9f) {
        const float rx = cx - g.a[0], ry = cy - g.a[1];
        const float w = g.b[6];
        const float sv0 = g.b[4] + (-w * ry), sv1 = g.b[5] + w * rx;
        const float rv0 = vx - sv0, rv1 = vy - sv1;
        const float vn = wdot(rv0, rv1, n0, n1);
        const float tv0 = rv0 - vn * n0, tv1 = rv1 - vn * n1;
        const float tlen = wlen(tv0, tv1);
        if (tlen > 1e-9f) {
            const float fd0 = -tv0 / tlen, fd1 = -tv1 / tlen;
            const float nfl = wlen(nf0, nf1);
            t0 = t0 + (fd0 * mu) * nfl;
            t1 = t1 + (fd1 * mu) * nfl;
        }
    }
    fx = fx + t0;
    fy = fy + t1;
    return true;
}

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
        // above the highest sample plus the rounding margin (below the lowest, reversed),
        // the signed distance is >= 0 or NaN: no contact, without the query's loads
        if (g.reverse == 1 ? py < g.b[0] : py > g.b[0]) return false;
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

// The tile block b processes in its round-th pass, or a value >= ntiles once it has
// none left.  XCD-aware: hardware places block b on XCD b % 8, and each XCD takes
// one contiguous eighth of the live tiles (its blocks striding through it), so halo
// elements and nodes shared by neighbouring tiles are reused from that XCD's L2 and
// every XCD gets work whatever the grid: a graph's grid covers the tile table's
// capacity, often twice the live tiles, and an eighth of the grid per XCD would
// leave the upper XCDs with nothing but exits.
__device__ __forceinline__ int tile_of_block(int round, int ntiles) {
    const int G = gridDim.x;
    const int b = blockIdx.x;
    if ((G & 7) != 0) return b + round * G;
    const int chunk = (ntiles + 7) >> 3;         // tiles per XCD
    const int j = (b >> 3) + round * (G >> 3);   // this block's position in its XCD's run
    if (j >= chunk) return ntiles;
    return (b & 7) * chunk + j;                  // >= ntiles only in the last XCD's short run
}

// Workgroup barrier that orders LDS only.  Threads of a tile never exchange data
// through global memory within a launch (positions, halo copies and J2 state are
// double-buffered and read in the next launch), so global loads issued before
// the barrier may stay in flight across it; __syncthreads would wait for them.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// One DPP step of a wave reduction: v combined with the value DPP control CTRL moves
// into each lane of the rows ROW_MASK enables (lanes without a source keep v).
template <int CTRL, int ROW_MASK, bool IS_MAX>
__device__ __forceinline__ int32_t dpp_step(int32_t v) {
    const int32_t o = __builtin_amdgcn_update_dpp(v, v, CTRL, ROW_MASK, 0xF, false);
    return IS_MAX ? max(v, o) : min(v, o);
}

template <bool IS_MAX>
__device__ __forceinline__ int32_t wave_reduce(int32_t v) {
    // inclusive scan inside each row of 16 lanes (row_shr 1, 2, 4, 8), then lane 15 of
    // rows 0 and 2 into rows 1 and 3 (row_bcast 15), then lane 31 into rows 2 and 3
    // (row_bcast 31): lane 63 holds the result, read back as a uniform value
    v = dpp_step<0x111, 0xF, IS_MAX>(v);
    v = dpp_step<0x112, 0xF, IS_MAX>(v);
    v = dpp_step<0x114, 0xF, IS_MAX>(v);
    v = dpp_step<0x118, 0xF, IS_MAX>(v);
    v = dpp_step<0x142, 0xA, IS_MAX>(v);
    v = dpp_step<0x143, 0xC, IS_MAX>(v);
    return __builtin_amdgcn_readlane(v, 63);
}

// Min of v[0..1] and max of v[2..3] over the 64 lanes of a wave (ordered ints);
// every lane ends with the result.
__device__ __forceinline__ void wave_minmax(int32_t v[4]) {
    v[0] = wave_reduce<false>(v[0]);
    v[1] = wave_reduce<false>(v[1]);
    v[2] = wave_reduce<true>(v[2]);
    v[3] = wave_reduce<true>(v[3]);
}

template <int BLOCK>
__device__ __forceinline__ void block_minmax(int32_t v[4], int32_t* red) {
    // v: lb.x, lb.y (min), ub.x, ub.y (max) as ordered ints
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    wave_minmax(v);
    if (lane == 0) {
        red[wid * 4 + 0] = v[0];
        red[wid * 4 + 1] = v[1];
        red[wid * 4 + 2] = v[2];
        red[wid * 4 + 3] = v[3];
    }
    lds_barrier();
    if (threadIdx.x == 0) {
        for (int w = 1; w < BLOCK / 64; ++w) {
            v[0] = min(v[0], red[w * 4 + 0]);
            v[1] = min(v[1], red[w * 4 + 1]);
            v[2] = max(v[2], red[w * 4 + 2]);
            v[3] = max(v[3], red[w * 4 + 3]);
        }
    }
}

}  // namespace fw
