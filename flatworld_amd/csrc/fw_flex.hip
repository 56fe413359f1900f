// fw_flex.hip -- FEM-FEM penalty contact between the FEM bodies of one loop:
// the spatial hash over their boundary segments (flatworld/spatialmanager.py,
// flatworld/femspringmanager.py:421-487, 1035-1073) and the batched contact pass
// (flatworld/mixedcontact.py:886-1025).  The hash build and the contribution
// lists take one workgroup per contact group; the work-item pass spreads the work
// items of every group over the GPU, two lanes per item.
//
// The reference runs both with atomics (the hash scatter, the Fext updates) and
// with one query buffer shared by every thread.  Here every result that order
// decides equals the reference's kernels run serially in thread order: a hash
// cell lists its segments in ascending order, every work item walks its own
// query, and a node's contributions are listed in ascending work item.  The
// substep kernel adds each node's list through the scene's FEM-FEM ground record,
// which sits after the plane records and before the height-field and voxel
// records, as the reference launches its contact kernels
// (flatworld/mixedcontact.py:1654-1754).
//
// Within a per-group kernel, phases hand data between threads of the workgroup
// through global memory behind __syncthreads(), whose workgroup-scope release and
// acquire fences cover global memory: every wave of a workgroup runs on one CU.
// Integer atomics only count and reserve places; every order-sensitive result is
// then sorted.
#include <hip/hip_runtime.h>

#include <climits>
#include <mutex>

#include "flatworld_fem.h"
#include "fw_types.h"

namespace fw {
namespace {

constexpr int FLEX_BLOCK = 256;  // threads per workgroup; per-group loops stride over segments, cells and slots
constexpr uint32_t FLEX_NO_CONTACT = 2u;  // FW_STEP_NO_CONTACT

__device__ __forceinline__ float dot2(float ax, float ay, float bx, float by) { return (0.0f + ax * bx) + ay * by; }
__device__ __forceinline__ float len2(float x, float y) { return sqrtf(dot2(x, y, x, y)); }
__device__ __forceinline__ float fmin_(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float fmax_(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ int imin_(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int imax_(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ int mod3_(int k) {
    const int r = k % 3;
    return r < 0 ? r + 3 : r;
}

__device__ __forceinline__ void phase_barrier() { __syncthreads(); }

// A load through a pointer read from a table, made a global-memory load: left to
// itself the compiler cannot tell such a pointer from an LDS one and emits a flat
// load (64-bit address arithmetic, and a wait on the LDS counter as well).
template <typename T>
__device__ __forceinline__ T gld(const T* p) {
    return fw_ld(p);
}
template <typename T>
__device__ __forceinline__ T* gld(T* const* p) {
    return reinterpret_cast<T*>(*(const FW_GLOBAL uint64_t*)p);
}
template <typename T>
__device__ __forceinline__ void gst(T* p, T v) {
    fw_st(p, v);
}

// The per-substep scratch of a group's contact pass: the group's own arrays (scr
// null), or its part of the compact layout at scr (fw_step_noalloc).
struct FlexScratch {
    int2* res_i;
    float4* res_f;
    int32_t* slot_start;
    int32_t* slot_cur;
    int32_t* ent_key;
    float2* ent_f;
    int32_t* list_total;
};

__device__ __forceinline__ FlexScratch flex_scratch(const FlexDev& F, char* scr) {
    if (!scr) return {F.res_i, F.res_f, F.slot_start, F.slot_cur, F.ent_key, F.ent_f, F.list_total};
    const FlexScratchLayout L = flex_scratch_layout(F.n_items, F.n_slots);
    char* p = scr + F.scr_off;
    return {reinterpret_cast<int2*>(p + L.res_i),        reinterpret_cast<float4*>(p + L.res_f),
            reinterpret_cast<int32_t*>(p + L.slot_start), reinterpret_cast<int32_t*>(p + L.slot_cur),
            reinterpret_cast<int32_t*>(p + L.ent_key),    reinterpret_cast<float2*>(p + L.ent_f),
            reinterpret_cast<int32_t*>(p + L.list_total)};
}

// pointToEdgeContact (flatworld/contact_detection.py:15-61); returns is_inside.
__device__ __forceinline__ bool point_edge(float px, float py, float ax, float ay, float bx, float by, float& pen,
                                           float& nx, float& ny, float& w0, float& w1) {
    const float ex = bx - ax, ey = by - ay;
    const float len = len2(ex, ey);
    bool inside = false;
    pen = 1e9f;
    nx = ny = 0.0f;
    w0 = w1 = 0.0f;
    if (len > 1e-9f) {
        const float dx = ex / len, dy = ey / len;
        nx = dy;
        ny = -dx;
        const float tx = px - ax, ty = py - ay;
        float t = dot2(tx, ty, dx, dy);
        t = t / len;
        if (0.0f <= t && t <= 1.0f) {
            inside = true;
            pen = dot2(tx, ty, nx, ny);
            w0 = 1.0f - t;
            w1 = t;
        }
    }
    return inside;
}

// detectPointToMeshBoundary (flatworld/contact_detection.py:64-90): the
// penetration when the point projects inside within `limit`, else 1e9.
__device__ __forceinline__ float detect(float px, float py, float2 a, float2 b, float limit, float& nx, float& ny,
                                        float& w0, float& w1) {
    float pen, n0, n1;
    const bool inside = point_edge(px, py, a.x, a.y, b.x, b.y, pen, n0, n1, w0, w1);
    if (fabsf(pen) < limit && inside) {
        nx = n0;
        ny = n1;
        return pen;
    }
    nx = ny = 0.0f;
    return 1e9f;
}

struct CellRange {
    int lx, ly, ux, uy;
};

// Cells of a box, clamped to the grid (flatworld/spatialmanager.py:160-175, 276-291).
__device__ __forceinline__ CellRange cell_range(const FlexGrid& g, float lbx, float lby, float ubx, float uby) {
    const float dlx = lbx - g.glb[0], dly = lby - g.glb[1];
    const float dux = ubx - g.glb[0], duy = uby - g.glb[1];
    CellRange r;
    r.lx = imin_(imax_((int)floorf(dlx / g.gs[0]), 0), g.cn[0] - 1);
    r.ly = imin_(imax_((int)floorf(dly / g.gs[1]), 0), g.cn[1] - 1);
    r.ux = imin_(imax_((int)floorf(dux / g.gs[0]), 0), g.cn[0] - 1);
    r.uy = imin_(imax_((int)floorf(duy / g.gs[1]), 0), g.cn[1] - 1);
    return r;
}

// The boundary AABB a body's gate reads at `local`: min / max over the sub-slots of
// buffer ar3 (ordered-int encoded).
__device__ __forceinline__ void scene_aabb(const FlexBody& B, int ar3, float bb[4]) {
    const int kk = B.aabb_k;
    const int32_t* e = B.aabb + ar3 * kk * FW_AABB_SUB_INTS;
    int32_t r[4] = {gld(e + 0), gld(e + 1), gld(e + 2), gld(e + 3)};
    for (int j = 1; j < kk; ++j) {
        const int32_t* f = e + j * FW_AABB_SUB_INTS;
        r[0] = imin_(r[0], gld(f + 0));
        r[1] = imin_(r[1], gld(f + 1));
        r[2] = imax_(r[2], gld(f + 2));
        r[3] = imax_(r[3], gld(f + 3));
    }
    for (int k = 0; k < 4; ++k) bb[k] = fw_o2f(r[k]);
}

// Inclusive prefix sum over the 64 lanes of a wave: inside each row of 16 lanes
// (row_shr 1, 2, 4, 8), then lane 15 of rows 0 and 2 added to rows 1 and 3 (row_bcast
// 15), then lane 31 added to rows 2 and 3 (row_bcast 31); a lane without a source
// adds 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int dpp_add(int v) {
    return v + __builtin_amdgcn_update_dpp(0, v, CTRL, ROW_MASK, 0xF, false);
}
__device__ __forceinline__ int wave_inclusive_scan(int v) {
    v = dpp_add<0x111, 0xF>(v);
    v = dpp_add<0x112, 0xF>(v);
    v = dpp_add<0x114, 0xF>(v);
    v = dpp_add<0x118, 0xF>(v);
    v = dpp_add<0x142, 0xA>(v);
    v = dpp_add<0x143, 0xC>(v);
    return v;
}

// Workgroup exclusive scan of one value per thread; *total gets the sum.  lds holds
// at least FLEX_BLOCK / 64 ints.
__device__ int block_exclusive_scan(int v, int* lds, int* total) {
    const int tid = threadIdx.x;
    const int lane = tid & 63, wid = tid >> 6;
    const int incl = wave_inclusive_scan(v);
    if (lane == 63) lds[wid] = incl;
    __syncthreads();
    int before = 0, all = 0;
#pragma unroll
    for (int j = 0; j < FLEX_BLOCK / 64; ++j) {
        const int t = lds[j];
        before += j < wid ? t : 0;
        all += t;
    }
    *total = all;
    __syncthreads();
    return before + incl - v;
}

// In-place exclusive scan of a[0, n) (copy, when given, receives the same); returns the sum.
__device__ int scan_global(int32_t* a, int32_t* copy, int n, int* lds) {
    int carry = 0;
    for (int base = 0; base < n; base += FLEX_BLOCK) {
        const int i = base + (int)threadIdx.x;
        const int v = i < n ? a[i] : 0;
        int tot;
        const int ex = block_exclusive_scan(v, lds, &tot);
        if (i < n) {
            a[i] = carry + ex;
            if (copy) copy[i] = carry + ex;
        }
        carry += tot;
    }
    return carry;
}

// populate_fem_spatial_hash + SpatialHashManager.build at the state substep k of
// the launch starts from (velocity buffer 0).
// One workgroup per group: the group at table row `row`, or row blockIdx.x when row < 0.
// With `sched`, a group rebuilds only when one of its scheduled steps is this substep's
// global step (fw_flex_schedule_rebuild), so that a captured frame holds the launch
// whatever the schedule.
__global__ __launch_bounds__(FLEX_BLOCK) void flex_hash_kernel(const FlexDev* groups, const SceneDev* scenes,
                                                               const ArenaCtl* ctl, int row, int k, int sched) {
    const FlexDev& F = groups[row >= 0 ? row : (int)blockIdx.x];
    const int step = ctl->clock + k;
    if (sched) {
        bool go = false;
#pragma unroll
        for (int r = 0; r < FW_FLEX_SCHED; ++r) go |= F.sched[r] == step;
        if (!go) return;
    }
    __shared__ int lds[FLEX_BLOCK];
    __shared__ float sh_bb[4];
    __shared__ int sh_est;
    __shared__ FlexGrid sh_g;
    const int tid = threadIdx.x;
    const int local = step - F.s0;
    const int rd = (local + 1) & 1;
    const int ar3 = mod3_(local);
    const int ns = F.n_segs;

    // global bounds: union of the bodies' AABBs, widened by 1% of the larger extent
    if (tid == 0) {
        float l0 = 1e9f, l1 = 1e9f, u0 = -1e9f, u1 = -1e9f;
        for (int b = 0; b < F.n_bodies; ++b) {
            float bb[4];
            scene_aabb(F.bodies[b], ar3, bb);
            l0 = fmin_(l0, bb[0]);
            l1 = fmin_(l1, bb[1]);
            u0 = fmax_(u0, bb[2]);
            u1 = fmax_(u1, bb[3]);
        }
        const float extent = fmax_(u0 - l0, u1 - l1);
        const float expand = extent * 0.01f + 0.0f;
        sh_bb[0] = l0 - expand;
        sh_bb[1] = l1 - expand;
        sh_bb[2] = u0 + expand;
        sh_bb[3] = u1 + expand;
        sh_est = fw_f2o(1e30f);
    }
    __syncthreads();

    // segment boxes (two nodes) and the smallest box diagonal
    for (int s = tid; s < ns; s += FLEX_BLOCK) {
        const SceneDev* Sb = scenes + F.body_scene[F.seg_body[s]];
        const float2* pos = Sb->pos[rd];
        const int2 sn = F.seg_node[s];
        const float2 a = pos[sn.x], c = pos[sn.y];
        float lb0 = 1e30f, lb1 = 1e30f, ub0 = -1e30f, ub1 = -1e30f;
        lb0 = fmin_(lb0, a.x);
        lb1 = fmin_(lb1, a.y);
        ub0 = fmax_(ub0, a.x);
        ub1 = fmax_(ub1, a.y);
        lb0 = fmin_(lb0, c.x);
        lb1 = fmin_(lb1, c.y);
        ub0 = fmax_(ub0, c.x);
        ub1 = fmax_(ub1, c.y);
        const float4 box = make_float4(lb0 - 0.0f, lb1 - 0.0f, ub0 + 0.0f, ub1 + 0.0f);
        F.seg_box[s] = box;
        atomicMin(&sh_est, fw_f2o(len2(box.z - box.x, box.w - box.y)));
    }
    __syncthreads();

    // _build_grid_kernel (flatworld/spatialmanager.py:217-245); pow(max_cells, 0.5)
    // is the correctly rounded square root
    if (tid == 0) {
        FlexGrid g;
        float cell = fw_o2f(sh_est) * 1.5f;
        const float ex = sh_bb[2] - sh_bb[0], ey = sh_bb[3] - sh_bb[1];
        const float mpd = fmax_(sqrtf((float)F.max_cells), 1.0f);
        cell = fmax_(cell, ex / mpd);
        cell = fmax_(cell, ey / mpd);
        const float cs = cell + 1e-6f;
        const int nx = (int)fmax_(ceilf(ex / cs), 1.0f);
        const int ny = (int)fmax_(ceilf(ey / cs), 1.0f);
        g.glb[0] = sh_bb[0];
        g.glb[1] = sh_bb[1];
        g.gs[0] = ex / (float)nx;
        g.gs[1] = ey / (float)ny;
        g.cn[0] = nx;
        g.cn[1] = ny;
        g.tc = imin_(nx * ny, F.max_cells);
        g.exact = nx * ny <= F.max_cells;
        sh_g = g;
    }
    __syncthreads();
    const FlexGrid g = sh_g;
    if (ns == 0) {
        if (tid == 0) *F.grid = g;
        return;
    }

    // pairs per segment, then their serial index base (pairs past max_pairs are
    // dropped in that order); each body's union of segment cell ranges
    for (int b = tid; b < F.n_bodies; b += FLEX_BLOCK) F.body_cells[b] = make_int4(INT_MAX, INT_MAX, INT_MIN, INT_MIN);
    phase_barrier();
    for (int s = tid; s < ns; s += FLEX_BLOCK) {
        const float4 box = F.seg_box[s];
        const CellRange r = cell_range(g, box.x, box.y, box.z, box.w);
        F.seg_cells[s] = make_int4(r.lx, r.ly, r.ux, r.uy);
        int* bc = reinterpret_cast<int*>(F.body_cells + F.seg_body[s]);
        atomicMin(bc + 0, r.lx);
        atomicMin(bc + 1, r.ly);
        atomicMax(bc + 2, r.ux);
        atomicMax(bc + 3, r.uy);
        int cnt = 0;
        for (int ix = r.lx; ix <= r.ux; ++ix)
            for (int iy = r.ly; iy <= r.uy; ++iy) cnt += ix + iy * g.cn[0] < g.tc;
        F.seg_pairs[s] = cnt;
    }
    for (int c = tid; c < g.tc; c += FLEX_BLOCK) F.cell_end[c] = 0;
    phase_barrier();
    const int total = scan_global(F.seg_pairs, nullptr, ns, lds);
    phase_barrier();
    for (int s = tid; s < ns; s += FLEX_BLOCK) {
        const int4 r = F.seg_cells[s];
        int pi = F.seg_pairs[s];
        for (int ix = r.x; ix <= r.z; ++ix)
            for (int iy = r.y; iy <= r.w; ++iy) {
                const int cid = ix + iy * g.cn[0];
                if (cid < g.tc) {
                    if (pi < F.max_pairs) atomicAdd(&F.cell_end[cid], 1);
                    ++pi;
                }
            }
    }
    phase_barrier();
    scan_global(F.cell_end, F.cell_start, g.tc, lds);
    phase_barrier();
    for (int s = tid; s < ns; s += FLEX_BLOCK) {
        const int4 r = F.seg_cells[s];
        int pi = F.seg_pairs[s];
        for (int ix = r.x; ix <= r.z; ++ix)
            for (int iy = r.y; iy <= r.w; ++iy) {
                const int cid = ix + iy * g.cn[0];
                if (cid < g.tc) {
                    if (pi < F.max_pairs) F.sorted[atomicAdd(&F.cell_end[cid], 1)] = s;
                    ++pi;
                }
            }
    }
    phase_barrier();
    // a cell's entries in ascending segment: the serial scatter order
    for (int c = tid; c < g.tc; c += FLEX_BLOCK) {
        const int b = F.cell_start[c], e = F.cell_end[c];
        for (int i = b + 1; i < e; ++i) {
            const int v = F.sorted[i];
            int j = i - 1;
            while (j >= b && F.sorted[j] > v) {
                F.sorted[j + 1] = F.sorted[j];
                --j;
            }
            F.sorted[j + 1] = v;
        }
        F.cell_rng[c] = make_int2(b, e);
        const int cx = c % g.cn[0], cy = c / g.cn[0];
        for (int i = b; i < e; ++i) {
            const int sg = F.sorted[i];
            const int4 r = F.seg_cells[sg];
            F.cell_ent[i] = sg << 2 | (cx == r.x) << 1 | (cy == r.y);
        }
    }
    if (tid == 0) {
        FlexGrid out = g;
        out.exact = g.exact && total <= F.max_pairs;
        *F.grid = out;
    }
}

// Whether the query walk met segment s in an earlier cell or earlier in this
// cell.  With every segment in every cell of its box (grid.exact), that is
// whether (ix, iy) is not the first cell of the walk inside the segment's box;
// otherwise the walk is repeated up to this entry.
__device__ __forceinline__ bool seen_before(const FlexDev& F, const FlexGrid& g, const CellRange& q, int s, int ix,
                                            int iy, int pos) {
    if (g.exact) {
        const int4 r = F.seg_cells[s];
        return !(ix == imax_(r.x, q.lx) && iy == imax_(r.y, q.ly));
    }
    for (int jx = q.lx; jx <= q.ux; ++jx)
        for (int jy = q.ly; jy <= q.uy; ++jy) {
            const int cid = jx + jy * g.cn[0];
            if (cid < 0 || cid >= g.tc) continue;
            const int2 cr = F.cell_rng[cid];
            for (int p = cr.x; p < cr.y; ++p) {
                if (jx == ix && jy == iy && p == pos) return false;
                if (F.sorted[p] == s) return true;
            }
        }
    return false;
}

constexpr int WALK_BATCH = 8;  // cells whose ranges are loaded together in a query walk
constexpr int FLEX_LPI = 2;    // lanes per work item in the contact pass

// The best candidate so far: the most negative penetration, the earliest walk
// position on ties (key = walk cell << 32 | entry).
struct Best {
    float pen = 1e9f, nx = 0.0f, ny = 0.0f, w0 = 0.0f, w1 = 0.0f;
    int s = -1;
    long long key = 0x7FFFFFFFFFFFFFFFll;
    bool found = false;
    __device__ void take(float p, float x, float y, float a, float c, int sg, long long k) {
        pen = p;
        nx = x;
        ny = y;
        w0 = a;
        w1 = c;
        s = sg;
        key = k;
        found = true;
    }
};

// Walk the query cells lane, lane + stride, ... (x outer, y inner), ranges loaded
// WALK_BATCH at a time, and keep the best qualifying candidate among the segments
// [tlo, thi) (the target body's) met for the first time; `met` counts those first meetings.  A single
// lane (stride 1) stops evaluating once max_query segments were met, as the
// reference's query list stops growing.  The candidate rule is the contact
// kernel's (deepest penetration below 1e-4) or, with NEAREST, the tie
// initialisation's (smallest |penetration|, flatworld/mixedcontact.py:757-762).
template <bool NEAREST>
__device__ void walk(const FlexDev& F, const FlexGrid& g, const CellRange& q, float2 p, int tlo, int thi, const float2* tpos,
                     float qbuf, int lane, int stride, bool exact, int max_query, int& met, Best& best) {
    const int nyq = q.uy - q.ly + 1, ncell = (q.ux - q.lx + 1) * nyq;
    for (int i0 = 0; i0 * stride + lane < ncell; i0 += WALK_BATCH) {
        int2 cr[WALK_BATCH];
#pragma unroll
        for (int u = 0; u < WALK_BATCH; ++u) {
            const int t = (i0 + u) * stride + lane;
            const int cid = (q.lx + t / nyq) + (q.ly + t % nyq) * g.cn[0];
            cr[u] = (t < ncell && cid >= 0 && cid < g.tc) ? gld(F.cell_rng + cid) : make_int2(0, 0);
        }
        for (int u = 0; u < WALK_BATCH; ++u) {
            const int t = (i0 + u) * stride + lane;
            if (t >= ncell) break;
            const int ix = q.lx + t / nyq, iy = q.ly + t % nyq;
            const int e = cr[u].y;
            for (int pp = cr[u].x; pp < e; ++pp) {
                const int ce = gld(F.cell_ent + pp);
                const int s = ce >> 2;
                if (s < tlo || s >= thi) continue;  // a segment of another body
                // first meeting of s in the walk: the first cell of the walk inside the
                // segment's box, i.e. its first column or the walk's, and its first row
                // or the walk's
                const bool seen = exact ? !((ix == q.lx || (ce & 2)) && (iy == q.ly || (ce & 1)))
                                        : seen_before(F, g, q, s, ix, iy, pp);
                if (seen) continue;
                if (stride == 1 && met >= max_query) continue;
                ++met;
                const int2 sn = gld(F.seg_node + s);
                float nx, ny, w0, w1;
                const float pen = detect(p.x, p.y, gld(tpos + sn.x), gld(tpos + sn.y), qbuf, nx, ny, w0, w1);
                const bool better = NEAREST ? (fabsf(pen) < fabsf(best.pen) && fabsf(pen) < qbuf)
                                            : (pen < best.pen && pen < 1e-4f && fabsf(pen) < qbuf);
                if (better)
                    best.take(pen, nx, ny, w0, w1, s, ((long long)t << 32) | (long long)pp);
            }
        }
    }
}

// The same candidates as walk() on an exact grid, found from the target body's
// segments instead of the query's cells: with every segment listed in every cell of
// its (clamped) cell box, the walk meets segment s exactly when that box meets the
// query's cells, first at the cell (max(box.lx, q.lx), max(box.ly, q.ly)), walk
// position t, and inside a cell the entries come in ascending segment, so the key
// (t, s) orders candidates as the walk's (t, entry) does.  Lanes take every
// stride-th segment of the body; `met` counts the segments met.
__device__ void scan_segments(const FlexDev& F, const CellRange& q, float2 p, int be_off, int be_cnt,
                              const float2* tpos, float qbuf, int lane, int stride, int& met, Best& best) {
    const int nyq = q.uy - q.ly + 1;
    constexpr int SB = 4;  // cell boxes loaded together
    for (int j0 = lane; j0 < be_cnt; j0 += SB * stride) {
        int4 rb[SB];
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            const int j = j0 + u * stride;
            rb[u] = j < be_cnt ? gld(F.seg_cells + be_off + j) : make_int4(1, 1, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            const int4 r = rb[u];
            if (r.x > q.ux || r.z < q.lx || r.y > q.uy || r.w < q.ly) continue;
            const int s = be_off + j0 + u * stride;
            ++met;
            const int t = (imax_(r.x, q.lx) - q.lx) * nyq + (imax_(r.y, q.ly) - q.ly);
            const long long key = ((long long)t << 32) | (long long)s;
            const int2 sn = gld(F.seg_node + s);
            float nx, ny, w0, w1;
            const float pen = detect(p.x, p.y, gld(tpos + sn.x), gld(tpos + sn.y), qbuf, nx, ny, w0, w1);
            if (pen < 1e-4f && fabsf(pen) < qbuf && (pen < best.pen || (pen == best.pen && key < best.key)))
                best.take(pen, nx, ny, w0, w1, s, key);
        }
    }
}

// Pair activation: the BVH overlap test (flatworld/explicitloop.py:477-487), or
// every pair when the loop skips its BVH.
__device__ __forceinline__ bool pair_active(const FlexDev& F, const SceneDev* scenes, int q, int local) {
    if (!F.gated) return true;
    const bool pad = ((local - scenes[F.body_scene[0]].hot.gate_base) % 100) == 0;
    const int ar3 = mod3_(local);
    const int2 pb = gld(F.pair_body + q);
    float a[4], b[4];
    scene_aabb(F.bodies[pb.x], ar3, a);
    scene_aabb(F.bodies[pb.y], ar3, b);
    if (pad) {
        for (int c = 0; c < 2; ++c) {
            const float ba = 0.1f * (a[c + 2] - a[c]);
            const float la = a[c] - ba, ua = a[c + 2] + ba;
            const float bb = 0.1f * (b[c + 2] - b[c]);
            const float lb = b[c] - bb, ub = b[c + 2] + bb;
            a[c] = la;
            a[c + 2] = ua;
            b[c] = lb;
            b[c + 2] = ub;
        }
    }
    return a[0] <= b[2] && a[2] >= b[0] && a[1] <= b[3] && a[3] >= b[1];
}

// _eval_fixed_tie_point and _fem_fem_tie_elem_normal (flatworld/mixedcontact.py:136-163):
// the weighted point of the edge and its left unit normal, (0, 1) for an edge
// shorter than 1e-10.
__device__ __forceinline__ float2 tie_point(float2 a, float2 c, float w0, float w1) {
    return make_float2((0.0f + a.x * w0) + c.x * w1, (0.0f + a.y * w0) + c.y * w1);
}
__device__ __forceinline__ float2 tie_normal(float2 a, float2 c) {
    const float ex = c.x - a.x, ey = c.y - a.y;
    const float elen = len2(ex, ey);
    float tx = 1.0f, ty = 0.0f;
    if (elen > 1e-10f) {
        tx = ex / elen;
        ty = ey / elen;
    }
    return make_float2(-ty, tx);
}

// _initialize_fem_fem_ties_once_kernel (flatworld/mixedcontact.py:687-770) at the
// state the next substep starts from: every tied, unresolved work item of group
// row `row` takes the edge of its target body nearest to it (smallest
// |penetration| within the query buffer) and keeps the edge, the weights and the
// normal gap.  One lane per work item, each with its own query walk.
__global__ __launch_bounds__(FLEX_BLOCK) void flex_ties_kernel(const FlexDev* groups, int row, const SceneDev* scenes,
                                                               const ArenaCtl* ctl) {
    const FlexDev& F = groups[row];
    const int w = (int)(blockIdx.x * FLEX_BLOCK + threadIdx.x);
    if (w >= F.n_items || !F.item_tied[w] || F.tie_seg[w] >= 0) return;
    const SceneDev* S0 = scenes + F.body_scene[0];
    const int rd = (ctl->clock - S0->s0 + 1) & 1;
    const FlexGrid g = *F.grid;
    const float qbuf = fmax_(fmax_(1e-4f, fmax_(g.gs[0], g.gs[1])), 0.001f);
    const float2 p = F.bodies[F.item_body[w]].pos[rd][F.item_node[w]];
    const int tb = F.item_target[w];
    const float2* tpos = F.bodies[tb].pos[rd];
    if (g.tc <= 0) return;
    const CellRange q = cell_range(g, p.x - qbuf, p.y - qbuf, p.x + qbuf, p.y + qbuf);
    int met = 0;
    Best b;
    walk<true>(F, g, q, p, F.body_seg_off[tb], F.body_seg_off[tb + 1], tpos, qbuf, 0, 1, g.exact, F.max_query, met, b);
    if (!b.found) return;
    const int2 sn = F.seg_node[b.s];
    const float2 a = tpos[sn.x], c = tpos[sn.y];
    const float2 cp = tie_point(a, c, b.w0, b.w1);
    const float2 n = tie_normal(a, c);
    F.tie_seg[w] = b.s;
    F.tie_wg[w] = make_float4(b.w0, b.w1, dot2(p.x - cp.x, p.y - cp.y, n.x, n.y), 0.0f);
}

// The substep a group's contact pass serves, or -1 when it has none this launch.
__device__ __forceinline__ int group_local(const FlexDev& F, const ArenaCtl* ctl, int k, uint32_t flags) {
    const int local = ctl->clock + k - F.s0;
    if (local < 0 || local >= F.horizon || (flags & FLEX_NO_CONTACT)) return -1;
    return local;
}

// _batched_flexflex_contact_kernel_sh (flatworld/mixedcontact.py:886-1025), one
// work item per FLEX_LPI lanes over the work items of every group (item_off:
// each group's first global item, n_groups + 1 entries), so that the items of a
// group with long queries spread over many workgroups.  The lanes of an item
// split its query walk and reduce; lane 0 does its stores and counts its
// contributions per slot.
__global__ __launch_bounds__(FLEX_BLOCK, 4) void flex_items_kernel(const FlexDev* groups, const int2* item_map,
                                                                   int total_items, const SceneDev* scenes,
                                                                   const ArenaCtl* ctl, int k, uint32_t flags,
                                                                   char* scr) {
    const int gw = (int)((blockIdx.x * FLEX_BLOCK + threadIdx.x) / FLEX_LPI);
    const int sub = threadIdx.x % FLEX_LPI;
    if (gw >= total_items) return;
    const int2 gi = item_map[gw];  // (group row, work item in the group)
    const FlexDev& F = groups[gi.x];
    const int w = gi.y;
    const int local = group_local(F, ctl, k, flags);
    if (local < 0) return;
    const FlexScratch X = flex_scratch(F, scr);
    const int rd = (local + 1) & 1;
    int2 ri = make_int2(0, -1);
    if (F.n_tied > 0 && gld(F.item_tied + w)) {
        // _apply_fem_fem_ties_fixed_kernel (flatworld/mixedcontact.py:773-811): a
        // spring toward the resolved edge point at the initial normal gap, every
        // substep, whatever the pair gate; tied items never slide
        const int ts = gld(F.tie_seg + w);
        if (sub != 0) return;
        if (ts >= 0) {
            const float2 x = gld(gld(&F.bodies[gld(F.item_body + w)].pos[rd]) + gld(F.item_node + w));
            const float2* tpos = gld(&F.bodies[gld(F.item_target + w)].pos[rd]);
            const int2 sn = gld(F.seg_node + ts);
            const float2 a = gld(tpos + sn.x), c = gld(tpos + sn.y);
            const float4 wg = gld(F.tie_wg + w);
            const float2 cp = tie_point(a, c, wg.x, wg.y);
            const float2 n = tie_normal(a, c);
            const float k = gld(F.item_penalty + w);
            const float gx = cp.x + wg.z * n.x, gy = cp.y + wg.z * n.y;
            gst(X.res_f + w, make_float4(k * (gx - x.x), k * (gy - x.y), wg.x, wg.y));
            const int2 ss = gld(F.seg_slot + ts);
            atomicAdd(&X.slot_start[gld(F.item_slot + w)], 1);
            atomicAdd(&X.slot_start[ss.x], 1);
            atomicAdd(&X.slot_start[ss.y], 1);
            ri = make_int2(1, ts);
        }
        gst(X.res_i + w, ri);
        return;
    }
    const FlexGrid g = *F.grid;
    const float qbuf = fmax_(fmax_(1e-4f, fmax_(g.gs[0], g.gs[1])), 0.016f);
    if (pair_active(F, scenes, gld(F.item_pair + w), local)) {
        const FlexBody* SB = F.bodies + gld(F.item_body + w);
        const int nid = gld(F.item_node + w);
        const float2 p = gld(gld(&SB->pos[rd]) + nid);
        const float2 vel = gld(gld(&SB->vel) + nid);
        const int tb = gld(F.item_target + w);
        const FlexBody* ST = F.bodies + tb;
        const float2* tpos = gld(&ST->pos[rd]);
        const int be_off = gld(F.body_seg_off + tb), be_cnt = gld(F.body_seg_off + tb + 1) - be_off;
        const float penalty = gld(F.item_penalty + w), friction = gld(F.item_friction + w);

        Best b;
        const int cached = gld(F.cache + w);
        if (cached >= be_off && cached < be_off + be_cnt) {
            const int2 sn = gld(F.seg_node + cached);
            const float2 a = gld(tpos + sn.x), c = gld(tpos + sn.y);
            float lb0 = 1e30f, lb1 = 1e30f, ub0 = -1e30f, ub1 = -1e30f;
            lb0 = fmin_(lb0, a.x);
            lb1 = fmin_(lb1, a.y);
            ub0 = fmax_(ub0, a.x);
            ub1 = fmax_(ub1, a.y);
            lb0 = fmin_(lb0, c.x);
            lb1 = fmin_(lb1, c.y);
            ub0 = fmax_(ub0, c.x);
            ub1 = fmax_(ub1, c.y);
            const float margin = fmax_(1e-4f, 1e-2f * len2(ub0 - lb0, ub1 - lb1));
            lb0 = lb0 - margin;
            lb1 = lb1 - margin;
            ub0 = ub0 + margin;
            ub1 = ub1 + margin;
            // _point_outside_aabb (flatworld/mixedcontact.py:113-119)
            const bool outside = fmin_(p.x - lb0, p.y - lb1) < 0.0f || fmin_(ub0 - p.x, ub1 - p.y) < 0.0f;
            if (!outside) {
                float nx, ny, w0, w1;
                const float pen = detect(p.x, p.y, a, c, margin, nx, ny, w0, w1);
                if (pen < 1e-4f && fabsf(pen) < margin) b.take(pen, nx, ny, w0, w1, cached, 0);
            }
        }
        int cache_new = cached;
        // query_point_with_buffer (flatworld/spatialmanager.py:131-199), walked in place; a
        // query whose cells miss every cell of the target body's segments meets none of them
        CellRange q{0, 0, -1, -1};
        bool reach = false;
        if (g.tc > 0) {
            q = cell_range(g, p.x - qbuf, p.y - qbuf, p.x + qbuf, p.y + qbuf);
            const int4 tbc = gld(F.body_cells + tb);
            reach = !(tbc.x > q.ux || tbc.z < q.lx || tbc.y > q.uy || tbc.w < q.ly);
        }
        if (!b.found && reach) {
            bool serial = !g.exact;
            if (g.exact) {
                // the item's lanes take every FLEX_LPI-th cell of the walk; the best
                // candidate is the minimum of (penetration, walk position), the serial
                // walk's choice, as long as no more than max_query distinct segments
                // are met (else the cap decides and the serial walk is repeated)
                int met = 0;
                Best lb;
                // the shorter of the two: the query's cells or the target body's segments
                const int ncell = (q.ux - q.lx + 1) * (q.uy - q.ly + 1);
                if (be_cnt <= 2 * ncell)
                    scan_segments(F, q, p, be_off, be_cnt, tpos, qbuf, sub, FLEX_LPI, met, lb);
                else
                    walk<false>(F, g, q, p, be_off, be_off + be_cnt, tpos, qbuf, sub, FLEX_LPI, true, F.max_query, met, lb);
#pragma unroll
                for (int off = 1; off < FLEX_LPI; off <<= 1) met += __shfl_xor(met, off, FLEX_LPI);
                if (met > F.max_query) {
                    serial = true;
                } else {
                    int bl = sub;
                    float bp = lb.pen;
                    long long bk = lb.key;
#pragma unroll
                    for (int off = 1; off < FLEX_LPI; off <<= 1) {
                        const float op = __shfl_xor(bp, off, FLEX_LPI);
                        const long long ok = __shfl_xor(bk, off, FLEX_LPI);
                        const int ol = __shfl_xor(bl, off, FLEX_LPI);
                        if (op < bp || (op == bp && ok < bk)) {
                            bp = op;
                            bk = ok;
                            bl = ol;
                        }
                    }
                    const int ws = __shfl(lb.s, bl, FLEX_LPI);
                    const float wnx = __shfl(lb.nx, bl, FLEX_LPI), wny = __shfl(lb.ny, bl, FLEX_LPI);
                    const float ww0 = __shfl(lb.w0, bl, FLEX_LPI), ww1 = __shfl(lb.w1, bl, FLEX_LPI);
                    if (ws >= 0) {
                        b.take(bp, wnx, wny, ww0, ww1, ws, bk);
                        cache_new = ws;
                    }
                }
            }
            if (serial) {
                int met = 0;
                Best sb;
                walk<false>(F, g, q, p, be_off, be_off + be_cnt, tpos, qbuf, 0, 1, g.exact, F.max_query, met, sb);
                if (sb.found) {
                    b = sb;
                    cache_new = sb.s;
                }
            }
        }
        if (!b.found) cache_new = -1;
        if (sub == 0) gst(F.cache + w, cache_new);
        if (b.found && b.pen < 0.0f) {
            const float nfx = ((-b.nx) * penalty) * b.pen;
            const float nfy = ((-b.ny) * penalty) * b.pen;
            float tx = nfx, ty = nfy;
            if (friction > 1e-9f) {
                const int2 sn = gld(F.seg_node + b.s);
                const float2* tv = gld(&ST->vel);
                const float2 v0 = gld(tv + sn.x), v1 = gld(tv + sn.y);
                const float svx = v0.x * b.w0 + v1.x * b.w1;
                const float svy = v0.y * b.w0 + v1.y * b.w1;
                const float rvx = vel.x - svx, rvy = vel.y - svy;
                const float d = dot2(rvx, rvy, b.nx, b.ny);
                const float tgx = rvx - d * b.nx, tgy = rvy - d * b.ny;
                const float tlen = len2(tgx, tgy);
                if (tlen > 1e-9f) {
                    const float fdx = (-tgx) / tlen, fdy = (-tgy) / tlen;
                    const float nl = len2(nfx, nfy);
                    tx = tx + (fdx * friction) * nl;
                    ty = ty + (fdy * friction) * nl;
                }
            }
            ri = make_int2(1, b.s);
            if (sub == 0) {
                gst(X.res_f + w, make_float4(tx, ty, b.w0, b.w1));
                const int2 ss = gld(F.seg_slot + b.s);
                atomicAdd(&X.slot_start[gld(F.item_slot + w)], 1);
                atomicAdd(&X.slot_start[ss.x], 1);
                atomicAdd(&X.slot_start[ss.y], 1);
            }
        }
    }
    if (sub == 0) gst(X.res_i + w, ri);
}

// The node-major contribution lists of one group (one workgroup): slot ranges by
// a scan of the counts, entries scattered and sorted by work item, written to
// each body's FEM-FEM block; the counts are then cleared for the next substep.
__global__ __launch_bounds__(FLEX_BLOCK) void flex_lists_kernel(const FlexDev* groups, const SceneDev* scenes,
                                                                const ArenaCtl* ctl, int k, uint32_t flags,
                                                                char* scr) {
    const FlexDev& F = groups[blockIdx.x];
    __shared__ int lds[FLEX_BLOCK];
    const int tid = threadIdx.x;
    if (group_local(F, ctl, k, flags) < 0) return;
    const FlexScratch X = flex_scratch(F, scr);
    const int last = *X.list_total;
    const int total = scan_global(X.slot_start, X.slot_cur, F.n_slots, lds);
    phase_barrier();
    // no entries now or in the previous substep: every node's range already reads
    // zero entries and the counts are zero
    if (total == 0 && last == 0) return;
    for (int w = tid; w < F.n_items; w += FLEX_BLOCK) {
        const int2 ri = X.res_i[w];
        if (!ri.x) continue;
        const float4 rf = X.res_f[w];
        const int2 ss = F.seg_slot[ri.y];
        const int sl[3] = {F.item_slot[w], ss.x, ss.y};
        const float2 val[3] = {make_float2(rf.x, rf.y), make_float2(-(rf.x * rf.z), -(rf.y * rf.z)),
                               make_float2(-(rf.x * rf.w), -(rf.y * rf.w))};
        for (int c = 0; c < 3; ++c) {
            const int at = atomicAdd(&X.slot_cur[sl[c]], 1);
            X.ent_key[at] = (F.n_tied > 0 && F.item_tied[w]) ? w : F.n_items + w;  // tie springs before contacts
            X.ent_f[at] = val[c];
        }
    }
    phase_barrier();
    // each slot's entries in ascending work item, into its body's ranges
    for (int sl = tid; sl < F.n_slots; sl += FLEX_BLOCK) {
        const int b0 = X.slot_start[sl], e = X.slot_cur[sl];
        for (int i = b0 + 1; i < e; ++i) {
            const int kv = X.ent_key[i];
            const float2 fv = X.ent_f[i];
            int j = i - 1;
            while (j >= b0 && X.ent_key[j] > kv) {
                X.ent_key[j + 1] = X.ent_key[j];
                X.ent_f[j + 1] = X.ent_f[j];
                --j;
            }
            X.ent_key[j + 1] = kv;
            X.ent_f[j + 1] = fv;
        }
        const int b = F.slot_body[sl];
        const int base = X.slot_start[F.body_slot_off[b]];
        int2* rng = reinterpret_cast<int2*>(F.body_flex[b]);
        float2* ent = reinterpret_cast<float2*>(rng + F.body_nn[b]);
        rng[F.slot_node[sl]] = make_int2(b0 - base, e - b0);
        for (int i = b0; i < e; ++i) ent[i - base] = X.ent_f[i];
    }
    phase_barrier();
    for (int sl = tid; sl < F.n_slots; sl += FLEX_BLOCK) X.slot_start[sl] = 0;
    if (tid == 0) *X.list_total = total;
}

// Start of a fw_step_noalloc call: the compact scratch holds whatever the last call
// on this GPU left, so every group's slot counts are cleared and its entry total set
// to 1, which makes the first lists pass rewrite every node's range; the group's own
// entry total is set to 1 as well, so that a later fw_step rewrites the ranges this
// call leaves behind.
__global__ __launch_bounds__(FLEX_BLOCK) void flex_scratch_init_kernel(const FlexDev* groups, char* scr) {
    const FlexDev& F = groups[blockIdx.x];
    const FlexScratch X = flex_scratch(F, scr);
    for (int sl = threadIdx.x; sl < F.n_slots; sl += FLEX_BLOCK) X.slot_start[sl] = 0;
    if (threadIdx.x == 0) {
        *X.list_total = 1;
        *F.list_total = 1;
    }
}

}  // namespace

// Scratch of fw_step_noalloc: one fixed region per GPU, shared by every call on it.
static __device__ __align__(256) char scratch[FW_NOALLOC_SCRATCH_BYTES];

}  // namespace fw

// Device address of the no-allocation scratch on `device` (the current device):
// looked up once per device with hipGetSymbolAddress and kept; the array's host-side
// name is only a stand-in and is never passed to a kernel.
extern "C" hipError_t fw_flex_noalloc_scratch(int device, char** out) {
    static std::mutex mu;
    static char* ptrs[64] = {};
    if (device < 0 || device >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!ptrs[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(fw::scratch));
        if (e != hipSuccess) return e;
        ptrs[device] = static_cast<char*>(p);
    }
    *out = ptrs[device];
    return hipSuccess;
}

extern "C" hipError_t fw_launch_flex_scratch_init(const FlexDev* groups, int n_groups, char* scr, hipStream_t stream) {
    if (n_groups <= 0) return hipSuccess;
    hipLaunchKernelGGL(fw::flex_scratch_init_kernel, dim3(n_groups), dim3(fw::FLEX_BLOCK), 0, stream, groups, scr);
    return hipGetLastError();
}

// row >= 0: that group now; row -1: every group now; sched != 0 (row -1): every group
// whose schedule names substep k of the launch.
extern "C" hipError_t fw_launch_flex_hash(const FlexDev* groups, int n_groups, const SceneDev* scenes,
                                          const ArenaCtl* ctl, int row, int k, int sched, hipStream_t stream) {
    if (n_groups <= 0) return hipSuccess;
    hipLaunchKernelGGL(fw::flex_hash_kernel, dim3(row >= 0 ? 1 : n_groups), dim3(fw::FLEX_BLOCK), 0, stream, groups,
                       scenes, ctl, row, k, sched);
    return hipGetLastError();
}

extern "C" hipError_t fw_launch_flex_ties(const FlexDev* groups, int row, int n_items, const SceneDev* scenes,
                                          const ArenaCtl* ctl, hipStream_t stream) {
    if (n_items <= 0) return hipSuccess;
    hipLaunchKernelGGL(fw::flex_ties_kernel, dim3((n_items + fw::FLEX_BLOCK - 1) / fw::FLEX_BLOCK),
                       dim3(fw::FLEX_BLOCK), 0, stream, groups, row, scenes, ctl);
    return hipGetLastError();
}

extern "C" hipError_t fw_launch_flex_contact(const FlexDev* groups, int n_groups, const int2* item_map,
                                             int total_items, const SceneDev* scenes, const ArenaCtl* ctl, int k,
                                             uint32_t flags, char* scr, hipStream_t stream) {
    if (n_groups <= 0) return hipSuccess;
    if (total_items > 0) {
        const int blocks = (int)(((int64_t)total_items * fw::FLEX_LPI + fw::FLEX_BLOCK - 1) / fw::FLEX_BLOCK);
        hipLaunchKernelGGL(fw::flex_items_kernel, dim3(blocks), dim3(fw::FLEX_BLOCK), 0, stream, groups, item_map,
                           total_items, scenes, ctl, k, flags, scr);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(fw::flex_lists_kernel, dim3(n_groups), dim3(fw::FLEX_BLOCK), 0, stream, groups, scenes, ctl, k,
                       flags, scr);
    return hipGetLastError();
}

