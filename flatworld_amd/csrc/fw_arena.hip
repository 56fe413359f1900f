// fw_arena.hip -- host runtime behind include/flatworld_fem.h.
//
// Owns one HBM region per arena and carves one contiguous block per scene out
// of it (first-fit free list).  At scene admission it
//   1. orders nodes for locality (Morton order of the rest positions when the
//      scene needs more than one tile) and cuts them into tiles of at most
//      tile_nodes owned nodes whose incident elements fit tile_slots LDS slots;
//   2. gives every element to the tile holding its lowest-ordered node and
//      stores elements grouped by owner tile, so each tile's owned elements
//      are one contiguous range and only the few elements it shares with
//      neighbours need an index list (halo);
//   3. builds, per tile, the node -> (slot, corner) incidence lists sorted by
//      the caller's element id, which fixes every nodal sum to the order of
//      the reference's serial atomic adds;
//   4. writes the scene block image into pinned memory and copies it with one
//      asynchronous DMA, then launches the prepare kernel.
// Scenes are never serialised behind each other: a substep is one launch over
// the tile table of the whole arena (or, for comparison, one launch per scene).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/flatworld_fem.h"
#include "fw_types.h"

extern "C" hipError_t fw_launch_substep(const SceneDev* scenes, const TileDev* tiles, const ArenaCtl* ctl, int k,
                                        int tile_base, int tile_count, float dt_override, float damping_override,
                                        uint32_t flags, int grid, int block, int slots, int npt, int kind,
                                        hipStream_t stream);
extern "C" hipError_t fw_launch_prepare(const SceneDev* scenes, const TileDev* tiles, int tile_base, int tile_count,
                                        int block, int slots, hipStream_t stream);
extern "C" hipError_t fw_launch_advance_clock(ArenaCtl* ctl, int n, hipStream_t stream);
extern "C" hipError_t fw_substep_occupancy(int block, int slots, int npt, int* per_cu);
extern "C" hipError_t fw_launch_copy_probe(const void* src, void* dst, int64_t n_float4, int grid, hipStream_t stream);
extern "C" hipError_t fw_launch_stream_copy(const void* src, void* dst, int64_t n_float4, int grid, int ilp,
                                            hipStream_t stream);
extern "C" hipError_t fw_launch_mass_scaling(const SceneDev* scenes, const TileDev* tiles, int tile_count, int buf,
                                             float dt_target, int block, int slots, hipStream_t stream);
extern "C" hipError_t fw_launch_flex_hash(const FlexDev* groups, int n_groups, const SceneDev* scenes,
                                          const ArenaCtl* ctl, int row, int k, int sched, hipStream_t stream);
extern "C" hipError_t fw_launch_flex_ties(const FlexDev* groups, int row, int n_items, const SceneDev* scenes,
                                          const ArenaCtl* ctl, hipStream_t stream);
extern "C" hipError_t fw_launch_flex_contact(const FlexDev* groups, int n_groups, const int2* item_map,
                                             int total_items, const SceneDev* scenes, const ArenaCtl* ctl, int k,
                                             uint32_t flags, char* scr, hipStream_t stream);
extern "C" hipError_t fw_flex_noalloc_scratch(int device, char** out);
extern "C" hipError_t fw_launch_rigid_pass(const SceneDev* scenes, const int2* rows, int n_rows, const ArenaCtl* ctl,
                                           int k, uint32_t flags, hipStream_t stream);
extern "C" hipError_t fw_launch_flex_scratch_init(const FlexDev* groups, int n_groups, char* scr, hipStream_t stream);

namespace {

thread_local std::string g_err;
thread_local const char* g_err_static = nullptr;  // a message with static storage (no allocation)

int fail(int code, const std::string& msg) {
    g_err = msg;
    g_err_static = nullptr;
    return code;
}

// fail() for the paths that must not allocate: the message has static storage.
int fail_static(int code, const char* msg) {
    g_err_static = msg;
    return code;
}

#define HIPCHK(expr)                                                                               \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(FW_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));              \
    } while (0)

constexpr int64_t ALIGN = 256;
inline int64_t align_up(int64_t v) { return (v + ALIGN - 1) / ALIGN * ALIGN; }

// ------------------------------------------------------------------ tiling

struct Tiling {
    std::vector<int32_t> node_perm;  // internal -> caller node id
    std::vector<int32_t> node_inv;   // caller -> internal
    std::vector<int32_t> elem_perm;  // internal -> caller element id
    std::vector<int32_t> elem_inv;
    std::vector<TileDev> tiles;      // pointers and scene fields left 0; offsets relative to the scene
    std::vector<uint32_t> rec;       // 2 per slot (one uint2): corner words (local node id | incidence rank)
    std::vector<int32_t> hel;        // per halo slot: storage element index
    std::vector<int32_t> hnode;      // per tile: storage ids of the halo nodes, ascending
    std::vector<uint16_t> csr_ptr;
    std::vector<int32_t> push;       // 2 per halo copy, grouped by source tile: (local owned node, copy index)
    int64_t halo_slots = 0;
};

uint32_t spread16(uint32_t v) {
    v &= 0xFFFF;
    v = (v | (v << 8)) & 0x00FF00FF;
    v = (v | (v << 4)) & 0x0F0F0F0F;
    v = (v | (v << 2)) & 0x33333333;
    v = (v | (v << 1)) & 0x55555555;
    return v;
}

// Corners per element: three for a triangle, two for a spring.
inline int corners_of(const fw_scene_desc* d) { return d->mat_type == FW_MAT_SPRING ? 2 : 3; }

// Temporaries of one tiling, kept per thread: a server prepares a stream of jobs on
// worker threads, and reusing warm buffers spares every job the page faults (and
// the cross-thread unmapping) of a dozen freshly allocated multi-megabyte arrays.
struct TileScratch {
    std::vector<int32_t> inc_ptr, inc, iptr, iinc, cint, stamp, seen, nstamp, nseen, tile_start, node_tile, owner,
        own_count, slot_of, lid, nst, slots, hl, hn, src_cnt;
    std::vector<uint64_t> kv, kv_tmp;
    std::vector<uint32_t> cnt;
    std::vector<uint16_t> rank;
};

TileScratch& tile_scratch() {
    static thread_local TileScratch s;
    return s;
}

// K corners per element, a compile-time constant: the incidence codes K * e + c are
// divided by K in every loop below.
//
// Node-indexed work runs in the internal (Morton) order: incidence lists are
// regrouped by internal node and connectivity is translated to internal node ids
// once, so the tile cut, the halo search and the slot records walk memory in the
// order they are produced instead of jumping through the caller's numbering.
template <int K>
int build_tiling_k(const fw_scene_desc* d, int tile_nodes, int tile_slots, int node_cap, Tiling& T) {
    const int N = d->n_nodes, E = d->n_elems;
    TileScratch& S = tile_scratch();
    // incidence (caller ids): node -> (element, corner), element-major so lists are ascending
    std::vector<int32_t>& inc_ptr = S.inc_ptr;
    inc_ptr.assign(N + 1, 0);
    for (int64_t i = 0; i < K * (int64_t)E; ++i) {
        int32_t n = d->conn[i];
        if (n < 0 || n >= N) return fail(FW_E_INVALID, "connectivity references node " + std::to_string(n) + " outside [0, n_nodes)");
        inc_ptr[n + 1]++;
    }
    for (int i = 0; i < N; ++i) inc_ptr[i + 1] += inc_ptr[i];
    std::vector<int32_t>& inc = S.inc;
    inc.resize(K * (size_t)E);
    {
        std::vector<int32_t>& fillp = S.nseen;  // reused below, re-initialised there
        fillp.assign(inc_ptr.begin(), inc_ptr.end() - 1);
        for (int32_t e = 0; e < E; ++e)
            for (int c = 0; c < K; ++c) inc[fillp[d->conn[K * e + c]]++] = K * e + c;
    }

    // node order
    T.node_perm.resize(N);
    const bool single = N <= tile_nodes && E <= tile_slots && N <= node_cap;
    if (single) {
        for (int i = 0; i < N; ++i) T.node_perm[i] = i;
    } else {
        float lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
        for (int i = 0; i < N; ++i)
            for (int k = 0; k < 2; ++k) {
                float v = d->x0[2 * i + k];
                if (std::isfinite(v)) {
                    lo[k] = std::min(lo[k], v);
                    hi[k] = std::max(hi[k], v);
                }
            }
        double scale[2];
        for (int k = 0; k < 2; ++k) {
            const double span = (double)hi[k] - (double)lo[k];
            scale[k] = span > 0 ? 65535.0 / span : 0.0;
        }
        // (Morton key << 32 | node id), radix-sorted on the key by three stable 11-bit
        // passes: equal keys keep ascending node ids
        std::vector<uint64_t>& kv = S.kv;
        std::vector<uint64_t>& tmp = S.kv_tmp;
        kv.resize(N);
        tmp.resize(N);
        for (int i = 0; i < N; ++i) {
            uint32_t q[2];
            for (int k = 0; k < 2; ++k) {
                double t = ((double)d->x0[2 * i + k] - lo[k]) * scale[k];
                if (!(t >= 0.0)) t = 0.0;
                if (t > 65535.0) t = 65535.0;
                q[k] = (uint32_t)t;
            }
            const uint32_t key = spread16(q[0]) | (spread16(q[1]) << 1);
            kv[i] = ((uint64_t)key << 32) | (uint32_t)i;
        }
        std::vector<uint32_t>& cnt = S.cnt;
        cnt.resize(2049);
        for (int pass = 0; pass < 3; ++pass) {
            const int sh = 32 + 11 * pass;
            std::fill(cnt.begin(), cnt.end(), 0u);
            for (int i = 0; i < N; ++i) cnt[((kv[i] >> sh) & 0x7FFu) + 1]++;
            for (int b = 0; b < 2048; ++b) cnt[b + 1] += cnt[b];
            for (int i = 0; i < N; ++i) tmp[cnt[(kv[i] >> sh) & 0x7FFu]++] = kv[i];
            kv.swap(tmp);
        }
        for (int i = 0; i < N; ++i) T.node_perm[i] = (int32_t)(uint32_t)kv[i];
    }
    T.node_inv.resize(N);
    for (int i = 0; i < N; ++i) T.node_inv[T.node_perm[i]] = i;

    // incidence lists regrouped by internal node, and connectivity in internal node ids
    std::vector<int32_t>& iptr = S.iptr;
    std::vector<int32_t>& iinc = S.iinc;
    iptr.resize(N + 1);
    iinc.resize(K * (size_t)E);
    iptr[0] = 0;
    for (int p = 0; p < N; ++p) {
        const int n = T.node_perm[p];
        const int a = inc_ptr[n], b = inc_ptr[n + 1];
        std::memcpy(iinc.data() + iptr[p], inc.data() + a, 4 * (size_t)(b - a));
        iptr[p + 1] = iptr[p] + (b - a);
    }
    std::vector<int32_t>& cint = S.cint;
    cint.resize(K * (size_t)E);
    for (int64_t i = 0; i < K * (int64_t)E; ++i) cint[i] = T.node_inv[d->conn[i]];

    // greedy tiles over internal node order: at most tile_nodes owned nodes, tile_slots
    // incident elements and node_cap nodes in LDS (owned + nodes of incident elements)
    std::vector<int32_t>& stamp = S.stamp;
    std::vector<int32_t>& seen = S.seen;
    std::vector<int32_t>& nstamp = S.nstamp;
    std::vector<int32_t>& nseen = S.nseen;
    stamp.assign(E, -1);
    seen.assign(E, -1);
    nstamp.assign(N, -1);  // indexed by internal node
    nseen.assign(N, -1);
    std::vector<int32_t>& tile_start = S.tile_start;  // internal node index where each tile starts
    tile_start.clear();
    {
        int tid = 0, cur_nodes = 0, cur_slots = 0, cur_lds = 0;
        if (N > 0) tile_start.push_back(0);
        for (int p = 0; p < N; ++p) {
            // what adding node p costs, relative to the current tile (add*) and to an empty one (all*)
            int add = 0, all = 0, addn = 0, alln = 0;
            if (nseen[p] != p) {
                nseen[p] = p;
                ++alln;
                if (nstamp[p] != tid) ++addn;
            }
            for (int j = iptr[p]; j < iptr[p + 1]; ++j) {
                const int e = iinc[j] / K;
                if (seen[e] == p) continue;
                seen[e] = p;
                ++all;
                if (stamp[e] != tid) ++add;
                for (int c = 0; c < K; ++c) {
                    const int q = cint[K * e + c];
                    if (nseen[q] == p) continue;
                    nseen[q] = p;
                    ++alln;
                    if (nstamp[q] != tid) ++addn;
                }
            }
            if (cur_nodes > 0 &&
                (cur_nodes + 1 > tile_nodes || cur_slots + add > tile_slots || cur_lds + addn > node_cap)) {
                ++tid;
                tile_start.push_back(p);
                cur_nodes = 0;
                cur_slots = 0;
                cur_lds = 0;
                add = all;
                addn = alln;
            }
            if (add > tile_slots || addn > node_cap)
                return fail(FW_E_INVALID, "node " + std::to_string(T.node_perm[p]) + " touches " + std::to_string(add) +
                                              " elements, more than one tile holds");
            nstamp[p] = tid;
            for (int j = iptr[p]; j < iptr[p + 1]; ++j) {
                const int e = iinc[j] / K;
                stamp[e] = tid;
                for (int c = 0; c < K; ++c) nstamp[cint[K * e + c]] = tid;
            }
            cur_nodes += 1;
            cur_slots += add;
            cur_lds += addn;
        }
    }
    const int ntiles = (int)tile_start.size();
    std::vector<int32_t>& node_tile = S.node_tile;
    node_tile.resize(N);
    for (int t = 0; t < ntiles; ++t) {
        int p1 = t + 1 < ntiles ? tile_start[t + 1] : N;
        for (int p = tile_start[t]; p < p1; ++p) node_tile[p] = t;
    }

    // element owner = tile of its lowest internal node; store grouped by owner
    std::vector<int32_t>& owner = S.owner;
    std::vector<int32_t>& own_count = S.own_count;
    owner.resize(E);
    own_count.assign(ntiles + 1, 0);
    for (int e = 0; e < E; ++e) {
        int m = cint[K * e];
        for (int c = 1; c < K; ++c) m = std::min(m, cint[K * e + c]);
        owner[e] = node_tile[m];
        own_count[owner[e] + 1]++;
    }
    for (int t = 0; t < ntiles; ++t) own_count[t + 1] += own_count[t];
    T.elem_perm.resize(E);
    {
        std::vector<int32_t>& fillp = S.slot_of;  // re-initialised below
        fillp.assign(own_count.begin(), own_count.end() - 1);
        for (int e = 0; e < E; ++e) T.elem_perm[fillp[owner[e]]++] = e;  // ascending caller id inside a tile
    }
    T.elem_inv.resize(E);
    for (int i = 0; i < E; ++i) T.elem_inv[T.elem_perm[i]] = i;

    // per tile: slot records, halo node list, node CSR
    T.tiles.resize(ntiles);
    std::fill(stamp.begin(), stamp.end(), -1);
    std::vector<int32_t>& slot_of = S.slot_of;
    std::vector<int32_t>& lid = S.lid;  // by internal node
    std::vector<int32_t>& nst = S.nst;  // by internal node
    slot_of.assign(E, -1);
    lid.assign(N, -1);
    nst.assign(N, -1);
    // scratch reused by every tile, and the outputs sized for a typical halo up front
    std::vector<int32_t>& slots = S.slots;
    std::vector<int32_t>& hl = S.hl;
    std::vector<int32_t>& hn = S.hn;
    std::vector<uint16_t>& rank = S.rank;
    T.rec.clear();
    T.hel.clear();
    T.csr_ptr.clear();
    T.hnode.clear();
    T.rec.reserve(2 * ((size_t)E + E / 4 + 64));
    T.hel.reserve((size_t)E / 4 + 64);
    T.csr_ptr.reserve((size_t)N + ntiles);
    T.hnode.reserve((size_t)N / 4 + 64);
    T.halo_slots = 0;
    const int max_rank = K == 3 ? FW_REC_MAX_RANK_TRI : FW_REC_MAX_RANK;
    for (int t = 0; t < ntiles; ++t) {
        TileDev& td = T.tiles[t];
        std::memset(&td, 0, sizeof(td));
        const int p0 = tile_start[t], p1 = t + 1 < ntiles ? tile_start[t + 1] : N;
        td.node0 = p0;
        td.n_nodes = p1 - p0;
        td.elem0 = own_count[t];
        td.n_own = own_count[t + 1] - own_count[t];
        {
            // AABB sub-slot: tiles share one only when the scene has more tiles than sub-slots
            const int kk = std::min(ntiles, FW_AABB_MAX_SUB);
            td.aabb_slot = t % kk;
            td.aabb_mode = ntiles <= kk ? 0 : (t < kk ? 2 : 1);
        }
        // slots: owned elements in storage order, then halo elements (ascending storage index)
        slots.clear();  // internal element index per slot
        for (int i = td.elem0; i < td.elem0 + td.n_own; ++i) slots.push_back(i);
        hl.clear();
        for (int j = iptr[p0]; j < iptr[p1]; ++j) {
            int e = iinc[j] / K;
            if (owner[e] != t && stamp[e] != t) {
                stamp[e] = t;
                hl.push_back(T.elem_inv[e]);
            }
        }
        std::sort(hl.begin(), hl.end());
        slots.insert(slots.end(), hl.begin(), hl.end());
        td.n_halo = (int32_t)hl.size();
        T.halo_slots += td.n_halo;
        const int nslots = (int)slots.size();
        if (nslots > tile_slots) return fail(FW_E_INVALID, "internal tiling error: slot overflow");
        for (int s = 0; s < nslots; ++s) slot_of[T.elem_perm[slots[s]]] = s;
        // local node ids: owned nodes first (p - p0), then halo nodes in ascending storage order
        for (int p = p0; p < p1; ++p) lid[p] = p - p0;
        hn.clear();
        for (int s = 0; s < nslots; ++s) {
            const int e = T.elem_perm[slots[s]];
            for (int c = 0; c < K; ++c) {
                const int q = cint[K * e + c];
                if ((q < p0 || q >= p1) && nst[q] != t) {
                    nst[q] = t;
                    hn.push_back(q);
                }
            }
        }
        std::sort(hn.begin(), hn.end());
        for (size_t k = 0; k < hn.size(); ++k) lid[hn[k]] = td.n_nodes + (int)k;
        if (td.n_nodes + (int)hn.size() > node_cap) return fail(FW_E_INVALID, "internal tiling error: node overflow");
        td.hnode_off = (int32_t)T.hnode.size();
        td.n_hnodes = (int32_t)hn.size();
        T.hnode.insert(T.hnode.end(), hn.begin(), hn.end());
        // node-major incidence positions: node i of the tile owns LDS entries
        // [ptr[i], ptr[i+1]) in ascending (caller element id, corner); a corner's
        // rank is its position in its node's list
        rank.assign(K * (size_t)nslots, 0);
        td.ptr_off = (int32_t)T.csr_ptr.size();
        int cnt = 0;
        T.csr_ptr.push_back(0);
        for (int p = p0; p < p1; ++p) {
            if (iptr[p + 1] - iptr[p] > max_rank + 1)
                return fail(FW_E_UNSUPPORTED, "node " + std::to_string(T.node_perm[p]) + " has " +
                                                  std::to_string(iptr[p + 1] - iptr[p]) + " element corners; the substep kernel "
                                                  "supports at most " + std::to_string(max_rank + 1));
            for (int j = iptr[p]; j < iptr[p + 1]; ++j) {  // ascending caller element id, then corner
                int e = iinc[j] / K, c = iinc[j] % K;
                rank[K * (size_t)slot_of[e] + c] = (uint16_t)(j - iptr[p]);
                ++cnt;
            }
            if (cnt > 65535) return fail(FW_E_INVALID, "internal tiling error: CSR overflow");
            T.csr_ptr.push_back((uint16_t)cnt);
        }
        // slot records: corner words (local id | rank << FW_REC_ID_BITS), for a spring the
        // rest length's bits in place of the third word; halo slots' storage element
        // indices in their own list
        td.rec_off = (int32_t)(T.rec.size() / 2);
        T.rec.resize(T.rec.size() + 2 * (size_t)nslots);
        uint32_t* rec = T.rec.data() + 2 * (size_t)td.rec_off;
        for (int s = 0; s < nslots; ++s) {
            const int e = T.elem_perm[slots[s]];
            uint32_t w[3] = {0u, 0u, 0u};
            uint32_t hi = 0u;  // triangles: rank bits 5..9 of each corner
            for (int c = 0; c < K; ++c) {
                const uint32_t l = (uint32_t)lid[cint[K * e + c]];
                if (l > FW_REC_ID_MASK) return fail(FW_E_INVALID, "internal tiling error: local id overflow");
                const uint32_t rk = rank[K * (size_t)s + c];
                w[c] = l | ((rk & 31u) << FW_REC_ID_BITS);
                hi |= ((rk >> 5) & 31u) << (16 + 5 * c);
            }
            if (K == 2) std::memcpy(&w[2], d->rest_len + e, 4);
            rec[2 * s + 0] = w[0] | (w[1] << 16);
            rec[2 * s + 1] = K == 3 ? (w[2] | hi) : w[2];
        }
        td.hel_off = (int32_t)T.hel.size();
        for (int s = td.n_own; s < nslots; ++s) T.hel.push_back(slots[s]);
    }
    // push lists: every halo copy is written by the tile owning its node, right after that
    // tile computes the node's new position; grouped by that tile, in (reading tile,
    // halo copy) order
    std::vector<int32_t>& src_cnt = S.src_cnt;
    src_cnt.assign(ntiles + 1, 0);
    for (size_t j = 0; j < T.hnode.size(); ++j) src_cnt[node_tile[T.hnode[j]] + 1]++;
    for (int t = 0; t < ntiles; ++t) src_cnt[t + 1] += src_cnt[t];
    T.push.resize(2 * T.hnode.size());
    for (int t = 0; t < ntiles; ++t) {
        T.tiles[t].push_off = src_cnt[t];
        T.tiles[t].n_push = src_cnt[t + 1] - src_cnt[t];
    }
    for (int t = 0; t < ntiles; ++t) {
        const TileDev& td = T.tiles[t];
        for (int h = 0; h < td.n_hnodes; ++h) {
            const int q = T.hnode[td.hnode_off + h];
            const int o = node_tile[q];
            const int at = src_cnt[o]++;
            T.push[2 * (size_t)at] = q - T.tiles[o].node0;
            T.push[2 * (size_t)at + 1] = td.hnode_off + h;
        }
    }
    return FW_OK;
}

int build_tiling(const fw_scene_desc* d, int tile_nodes, int tile_slots, int node_cap, Tiling& T) {
    return corners_of(d) == 2 ? build_tiling_k<2>(d, tile_nodes, tile_slots, node_cap, T)
                              : build_tiling_k<3>(d, tile_nodes, tile_slots, node_cap, T);
}

// Height-field block: nx float heights, then (8-byte aligned) nx float2 unit normals.
// Tilings of prepared scenes, kept by connectivity.  A tiling is valid for a scene
// exactly when the scene has the tiling's connectivity (the slot, LDS and rank limits
// depend on it alone) and, for springs, its rest lengths (they ride in the slot
// records); the node positions only choose the Morton order, i.e. which valid tiling
// is built.  Results do not depend on the tiling, so a job whose connectivity was seen
// before reuses that tiling and skips the tile cut.  Keys are compared in full; the
// least recently used entries go once the cache holds more than TC_MAX_BYTES.
struct TilingKey {
    uint64_t hash;
    int32_t n_nodes, n_elems, corners, tile_nodes, tile_slots, node_cap;
    std::vector<int32_t> conn;
    std::vector<float> rest;
};
struct TilingCacheEntry {
    std::shared_ptr<const TilingKey> key;
    std::shared_ptr<const Tiling> T;
    int64_t bytes;
    uint64_t used;
};
constexpr int64_t TC_MAX_BYTES = int64_t(512) << 20;
std::mutex g_tc_mu;
std::vector<TilingCacheEntry> g_tc;
int64_t g_tc_bytes = 0, g_tc_hits = 0, g_tc_misses = 0;
uint64_t g_tc_clock = 0;

int64_t tiling_bytes(const Tiling& T) {
    return 4 * (int64_t)(T.node_perm.size() + T.node_inv.size() + T.elem_perm.size() + T.elem_inv.size() +
                         T.rec.size() + T.hel.size() + T.hnode.size() + T.push.size()) +
           (int64_t)sizeof(TileDev) * (int64_t)T.tiles.size() + 2 * (int64_t)T.csr_ptr.size();
}

uint64_t tiling_key_hash(const fw_scene_desc* d) {
    uint64_t h = 1469598103934665603ull;  // FNV-1a over the connectivity words (and rest lengths)
    const int64_t n = corners_of(d) * (int64_t)d->n_elems;
    for (int64_t i = 0; i < n; ++i) h = (h ^ (uint32_t)d->conn[i]) * 1099511628211ull;
    if (d->mat_type == FW_MAT_SPRING)
        for (int64_t i = 0; i < d->n_elems; ++i) {
            uint32_t b;
            std::memcpy(&b, d->rest_len + i, 4);
            h = (h ^ b) * 1099511628211ull;
        }
    return h;
}

// The key's scalar fields (cheap, under the lock) and its arrays (outside it).
bool tiling_key_head(const TilingKey& k, uint64_t h, const fw_scene_desc* d, int tn, int ts, int cap) {
    return k.hash == h && k.n_nodes == d->n_nodes && k.n_elems == d->n_elems && k.corners == corners_of(d) &&
           k.tile_nodes == tn && k.tile_slots == ts && k.node_cap == cap;
}
bool tiling_key_arrays(const TilingKey& k, const fw_scene_desc* d) {
    if (std::memcmp(k.conn.data(), d->conn, sizeof(int32_t) * k.corners * (size_t)d->n_elems) != 0) return false;
    return k.corners != 2 || std::memcmp(k.rest.data(), d->rest_len, sizeof(float) * (size_t)d->n_elems) == 0;
}

// The tiling of d for shape (tn, ts, cap): a copy of a kept one, or built and kept.
int cached_tiling(const fw_scene_desc* d, int tn, int ts, int cap, Tiling& out) {
    const uint64_t h = tiling_key_hash(d);
    std::vector<TilingCacheEntry> cand;
    {
        std::lock_guard<std::mutex> lock(g_tc_mu);
        for (const auto& c : g_tc)
            if (tiling_key_head(*c.key, h, d, tn, ts, cap)) cand.push_back(c);
    }
    for (const auto& c : cand)
        if (tiling_key_arrays(*c.key, d)) {
            out = *c.T;
            std::lock_guard<std::mutex> lock(g_tc_mu);
            ++g_tc_hits;
            for (auto& e : g_tc)
                if (e.T == c.T) e.used = ++g_tc_clock;
            return FW_OK;
        }
    const int rc = build_tiling(d, tn, ts, cap, out);
    if (rc) return rc;
    auto key = std::make_shared<TilingKey>();
    key->hash = h;
    key->n_nodes = d->n_nodes;
    key->n_elems = d->n_elems;
    key->corners = corners_of(d);
    key->tile_nodes = tn;
    key->tile_slots = ts;
    key->node_cap = cap;
    key->conn.assign(d->conn, d->conn + key->corners * (size_t)d->n_elems);
    if (key->corners == 2) key->rest.assign(d->rest_len, d->rest_len + d->n_elems);
    TilingCacheEntry e;
    e.bytes = tiling_bytes(out) + 4 * (int64_t)(key->conn.size() + key->rest.size());
    e.key = std::move(key);
    e.T = std::make_shared<const Tiling>(out);
    std::lock_guard<std::mutex> lock(g_tc_mu);
    ++g_tc_misses;
    if (e.bytes > TC_MAX_BYTES) return FW_OK;
    while (!g_tc.empty() && g_tc_bytes + e.bytes > TC_MAX_BYTES) {
        size_t lru = 0;
        for (size_t i = 1; i < g_tc.size(); ++i)
            if (g_tc[i].used < g_tc[lru].used) lru = i;
        g_tc_bytes -= g_tc[lru].bytes;
        g_tc.erase(g_tc.begin() + (std::ptrdiff_t)lru);
    }
    e.used = ++g_tc_clock;
    g_tc_bytes += e.bytes;
    g_tc.push_back(std::move(e));
    return FW_OK;
}

int64_t hf_normals_offset(int nx) { return (4 * (int64_t)nx + 7) / 8 * 8; }
int64_t hf_bytes(int nx) { return hf_normals_offset(nx) + 8 * (int64_t)nx; }

// The height-field query's normal (flatworld/mixedcontact.py:193-214) depends on the
// query point only through the rounded sample index ir: a central difference over
// the clamped samples ir - 1, ir + 1, guarded by denom > 1e-6, negated when
// `reverse`, normalised when its length exceeds 1e-12.  Evaluated here once per
// sample with the same float32 operations (no contraction, IEEE division and square
// root), so the device looks the normal up instead of recomputing it per node.
void build_hf_normals(const GroundDev& G, const float* h, float2* nrm) {
    const int nx = G.nx;
    for (int ir = 0; ir < nx; ++ir) {
        int j0 = ir - 1, j1 = ir + 1;
        if (j0 < 0) j0 = 0;
        if (j1 >= nx) j1 = nx - 1;
        volatile float denom = G.hf_dxw * (float)(j1 - j0);
        float dhdx = 0.0f;
        if (denom > 1e-6f) dhdx = (h[j1] - h[j0]) / denom;
        float n0 = -dhdx, n1 = 1.0f;
        if (G.reverse == 1) {
            n0 = -n0;
            n1 = -n1;
        }
        volatile float sq = n0 * n0;
        volatile float nlen = std::sqrt(sq + n1 * n1);
        if (nlen > 1e-12f) {
            n0 = n0 / nlen;
            n1 = n1 / nlen;
        }
        nrm[ir] = make_float2(n0, n1);
    }
}

// Height beyond which a node cannot touch a height field: the interpolated height
// h0 + (h1 - h0) t (t in [0, 1]) stays within (|highest| + |lowest|) 2^-22 of the
// sample range in float32, and penetration needs z < h (z > h for a reversed field,
// whose normals point down), so a node above the returned level (below it, reversed)
// has a signed distance >= 0 or NaN: no contact.  The device skips the query there.
float hf_skip_level(const float* h, int nx, bool reversed) {
    double hi = -INFINITY, lo = INFINITY;
    for (int i = 0; i < nx; ++i) {
        if (h[i] > hi) hi = h[i];
        if (h[i] < lo) lo = h[i];
    }
    if (!(hi >= lo) || std::isinf(hi) || std::isinf(lo)) return reversed ? -INFINITY : INFINITY;
    const double m = (std::fabs(hi) + std::fabs(lo)) * std::ldexp(1.0, -22) + 1e-37;
    if (reversed) return std::nextafter((float)(lo - m), -INFINITY);
    return std::nextafter((float)(hi + m), INFINITY);
}

// ------------------------------------------------------------------ scene layout

// A mesh rigid record's data block (offsets from the block start): the rigid pass's
// float4 per node and the boundary-node ids as for a primitive, then the MeshDev
// header, the per-item state and the mesh arrays at the capacities reserved at
// admission.
struct MeshLayout {
    int64_t hdr, cache, skip, fric_i, fric_f, fric_pen, tie;
    int64_t coords, conn, elb, eub, margin, reg, cell, sorted;
    int64_t total;
};

MeshLayout mesh_layout(int64_t N, const fw_ground_desc& gd) {
    MeshLayout M;
    int64_t o = (20 * N + 15) / 16 * 16;
    auto take = [&](int64_t bytes) {
        int64_t r = o;
        o += (bytes + 15) / 16 * 16;
        return r;
    };
    M.hdr = take((int64_t)sizeof(MeshDev));
    M.cache = take(8 * N);
    M.skip = take(4 * N);
    M.fric_i = take(8 * N);
    M.fric_f = take(16 * N);
    M.fric_pen = take(4 * N);
    M.tie = take(16 * N);
    M.coords = take(8 * (int64_t)gd.mesh_nodes);
    M.conn = take(8 * (int64_t)gd.mesh_edges);
    M.elb = take(8 * (int64_t)gd.mesh_edges);
    M.eub = take(8 * (int64_t)gd.mesh_edges);
    M.margin = take(4 * (int64_t)gd.mesh_edges);
    M.reg = take(8 * (int64_t)std::max(gd.mesh_cap_reg, 1));
    M.cell = take(8 * (int64_t)std::max(gd.mesh_cap_cells, 1));
    M.sorted = take(4 * (int64_t)std::max(gd.mesh_cap_pairs, 1));
    M.total = o;
    return M;
}

// What the host keeps of a scene's mesh rigid record: its layout, counts and capacities.
struct MeshHost {
    bool used = false;
    MeshLayout ml{};
    int32_t nodes = 0, edges = 0, cap_reg = 0, cap_cells = 0, cap_pairs = 0, n_items = 0;
};

MeshDev mesh_header(const fw_ground_desc& gd, const MeshLayout& M, int32_t n_items) {
    MeshDev H;
    std::memset(&H, 0, sizeof(H));
    for (int k = 0; k < 2; ++k) {
        H.glb[k] = gd.mesh_glb[k];
        H.gs[k] = gd.mesh_gs[k];
        H.cn[k] = gd.mesh_cn[k];
    }
    H.tc = gd.mesh_tc;
    H.did = gd.mesh_did;
    H.subcycle = gd.mesh_subcycle;
    H.reset_at = gd.mesh_reset_at;
    H.n_items = n_items;
    H.o_coords = M.coords;
    H.o_conn = M.conn;
    H.o_elb = M.elb;
    H.o_eub = M.eub;
    H.o_margin = M.margin;
    H.o_reg = M.reg;
    H.o_cell = M.cell;
    H.o_sorted = M.sorted;
    H.o_cache = M.cache;
    H.o_skip = M.skip;
    H.o_fric_i = M.fric_i;
    H.o_fric_f = M.fric_f;
    H.o_fric_pen = M.fric_pen;
    H.o_tie = M.tie;
    return H;
}

struct Layout {
    int64_t pos0, pos1, rest, vel, mass, flag, bcv, disp, force, cflag;
    int64_t jinv, area, j2a, j2b, stress, strain, yflag;
    int64_t rec, hel, hpos0, hpos1, hrest, hvel0, hvel1, push, cptr, tiles, aabb;
    int32_t aabb_k;
    int64_t gdata[FW_MAX_GROUNDS];
    int64_t host_bytes;  // [0, host_bytes) is built on the host; the rest is cleared on the device
    int64_t total;
};

Layout make_layout(const fw_scene_desc* d, int64_t n_slots, int64_t n_hslots, int64_t n_hnode, int64_t n_ptr,
                   int64_t n_tiles) {
    Layout L;
    int64_t o = 0;
    auto take = [&](int64_t bytes) {
        int64_t r = o;
        o += align_up(bytes);
        return r;
    };
    const int64_t N = d->n_nodes, E = d->n_elems;
    // springs need velocities (and halo velocity copies) where triangles need rest
    // positions, and no element arrays: the rest length is in the slot record
    const bool spr = d->mat_type == FW_MAT_SPRING;
    // arrays the host fills, uploaded at admission
    L.pos0 = take(8 * N);
    L.rest = spr ? -1 : take(8 * N);
    L.vel = take(8 * N);
    L.flag = take(4 * N);
    L.bcv = take(8 * std::max<int64_t>(d->n_bc_values, 1));
    L.rec = take(8 * n_slots);
    L.hel = take(4 * std::max<int64_t>(n_hslots, 1));
    L.hpos0 = take(8 * n_hnode);
    L.hrest = spr ? -1 : take(8 * n_hnode);
    L.hvel0 = spr ? take(8 * n_hnode) : -1;
    L.push = take(8 * n_hnode);
    L.cptr = take(2 * n_ptr);
    L.tiles = take((int64_t)sizeof(TileDev) * n_tiles);
    L.aabb_k = (int32_t)std::max<int64_t>(1, std::min<int64_t>(n_tiles, FW_AABB_MAX_SUB));
    L.aabb = take(3 * (int64_t)L.aabb_k * FW_AABB_SUB_INTS * 4);
    for (int g = 0; g < FW_MAX_GROUNDS; ++g) {
        L.gdata[g] = -1;
        if (g >= d->n_grounds) continue;
        const fw_ground_desc& gd = d->grounds[g];
        if (gd.kind == FW_GROUND_HEIGHTFIELD) L.gdata[g] = take(hf_bytes(gd.hf_nx));
        if (gd.kind == FW_GROUND_VOXEL) L.gdata[g] = take(8 * (int64_t)gd.vox_nx * gd.vox_ny);
        // rigid: the rigid pass's float4 per node, then the boundary nodes' internal ids
        if (gd.kind == FW_GROUND_RIGID)
            L.gdata[g] = take(gd.rigid_type == FW_RIGID_MESH ? mesh_layout(N, gd).total : 16 * N + 4 * (int64_t)N);
    }
    L.host_bytes = o;
    // FEM-FEM contributions: int2 (first, count) per node, then the float2 entries;
    // cleared on the device, filled by the contact group before each substep
    for (int g = 0; g < d->n_grounds; ++g)
        if (d->grounds[g].kind == FW_GROUND_FLEX) L.gdata[g] = take(8 * N + 8 * (int64_t)d->grounds[g].flex_cap);
    // arrays that start at zero (cleared on the device) or that the device writes
    // before reading: the second position and halo buffers (the warm-up substep
    // writes them), mass / J^-1 / area (prepare kernel), U, J2 state, diagnostics
    L.pos1 = take(8 * N);
    L.hpos1 = take(8 * n_hnode);
    L.hvel1 = spr ? take(8 * n_hnode) : -1;
    L.mass = take(4 * N);
    L.disp = d->track_disp ? take(8 * N) : -1;
    L.force = take(8 * N);
    L.cflag = take(N);
    L.jinv = spr ? -1 : take(16 * E);
    L.area = spr ? -1 : take(4 * E);
    if (d->mat_type == FW_MAT_MISES) {
        L.j2a = take(16 * E);
        L.j2b = take(16 * E);
    } else {
        L.j2a = L.j2b = -1;
    }
    L.stress = spr ? -1 : take(16 * E);
    L.strain = spr ? -1 : take(16 * E);
    L.yflag = spr ? -1 : take(E);
    L.total = o;
    return L;
}

// A mesh record's arrays, checked so that no index the rigid pass follows leaves them.
int validate_mesh(const fw_ground_desc& gd) {
    auto bad = [](const char* what) { return fail(FW_E_INVALID, std::string("mesh rigid record: ") + what); };
    if (gd.mesh_nodes < 1 || gd.mesh_edges < 1) return bad("needs boundary nodes and edges");
    if (!gd.mesh_coords || !gd.mesh_conn || !gd.mesh_elb || !gd.mesh_eub || !gd.mesh_margin || !gd.mesh_reg ||
        !gd.mesh_cell || !gd.mesh_sorted)
        return bad("arrays are required");
    if (gd.mesh_cap_reg < gd.mesh_nreg || gd.mesh_cap_cells < gd.mesh_tc || gd.mesh_cap_pairs < gd.mesh_npairs ||
        gd.mesh_nreg < 0 || gd.mesh_tc < 0 || gd.mesh_npairs < 0)
        return bad("hash sizes exceed the capacities");
    if (gd.mesh_cn[0] < 1 || gd.mesh_cn[1] < 1 || (int64_t)gd.mesh_cn[0] * gd.mesh_cn[1] < gd.mesh_tc)
        return bad("cell counts do not cover the cells");
    if (!(gd.mesh_gs[0] > 0.0f) || !(gd.mesh_gs[1] > 0.0f)) return bad("cell size must be positive");
    if (gd.mesh_subcycle < 1) return bad("subcycle must be >= 1");
    for (int64_t e = 0; e < 2 * (int64_t)gd.mesh_edges; ++e)
        if (gd.mesh_conn[e] < 0 || gd.mesh_conn[e] >= gd.mesh_nodes) return bad("edge node outside the nodes");
    for (int64_t r = 0; r < gd.mesh_nreg; ++r)
        if (gd.mesh_reg[2 * r + 1] < 0 || gd.mesh_reg[2 * r + 1] >= gd.mesh_edges) return bad("registry edge outside");
    for (int64_t c = 0; c < gd.mesh_tc; ++c)
        if (gd.mesh_cell[2 * c] < 0 || gd.mesh_cell[2 * c] > gd.mesh_cell[2 * c + 1] || gd.mesh_cell[2 * c + 1] > gd.mesh_npairs)
            return bad("cell range outside the entries");
    for (int64_t q = 0; q < gd.mesh_npairs; ++q)
        if (gd.mesh_sorted[q] >= gd.mesh_nreg) return bad("cell entry outside the registry");
    return FW_OK;
}

int validate_rigid(const fw_ground_desc& gd) {
    if (gd.rigid_type == FW_RIGID_MESH) {
        if (gd.rigid_tied != 0 && gd.rigid_tied != 1) return fail(FW_E_INVALID, "rigid record: rigid_tied must be 0 or 1");
        return validate_mesh(gd);
    }
    if (gd.rigid_type != FW_RIGID_BALL && gd.rigid_type != FW_RIGID_BOX && gd.rigid_type != FW_RIGID_CAPSULE)
        return fail(FW_E_INVALID, "rigid record: type must be FW_RIGID_BALL, FW_RIGID_BOX, FW_RIGID_CAPSULE or FW_RIGID_MESH");
    if (gd.rigid_type == FW_RIGID_CAPSULE && gd.rigid_prim[0] == 0.0f && gd.rigid_prim[1] == 0.0f)
        return fail(FW_E_INVALID, "rigid record: a capsule needs a nonzero axis");
    if (gd.rigid_tied != 0 && gd.rigid_tied != 1) return fail(FW_E_INVALID, "rigid record: rigid_tied must be 0 or 1");
    return FW_OK;
}

// The device fields of a rigid record (GroundDev layout in fw_types.h).
void pack_rigid(GroundDev& G, const fw_ground_desc& gd) {
    G.a[0] = gd.rigid_center[0];
    G.a[1] = gd.rigid_center[1];
    G.a[2] = gd.rigid_prim[0];
    G.a[3] = gd.rigid_prim[1];
    G.nx = gd.rigid_type;
    G.ny = gd.rigid_tied;
    G.hf_dxw = gd.rigid_radius;
    for (int k = 0; k < 4; ++k) G.b[k] = gd.rigid_rot[k];
    G.b[4] = gd.rigid_vel[0];
    G.b[5] = gd.rigid_vel[1];
    G.b[6] = gd.rigid_omega;
    G.b[7] = gd.rigid_friction;
}

// The ground AABB the gate reads, raw and padded by 0.1 x extent.
void pack_aabb(GroundDev& G, const fw_ground_desc& gd) {
    for (int k = 0; k < 2; ++k) {
        G.lo[k] = gd.aabb_lo[k];
        G.hi[k] = gd.aabb_hi[k];
        float b = 0.1f * (gd.aabb_hi[k] - gd.aabb_lo[k]);
        G.plo[k] = gd.aabb_lo[k] - b;
        G.phi[k] = gd.aabb_hi[k] + b;
    }
}

int validate(const fw_scene_desc* d) {
    if (!d) return fail(FW_E_INVALID, "null scene description");
    if (d->n_nodes <= 0 || d->n_elems <= 0) return fail(FW_E_INVALID, "scene needs at least one node and one element");
    if (!d->x0 || !d->v0 || !d->conn || !d->flags || !d->bc_values)
        return fail(FW_E_INVALID, "scene arrays x0, v0, conn, flags and bc_values are required");
    if (d->mat_type < FW_MAT_ELASTIC || d->mat_type > FW_MAT_SPRING)
        return fail(FW_E_INVALID, "unknown material type " + std::to_string(d->mat_type));
    if (d->mat_type == FW_MAT_SPRING && !d->rest_len) return fail(FW_E_INVALID, "spring scene needs rest_len");
    if (d->n_grounds < 0 || d->n_grounds > FW_MAX_GROUNDS)
        return fail(FW_E_UNSUPPORTED, "a scene holds at most " + std::to_string(FW_MAX_GROUNDS) + " grounds");
    if (d->n_grounds > 0 && !d->grounds) return fail(FW_E_INVALID, "grounds pointer is null");
    if (d->n_bc_values <= 0 || d->n_bc_values > 65536) return fail(FW_E_INVALID, "n_bc_values must be in [1, 65536]");
    for (int i = 0; i < d->n_nodes; ++i)
        if ((int)(d->flags[i] >> FW_NODE_BCV_SHIFT) >= d->n_bc_values)
            return fail(FW_E_INVALID, "node flag indexes a BC value beyond the table");
    for (int g = 0; g < d->n_grounds; ++g) {
        const fw_ground_desc& gd = d->grounds[g];
        if (gd.kind == FW_GROUND_HEIGHTFIELD && (gd.hf_nx < 2 || !gd.heights))
            return fail(FW_E_INVALID, "height field needs at least 2 samples");
        if (gd.kind == FW_GROUND_VOXEL && (gd.vox_nx < 1 || gd.vox_ny < 1 || !gd.occ))
            return fail(FW_E_INVALID, "voxel grid needs nx, ny >= 1 and an occupancy array");
        if (gd.kind == FW_GROUND_VOXEL && (gd.vox_nx > 65534 || gd.vox_ny > 65534))
            return fail(FW_E_INVALID, "voxel grid is limited to 65534 cells per axis");
        if (gd.kind < FW_GROUND_PLANE || gd.kind > FW_GROUND_RIGID)
            return fail(FW_E_INVALID, "unknown ground kind " + std::to_string(gd.kind));
        if (gd.kind == FW_GROUND_RIGID) {
            int rc = validate_rigid(gd);
            if (rc) return rc;
        }
        if (gd.kind == FW_GROUND_FLEX) {
            if (d->mat_type == FW_MAT_SPRING) return fail(FW_E_UNSUPPORTED, "FEM-FEM contact records need a FEM scene");
            if (gd.flex_cap < 0) return fail(FW_E_INVALID, "FEM-FEM record needs flex_cap >= 0");
            for (int h = 0; h < g; ++h)
                if (d->grounds[h].kind == FW_GROUND_FLEX) return fail(FW_E_INVALID, "at most one FEM-FEM record per scene");
        }
    }
    return FW_OK;
}

// ------------------------------------------------------------------ pinned staging

// Per cell of a voxel grid, the four scan results of the signed-distance query
// (flatworld/mixedcontact.py:220-302): scanning from an occupied cell (i, j) along
// -x, +x, -z, +z for the first empty cell ends at i0 - 1, i1, j0 - 1, j1, stored as
// (i0, i1, j0, j1); a cell that is not occupied (value != 1) stores i0 = 0xFFFF.
// The device reads one 8-byte entry instead of walking the grid, and evaluates
// the same float expressions on the same integers.
void build_voxel_extents(const fw_ground_desc& gd, uint16_t* ext) {
    const int nx = gd.vox_nx, ny = gd.vox_ny;
    auto at = [&](int i, int j) { return gd.occ[(int64_t)i * ny + j]; };
    std::vector<uint16_t> lo(nx), hi(nx);
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) lo[i] = (i > 0 && at(i - 1, j) != 0) ? lo[i - 1] : (uint16_t)i;
        for (int i = nx - 1; i >= 0; --i) hi[i] = (i + 1 < nx && at(i + 1, j) != 0) ? hi[i + 1] : (uint16_t)(i + 1);
        for (int i = 0; i < nx; ++i) {
            uint16_t* e = ext + 4 * ((int64_t)i * ny + j);
            e[0] = at(i, j) == 1 ? lo[i] : (uint16_t)0xFFFF;
            e[1] = hi[i];
        }
    }
    std::vector<uint16_t> lo2(ny), hi2(ny);
    for (int i = 0; i < nx; ++i) {
        for (int j = 0; j < ny; ++j) lo2[j] = (j > 0 && at(i, j - 1) != 0) ? lo2[j - 1] : (uint16_t)j;
        for (int j = ny - 1; j >= 0; --j) hi2[j] = (j + 1 < ny && at(i, j + 1) != 0) ? hi2[j + 1] : (uint16_t)(j + 1);
        for (int j = 0; j < ny; ++j) {
            uint16_t* e = ext + 4 * ((int64_t)i * ny + j);
            e[2] = lo2[j];
            e[3] = hi2[j];
        }
    }
}

struct Staging {
    void* ptr = nullptr;
    int64_t size = 0;
    hipEvent_t ev = nullptr;
    bool busy = false;
    bool held = false;  // holds a retired scene's readback until fw_retire_wait copies it out
};

// A retired scene whose final state is on its way to the host (fw_scene_retire).
struct PendingRetire {
    Staging* st = nullptr;
    int64_t n_nodes = 0;
    std::vector<int32_t> perm;  // internal -> caller node order
    float* x_out = nullptr;
    float* v_out = nullptr;
};

}  // namespace

// A FEM-FEM contact group: its device block (arena memory) and table entry.
struct FlexHost {
    int64_t block_off = 0, block_bytes = 0;
    std::vector<int32_t> scenes;
    int32_t n_items = 0, max_cells = 0, max_pairs = 0, n_tied = 0;
    bool hash_built = false;  // set by the first rebuild; tie resolution walks that hash
    int32_t sched[FW_FLEX_SCHED];  // host copy of the device schedule (INT32_MIN: free)
    FlexDev dev{};
};

struct SceneHost {
    bool live = false;
    int flex_group = -1;  // contact group holding this scene, -1 for none
    int64_t block_off = 0, block_bytes = 0;
    const TileDev* tiles_dev = nullptr;  // the scene's own copy of its tile records (in its block)
    int32_t n_nodes = 0, n_elems = 0, mat_type = 0;
    int32_t s0 = 0, horizon = INT32_MAX;
    Tiling tiling;
    std::vector<TileDev> tiles;  // scene-relative, scene id filled in
    int tile_base = 0;           // position in the arena tile table
    SceneDev dev;
    MeshHost mesh[FW_MAX_GROUNDS];  // mesh rigid records
};

// A scene prepared on the host for an arena tuning (fw_scene_prepare).
// Page-locked image buffers of prepared scenes, shared by every thread and arena:
// a worker thread copies a prepared image into one (fw_prepared_stage), the commit
// DMAs from it directly, and the buffer is reused once that copy has completed.
struct PinnedBuf {
    void* ptr = nullptr;
    int64_t size = 0;
    hipEvent_t ev = nullptr;
    bool in_flight = false;  // a commit's copy may still read it (ev recorded behind it)
};

static std::mutex g_pin_mu;
static std::vector<PinnedBuf*> g_pin_free;  // idle or in flight; never freed before the process ends
// Held by a graph capture and by the pool's HIP calls from worker threads: a HIP call
// from another thread (an event query, a page-locked allocation) during a capture can
// invalidate the capture, so the two never overlap.
static std::mutex g_capture_mu;

static PinnedBuf* pinned_get(int64_t bytes) {
    std::lock_guard<std::mutex> cap(g_capture_mu);
    {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        for (size_t i = 0; i < g_pin_free.size(); ++i) {
            PinnedBuf* b = g_pin_free[i];
            if (b->size < bytes || b->size > 4 * bytes + (1 << 20)) continue;
            if (b->in_flight) {
                if (hipEventQuery(b->ev) != hipSuccess) continue;
                b->in_flight = false;
            }
            g_pin_free.erase(g_pin_free.begin() + (long)i);
            return b;
        }
    }
    auto* b = new PinnedBuf;
    b->size = std::max<int64_t>(bytes, 1 << 16);
    if (hipHostMalloc(&b->ptr, (size_t)b->size, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        if (b->ptr) (void)hipHostFree(b->ptr);
        delete b;
        return nullptr;
    }
    return b;
}

static void pinned_put(PinnedBuf* b) {
    std::lock_guard<std::mutex> lk(g_pin_mu);
    g_pin_free.push_back(b);
}

struct fw_prepared {
    fw_tuning tune{};
    Tiling T;               // tiles carry offsets; base pointer, scene id and s0 are set at the commit
    Layout L{};
    std::vector<char> img;  // the block's bytes, tile table excluded (written at the commit)
    PinnedBuf* pin = nullptr;  // the same bytes in page-locked memory, after fw_prepared_stage
    SceneDev D{};           // scene entry without device addresses
    int32_t n_nodes = 0, n_elems = 0, mat_type = 0, horizon = 0;
    MeshHost mesh[FW_MAX_GROUNDS];
    bool consumed = false;
};

struct fw_arena {
    int device = 0;
    hipStream_t stream = nullptr;
    char* base = nullptr;
    int64_t capacity = 0;
    std::map<int64_t, int64_t> free_list;  // offset -> size
    int64_t used = 0;
    fw_tuning tune{};
    int n_cu = 0;
    // device tables
    SceneDev* d_scenes = nullptr;
    int max_scenes = 0;
    TileDev* d_tiles = nullptr;
    int tile_cap = 0;
    ArenaCtl* d_ctl = nullptr;
    // host state
    std::vector<std::unique_ptr<SceneHost>> scenes;
    std::vector<int> free_ids;
    std::vector<TileDev> tile_table;  // triangle scenes' tiles first, then spring scenes'
    int n_tri_tiles = 0;
    bool tiles_dirty = false;
    int64_t clock = 0;  // host mirror of ctl->clock
    std::vector<std::unique_ptr<Staging>> staging;
    std::vector<PendingRetire> retiring;
    // graph cache: one executable graph per (substeps, flags), all captured for the same
    // launch grid and element kinds; dropped together when either changes
    struct GraphEntry {
        hipGraphExec_t exec;
        int n;
        uint32_t flags;
    };
    std::vector<GraphEntry> graphs;
    // GPU-busy accounting (fw_arena_busy): events bracketing each fw_step / fw_step_graph
    // call on the arena stream, resolved when queried
    bool track_busy = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> busy_pending;
    std::vector<hipEvent_t> busy_pool;
    double busy_ms = 0.0;
    int64_t busy_calls = 0;
    int graph_grid = 0;
    int graph_kinds = 0;  // bit 0: triangle launch captured, bit 1: spring launch captured
    // FEM-FEM contact groups (id -> group, nullptr once removed) and the device
    // table of the live ones, in id order; flex_row[id] is a group's table row
    std::vector<std::unique_ptr<FlexHost>> flex;
    std::vector<int> flex_row;
    FlexDev* d_flex = nullptr;
    int2* d_flex_items = nullptr;  // (table row, work item) of every global work item
    int32_t* d_flex_sched = nullptr;  // FW_FLEX_SCHED scheduled rebuild steps per table row
    int flex_cap = 0, n_flex_live = 0, n_flex_items = 0, flex_items_cap = 0;
    int64_t flex_scratch_bytes = 0;  // compact per-substep scratch of the live groups (fw_step_noalloc)
    // (scene, record) of every rigid contact record of the live scenes, for the rigid pass
    int2* d_rigid = nullptr;
    int rigid_cap = 0, n_rigid_live = 0;
    bool flex_dirty = false;
    // one-entry tiling cache: fw_scene_footprint followed by fw_scene_add (or a retry of
    // an add that found the arena full) reuses the tiling instead of rebuilding it
    struct {
        const void* x0 = nullptr;
        const void* conn = nullptr;
        int32_t n_nodes = -1, n_elems = -1;
        int64_t conn_sum = 0, x_sum = 0;
        std::unique_ptr<Tiling> tiling;
    } cache;
};

namespace {

// LDS nodes (owned + halo) a tile of the compiled shapes holds: FW_SHAPES gives the
// same count for every thread count of a slot count.
int node_cap_of_slots(int slots, bool spring = false) {
#define FW_CAP(B, S, N, P, NS) \
    if (slots == S) return spring ? NS : N;
    FW_SHAPES(FW_CAP)
#undef FW_CAP
    return 0;
}
// LDS nodes a tile of this scene may hold: spring tiles have their own, smaller LDS layout
int node_cap_of(const fw_tuning& t, const fw_scene_desc* d) {
    return node_cap_of_slots(t.tile_slots, d->mat_type == FW_MAT_SPRING);
}

// Tiling of a scene description, from the cache when the same arrays are presented again.
int tiling_for(fw_arena* a, const fw_scene_desc* d, std::unique_ptr<Tiling>* out) {
    int64_t cs = 0, xs = 0;
    const int64_t E3 = corners_of(d) * (int64_t)d->n_elems;
    for (int64_t i = 0; i < E3; i += std::max<int64_t>(1, E3 / 64)) cs = cs * 31 + d->conn[i];
    cs = cs * 31 + d->mat_type;
    for (int64_t i = 0; i < 2 * (int64_t)d->n_nodes; i += std::max<int64_t>(1, d->n_nodes / 32)) {
        int32_t b;
        std::memcpy(&b, d->x0 + i, 4);
        xs = xs * 31 + b;
    }
    auto& c = a->cache;
    if (c.tiling && c.x0 == d->x0 && c.conn == d->conn && c.n_nodes == d->n_nodes && c.n_elems == d->n_elems &&
        c.conn_sum == cs && c.x_sum == xs) {
        *out = std::move(c.tiling);
        return FW_OK;
    }
    auto T = std::make_unique<Tiling>();
    int rc = build_tiling(d, a->tune.tile_nodes, a->tune.tile_slots, node_cap_of(a->tune, d), *T);
    if (rc) return rc;
    c.x0 = d->x0;
    c.conn = d->conn;
    c.n_nodes = d->n_nodes;
    c.n_elems = d->n_elems;
    c.conn_sum = cs;
    c.x_sum = xs;
    *out = std::move(T);
    return FW_OK;
}

void cache_keep(fw_arena* a, std::unique_ptr<Tiling> T) { a->cache.tiling = std::move(T); }

int64_t arena_alloc(fw_arena* a, int64_t bytes) {
    for (auto it = a->free_list.begin(); it != a->free_list.end(); ++it) {
        if (it->second >= bytes) {
            int64_t off = it->first, sz = it->second;
            a->free_list.erase(it);
            if (sz > bytes) a->free_list[off + bytes] = sz - bytes;
            a->used += bytes;
            return off;
        }
    }
    return -1;
}

void arena_free(fw_arena* a, int64_t off, int64_t bytes) {
    a->used -= bytes;
    auto it = a->free_list.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != a->free_list.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        a->free_list.erase(nx);
    }
    if (it != a->free_list.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            a->free_list.erase(it);
        }
    }
}

// pinned buffer whose previous copy has finished
int get_staging(fw_arena* a, int64_t bytes, Staging** out) {
    for (auto& s : a->staging) {
        if (s->busy && hipEventQuery(s->ev) == hipSuccess) s->busy = false;
    }
    for (auto& s : a->staging) {
        if (!s->busy && !s->held && s->size >= bytes) {
            *out = s.get();
            return FW_OK;
        }
    }
    auto s = std::make_unique<Staging>();
    s->size = std::max<int64_t>(bytes, 4 << 20);
    HIPCHK(hipHostMalloc(&s->ptr, s->size, hipHostMallocDefault));
    HIPCHK(hipEventCreateWithFlags(&s->ev, hipEventDisableTiming));
    *out = s.get();
    a->staging.push_back(std::move(s));
    return FW_OK;
}

int upload(fw_arena* a, void* dst, const void* src, int64_t bytes) {
    if (bytes <= 0) return FW_OK;
    Staging* s = nullptr;
    int rc = get_staging(a, bytes, &s);
    if (rc) return rc;
    std::memcpy(s->ptr, src, bytes);
    HIPCHK(hipMemcpyAsync(dst, s->ptr, bytes, hipMemcpyHostToDevice, a->stream));
    HIPCHK(hipEventRecord(s->ev, a->stream));
    s->busy = true;
    return FW_OK;
}

void drop_graphs(fw_arena* a) {
    for (auto& g : a->graphs) (void)hipGraphExecDestroy(g.exec);
    a->graphs.clear();
}

int ensure_tiles(fw_arena* a) {
    if (!a->tiles_dirty) return FW_OK;
    a->tile_table.clear();
    for (int pass = 0; pass < 2; ++pass) {  // triangle scenes, then spring scenes
        for (auto& sp : a->scenes) {
            if (!sp || !sp->live || (sp->mat_type == FW_MAT_SPRING) != (pass == 1)) continue;
            sp->tile_base = (int)a->tile_table.size();
            a->tile_table.insert(a->tile_table.end(), sp->tiles.begin(), sp->tiles.end());
        }
        if (pass == 0) a->n_tri_tiles = (int)a->tile_table.size();
    }
    std::vector<int2> rigid_rows;
    for (size_t id = 0; id < a->scenes.size(); ++id) {
        const auto& sp = a->scenes[id];
        if (!sp || !sp->live) continue;
        for (int g = 0; g < sp->dev.hot.n_grounds; ++g)
            if (sp->dev.ground[g].kind == FW_GROUND_RIGID || sp->dev.ground[g].kind == FW_GROUND_RIGID_MESH_DEV)
                rigid_rows.push_back(make_int2((int)id, g));
    }
    if ((int)rigid_rows.size() > a->rigid_cap) {
        const int cap = std::max((int)rigid_rows.size(), std::max(64, 2 * a->rigid_cap));
        int2* nd = nullptr;
        HIPCHK(hipStreamSynchronize(a->stream));
        HIPCHK(hipMalloc(&nd, sizeof(int2) * (size_t)cap));
        if (a->d_rigid) HIPCHK(hipFree(a->d_rigid));
        a->d_rigid = nd;
        a->rigid_cap = cap;
        drop_graphs(a);  // the graphs hold the old table pointer
    }
    if (!rigid_rows.empty()) {
        int rc0 = upload(a, a->d_rigid, rigid_rows.data(), (int64_t)sizeof(int2) * (int64_t)rigid_rows.size());
        if (rc0) return rc0;
    }
    if (a->n_rigid_live != (int)rigid_rows.size()) drop_graphs(a);  // the graphs hold the rigid launch's grid
    a->n_rigid_live = (int)rigid_rows.size();
    const int nt = (int)a->tile_table.size();
    if (nt > a->tile_cap) {
        int cap = std::max(nt, std::max(1024, a->tile_cap * 2));
        TileDev* nd = nullptr;
        HIPCHK(hipStreamSynchronize(a->stream));
        HIPCHK(hipMalloc(&nd, sizeof(TileDev) * (size_t)cap));
        if (a->d_tiles) HIPCHK(hipFree(a->d_tiles));
        a->d_tiles = nd;
        a->tile_cap = cap;
        drop_graphs(a);  // the graphs hold the old table pointer
    }
    int rc = upload(a, a->d_tiles, a->tile_table.data(), (int64_t)sizeof(TileDev) * nt);
    if (rc) return rc;
    ArenaCtl ctl{(int32_t)a->clock, nt, a->n_tri_tiles};
    rc = upload(a, &a->d_ctl->n_tiles, &ctl.n_tiles, 2 * sizeof(int32_t));  // n_tiles, n_tri_tiles
    if (rc) return rc;
    a->tiles_dirty = false;
    return FW_OK;
}

// Every live group's scheduled rebuild steps, in table-row order, in one copy.
int upload_sched(fw_arena* a, int n_rows) {
    if (n_rows <= 0) return FW_OK;
    std::vector<int32_t> all((size_t)n_rows * FW_FLEX_SCHED, INT32_MIN);
    for (size_t g = 0; g < a->flex.size(); ++g) {
        const int r = g < a->flex_row.size() ? a->flex_row[g] : -1;
        if (!a->flex[g] || r < 0 || r >= n_rows) continue;
        for (int j = 0; j < FW_FLEX_SCHED; ++j) all[(size_t)r * FW_FLEX_SCHED + j] = a->flex[g]->sched[j];
    }
    return upload(a, a->d_flex_sched, all.data(), (int64_t)sizeof(int32_t) * (int64_t)all.size());
}

// Device table of the live FEM-FEM groups, in group-id order.
int ensure_flex(fw_arena* a) {
    if (!a->flex_dirty) return FW_OK;
    drop_graphs(a);  // the graphs hold the group count, the work-item grid and the table pointers
    std::vector<FlexDev> rows;
    a->flex_row.assign(a->flex.size(), -1);
    for (size_t g = 0; g < a->flex.size(); ++g) {
        if (!a->flex[g]) continue;
        a->flex_row[g] = (int)rows.size();
        rows.push_back(a->flex[g]->dev);
    }
    const int n = (int)rows.size();
    if (n > a->flex_cap) {
        const int cap = std::max(n, std::max(8, 2 * a->flex_cap));
        FlexDev* nd = nullptr;
        int32_t* ns = nullptr;
        HIPCHK(hipStreamSynchronize(a->stream));
        HIPCHK(hipMalloc(&nd, sizeof(FlexDev) * (size_t)cap));
        HIPCHK(hipMalloc(&ns, sizeof(int32_t) * FW_FLEX_SCHED * (size_t)cap));
        if (a->d_flex) HIPCHK(hipFree(a->d_flex));
        if (a->d_flex_sched) HIPCHK(hipFree(a->d_flex_sched));
        a->d_flex = nd;
        a->d_flex_sched = ns;
        a->flex_cap = cap;
    }
    for (int r = 0; r < n; ++r) rows[r].sched = a->d_flex_sched + (size_t)r * FW_FLEX_SCHED;
    std::vector<int32_t> item_off(n + 1, 0);
    for (int r = 0; r < n; ++r) item_off[r + 1] = item_off[r] + rows[r].n_items;
    if (item_off[n] > a->flex_items_cap) {
        const int cap = std::max(item_off[n], 2 * a->flex_items_cap);
        int2* ni = nullptr;
        HIPCHK(hipStreamSynchronize(a->stream));
        HIPCHK(hipMalloc(&ni, sizeof(int2) * (size_t)cap));
        if (a->d_flex_items) HIPCHK(hipFree(a->d_flex_items));
        a->d_flex_items = ni;
        a->flex_items_cap = cap;
    }
    std::vector<int2> item_map((size_t)item_off[n]);
    for (int r = 0; r < n; ++r)
        for (int w = 0; w < rows[r].n_items; ++w) item_map[(size_t)item_off[r] + w] = make_int2(r, w);
    int64_t scr = 0;  // the compact per-substep scratch layout of fw_step_noalloc
    for (int r = 0; r < n; ++r) {
        rows[r].scr_off = scr;
        scr += flex_scratch_layout(rows[r].n_items, rows[r].n_slots).bytes;
    }
    a->flex_scratch_bytes = scr;
    int rc = upload(a, a->d_flex, rows.data(), (int64_t)sizeof(FlexDev) * n);
    if (rc) return rc;
    rc = upload_sched(a, n);
    if (rc) return rc;
    if (item_off[n] > 0) {
        rc = upload(a, a->d_flex_items, item_map.data(), (int64_t)sizeof(int2) * item_off[n]);
        if (rc) return rc;
    }
    a->n_flex_live = n;
    a->n_flex_items = item_off[n];
    a->flex_dirty = false;
    return FW_OK;
}

int launch_grid(fw_arena* a, int ntiles) {
    int g = ntiles;
    if (a->tune.grid_per_cu > 0) g = std::min(ntiles, a->tune.grid_per_cu * a->n_cu);
    g = std::max(g, 1);
    return (g + 7) / 8 * 8;
}

SceneHost* get_scene(fw_arena* a, int id) {
    if (id < 0 || id >= (int)a->scenes.size() || !a->scenes[id] || !a->scenes[id]->live) return nullptr;
    return a->scenes[id].get();
}

// substeps this scene has completed (warm-up excluded), and the buffer holding its state
void scene_state(const fw_arena* a, const SceneHost* s, int64_t* done, int* buf, bool* warmed) {
    int64_t ran = a->clock - ((int64_t)s->s0 - 1);  // launches that covered this scene incl. warm-up
    if (ran <= 0) {
        *done = 0;
        *buf = 0;
        *warmed = false;
        return;
    }
    int64_t c = std::min<int64_t>(ran - 1, s->horizon);
    *done = c;
    *buf = (int)((c + 1) & 1);
    *warmed = true;
}

Layout layout_of(const fw_scene_desc* d, const Tiling& T) {
    return make_layout(d, (int64_t)(T.rec.size() / 2), (int64_t)T.hel.size(), (int64_t)T.hnode.size(),
                       (int64_t)T.csr_ptr.size(), (int64_t)T.tiles.size());
}

// Host-only half of adding a scene: validation, tiling and the packed image of
// the scene's block (everything but the device addresses).  Uses no arena state
// except the optional tiling cache, so scenes can be prepared on other threads.
int prepare_scene(const fw_scene_desc* d, const fw_tuning& tune, fw_arena* cache_arena,
                  std::unique_ptr<fw_prepared>* out) {
    int rc = validate(d);
    if (rc) return rc;
    auto P = std::make_unique<fw_prepared>();
    P->tune = tune;
    {
        std::unique_ptr<Tiling> Tp;
        if (cache_arena) {
            rc = tiling_for(cache_arena, d, &Tp);
        } else {
            Tp = std::make_unique<Tiling>();
            rc = cached_tiling(d, tune.tile_nodes, tune.tile_slots, node_cap_of(tune, d), *Tp);
        }
        if (rc) return rc;
        P->T = std::move(*Tp);
    }
    Tiling& T = P->T;
    const int N = d->n_nodes, E = d->n_elems;
    P->L = layout_of(d, T);
    const Layout& L = P->L;
    P->n_nodes = N;
    P->n_elems = E;
    P->mat_type = d->mat_type;
    P->horizon = d->horizon > 0 ? d->horizon : INT32_MAX;

    // scene entry, device addresses left for the commit
    SceneDev& D = P->D;
    std::memset(&D, 0, sizeof(D));
    D.n_nodes = N;
    D.n_elems = E;
    D.hot.mat_type = d->mat_type;
    D.hot.n_grounds = d->n_grounds;
    D.hot.E = d->E;
    D.hot.nu = d->mat_type == FW_MAT_SPRING ? d->spring_c : d->nu;  // springs: the damping coefficient
    D.rho = d->rho;
    D.hot.sigma_y = d->sigma_y;
    D.hot.H = d->H;
    D.hot.dt = d->dt;
    D.hot.damping = d->damping;
    D.horizon = P->horizon;
    D.live = 1;
    D.hot.aabb_k = L.aabb_k;
    auto o8 = [](int64_t off) { return (uint32_t)(off >> 8); };  // layout offsets are 256-byte aligned
    for (auto& td : T.tiles) {
        td.base = nullptr;
        td.o_rec = o8(L.rec);
        td.o_hel = o8(L.hel);
        td.o_bcv = o8(L.bcv);
        td.o_pos[0] = o8(L.pos0);
        td.o_pos[1] = o8(L.pos1);
        td.o_hpos[0] = o8(L.hpos0);
        td.o_hpos[1] = o8(L.hpos1);
        if (d->mat_type == FW_MAT_SPRING) {
            td.o_aux = o8(L.vel);
            td.o_haux[0] = o8(L.hvel0);
            td.o_haux[1] = o8(L.hvel1);
        } else {
            td.o_aux = o8(L.rest);
            td.o_haux[0] = td.o_haux[1] = o8(L.hrest);
        }
        td.o_vel = o8(L.vel);
        td.o_mass = o8(L.mass);
        td.o_flag = o8(L.flag);
        td.o_disp = L.disp >= 0 ? o8(L.disp) : 0;
        td.o_cptr = o8(L.cptr);
        td.o_push = o8(L.push);
        td.o_j2[0] = L.j2a >= 0 ? o8(L.j2a) : 0;
        td.o_j2[1] = L.j2b >= 0 ? o8(L.j2b) : 0;
        td.horizon = D.horizon;
    }

    // host image of the block
    P->img.assign((size_t)L.host_bytes, 0);
    char* img = P->img.data();
    float2* pos0 = (float2*)(img + L.pos0);
    float2* vel = (float2*)(img + L.vel);
    uint32_t* flag = (uint32_t*)(img + L.flag);
    for (int i = 0; i < N; ++i) {
        int o = T.node_perm[i];
        pos0[i] = make_float2(d->x0[2 * o], d->x0[2 * o + 1]);
        vel[i] = make_float2(d->v0[2 * o], d->v0[2 * o + 1]);
        flag[i] = d->flags[o];
    }
    if (L.rest >= 0) std::memcpy(img + L.rest, img + L.pos0, 8 * (size_t)N);
    {
        // every AABB sub-slot of the three buffers starts empty
        int32_t* ab = (int32_t*)(img + L.aabb);
        for (int j = 0; j < 3 * L.aabb_k; ++j) {
            ab[j * FW_AABB_SUB_INTS + 0] = fw_f2o(1e30f);
            ab[j * FW_AABB_SUB_INTS + 1] = fw_f2o(1e30f);
            ab[j * FW_AABB_SUB_INTS + 2] = fw_f2o(-1e30f);
            ab[j * FW_AABB_SUB_INTS + 3] = fw_f2o(-1e30f);
        }
    }
    {
        float2* hp = (float2*)(img + L.hpos0);
        for (size_t j = 0; j < T.hnode.size(); ++j) hp[j] = pos0[T.hnode[j]];
        if (L.hrest >= 0) std::memcpy(img + L.hrest, img + L.hpos0, 8 * T.hnode.size());
        if (L.hvel0 >= 0) {
            float2* hv = (float2*)(img + L.hvel0);
            for (size_t j = 0; j < T.hnode.size(); ++j) hv[j] = vel[T.hnode[j]];
        }
    }
    std::memcpy(img + L.bcv, d->bc_values, 8 * (size_t)d->n_bc_values);
    std::memcpy(img + L.rec, T.rec.data(), 4 * T.rec.size());
    std::memcpy(img + L.hel, T.hel.data(), 4 * T.hel.size());
    std::memcpy(img + L.push, T.push.data(), 4 * T.push.size());
    std::memcpy(img + L.cptr, T.csr_ptr.data(), 2 * T.csr_ptr.size());
    for (int g = 0; g < d->n_grounds; ++g) {
        const fw_ground_desc& gd = d->grounds[g];
        GroundDev& G = D.ground[g];
        G.kind = gd.kind;
        G.penalty = gd.penalty;
        pack_aabb(G, gd);
        if (gd.kind == FW_GROUND_PLANE) {
            G.a[0] = gd.point[0];
            G.a[1] = gd.point[1];
            G.a[2] = gd.normal[0];
            G.a[3] = gd.normal[1];
        } else if (gd.kind == FW_GROUND_HEIGHTFIELD) {
            G.a[0] = gd.hf_lb_x;
            G.a[1] = gd.hf_ub_x;
            G.nx = gd.hf_nx;
            {
                // sample spacing as the device query computes it: float32 subtraction, then division
                volatile float span = gd.hf_ub_x - gd.hf_lb_x;
                G.hf_dxw = span / (float)(gd.hf_nx - 1);
            }
            G.reverse = gd.reverse;
            std::memcpy(img + L.gdata[g], gd.heights, 4 * (size_t)gd.hf_nx);
            build_hf_normals(G, gd.heights, (float2*)(img + L.gdata[g] + hf_normals_offset(gd.hf_nx)));
            G.b[0] = hf_skip_level(gd.heights, gd.hf_nx, gd.reverse == 1);
        } else if (gd.kind == FW_GROUND_FLEX) {
            G.nx = N;             // the entries follow the per-node ranges
            G.ny = gd.flex_cap;
        } else if (gd.kind == FW_GROUND_RIGID) {
            pack_rigid(G, gd);
            int32_t* bn = (int32_t*)(img + L.gdata[g] + 16 * (int64_t)N);
            int nb = 0;
            for (int i = 0; i < N; ++i)
                if (flag[i] & FW_NODE_BOUNDARY) bn[nb++] = i;
            G.reverse = nb;
            if (gd.rigid_type == FW_RIGID_MESH) {
                // header, per-item state at the reference's initial values
                // (flatworld/mixedcontact.py:2052-2093) and the mesh arrays
                G.kind = FW_GROUND_RIGID_MESH_DEV;
                const MeshLayout M = mesh_layout(N, gd);
                char* b = img + L.gdata[g];
                const MeshDev H = mesh_header(gd, M, nb);
                std::memcpy(b + M.hdr, &H, sizeof(H));
                int32_t* cache = (int32_t*)(b + M.cache);
                int32_t* skip = (int32_t*)(b + M.skip);
                int32_t* fi = (int32_t*)(b + M.fric_i);
                for (int i = 0; i < N; ++i) {
                    cache[2 * i] = cache[2 * i + 1] = -1;
                    skip[i] = -1;
                    fi[2 * i] = -1;
                    fi[2 * i + 1] = 0;
                }
                std::memcpy(b + M.coords, gd.mesh_coords, 8 * (size_t)gd.mesh_nodes);
                std::memcpy(b + M.conn, gd.mesh_conn, 8 * (size_t)gd.mesh_edges);
                std::memcpy(b + M.elb, gd.mesh_elb, 8 * (size_t)gd.mesh_edges);
                std::memcpy(b + M.eub, gd.mesh_eub, 8 * (size_t)gd.mesh_edges);
                std::memcpy(b + M.margin, gd.mesh_margin, 4 * (size_t)gd.mesh_edges);
                std::memcpy(b + M.reg, gd.mesh_reg, 8 * (size_t)gd.mesh_nreg);
                std::memcpy(b + M.cell, gd.mesh_cell, 8 * (size_t)gd.mesh_tc);
                std::memcpy(b + M.sorted, gd.mesh_sorted, 4 * (size_t)gd.mesh_npairs);
                MeshHost& mh = P->mesh[g];
                mh.used = true;
                mh.ml = M;
                mh.nodes = gd.mesh_nodes;
                mh.edges = gd.mesh_edges;
                mh.cap_reg = gd.mesh_cap_reg;
                mh.cap_cells = gd.mesh_cap_cells;
                mh.cap_pairs = gd.mesh_cap_pairs;
                mh.n_items = nb;
            }
        } else {
            G.a[0] = gd.vox_lb[0];
            G.a[1] = gd.vox_lb[1];
            G.a[2] = gd.vox_dx;
            G.a[3] = gd.vox_dz;
            G.nx = gd.vox_nx;
            G.ny = gd.vox_ny;
            build_voxel_extents(gd, (uint16_t*)(img + L.gdata[g]));
        }
    }
    *out = std::move(P);
    return FW_OK;
}

// Device half: allocate the block, patch the device addresses into the image,
// upload it and the scene entry, and compute area / J^-1 / lumped mass on the
// device.  The prepared scene's tiling moves into the arena.
int commit_scene(fw_arena* a, fw_prepared* p, int32_t* scene_id) {
    const Layout& L = p->L;
    int id_probe = !a->free_ids.empty() ? a->free_ids.back() : (int)a->scenes.size();
    if (id_probe >= a->max_scenes)
        return fail(FW_E_CAPACITY, "arena holds at most " + std::to_string(a->max_scenes) + " scenes");
    if (L.total > a->capacity - a->used)
        return fail(FW_E_CAPACITY, "arena full: scene needs " + std::to_string(L.total) + " bytes, " +
                                       std::to_string(a->capacity - a->used) + " free");
    Staging* st = nullptr;
    int rc = FW_OK;
    if (!p->pin) {
        rc = get_staging(a, L.host_bytes, &st);
        if (rc) return rc;
    }
    int64_t off = arena_alloc(a, L.total);
    if (off < 0)
        return fail(FW_E_CAPACITY, "arena full: scene needs " + std::to_string(L.total) + " bytes, " +
                                       std::to_string(a->capacity - a->used) + " free (fragmented)");
    int id;
    if (!a->free_ids.empty()) {
        id = a->free_ids.back();
        a->free_ids.pop_back();
    } else {
        id = (int)a->scenes.size();
    }
    char* blk = a->base + off;

    SceneDev D = p->D;
    D.pos[0] = (float2*)(blk + L.pos0);
    D.pos[1] = (float2*)(blk + L.pos1);
    D.vel = (float2*)(blk + L.vel);
    D.mass = (float*)(blk + L.mass);
    D.flag = (const uint32_t*)(blk + L.flag);
    D.bcv = (const float2*)(blk + L.bcv);
    D.disp = L.disp >= 0 ? (float2*)(blk + L.disp) : nullptr;
    D.force = (float2*)(blk + L.force);
    D.cflag = (uint8_t*)(blk + L.cflag);
    auto at = [&](int64_t off) -> char* { return off >= 0 ? blk + off : nullptr; };
    D.rest = (const float2*)at(L.rest);
    D.hpos[0] = (float2*)(blk + L.hpos0);
    D.hpos[1] = (float2*)(blk + L.hpos1);
    D.hvel[0] = (float2*)at(L.hvel0);
    D.hvel[1] = (float2*)at(L.hvel1);
    D.jinv = (float4*)at(L.jinv);
    D.area = (float*)at(L.area);
    D.j2[0] = (float4*)at(L.j2a);
    D.j2[1] = (float4*)at(L.j2b);
    D.stress = (float4*)at(L.stress);
    D.strain = (float4*)at(L.strain);
    D.yflag = (uint8_t*)at(L.yflag);
    D.hot.aabb = (int32_t*)(blk + L.aabb);
    for (int g = 0; g < D.hot.n_grounds; ++g)
        if (L.gdata[g] >= 0) D.ground[g].data = blk + L.gdata[g];
    // warm-up runs at the next substep launched (local step -1)
    D.s0 = (int32_t)(a->clock + 1);

    Tiling& T = p->T;
    for (auto& td : T.tiles) {
        td.base = blk;
        td.scene = id;
        td.s0 = D.s0;
    }
    if (p->pin) {  // staged on a worker thread: DMA straight from it
        PinnedBuf* pb = p->pin;
        p->pin = nullptr;
        std::memcpy((char*)pb->ptr + L.tiles, T.tiles.data(), sizeof(TileDev) * T.tiles.size());
        HIPCHK(hipMemcpyAsync(blk, pb->ptr, L.host_bytes, hipMemcpyHostToDevice, a->stream));
        HIPCHK(hipEventRecord(pb->ev, a->stream));
        pb->in_flight = true;
        pinned_put(pb);
    } else {
        std::memcpy(st->ptr, p->img.data(), (size_t)L.host_bytes);
        std::memcpy((char*)st->ptr + L.tiles, T.tiles.data(), sizeof(TileDev) * T.tiles.size());
        HIPCHK(hipMemcpyAsync(blk, st->ptr, L.host_bytes, hipMemcpyHostToDevice, a->stream));
        HIPCHK(hipEventRecord(st->ev, a->stream));
        st->busy = true;
    }
    HIPCHK(hipMemsetAsync(blk + L.host_bytes, 0, (size_t)(L.total - L.host_bytes), a->stream));

    rc = upload(a, a->d_scenes + id, &D, sizeof(SceneDev));
    if (rc) return rc;
    HIPCHK(fw_launch_prepare(a->d_scenes, (const TileDev*)(blk + L.tiles), 0, (int)T.tiles.size(), a->tune.block,
                             a->tune.tile_slots, a->stream));

    auto sh = std::make_unique<SceneHost>();
    sh->live = true;
    sh->block_off = off;
    sh->block_bytes = L.total;
    sh->tiles_dev = (const TileDev*)(blk + L.tiles);
    sh->n_nodes = p->n_nodes;
    sh->n_elems = p->n_elems;
    sh->mat_type = p->mat_type;
    sh->s0 = D.s0;
    sh->horizon = D.horizon;
    sh->tiles = T.tiles;
    sh->dev = D;
    for (int g = 0; g < FW_MAX_GROUNDS; ++g) sh->mesh[g] = p->mesh[g];
    sh->tiling = std::move(T);
    p->consumed = true;
    p->img.clear();
    p->img.shrink_to_fit();
    if (id == (int)a->scenes.size())
        a->scenes.push_back(std::move(sh));
    else
        a->scenes[id] = std::move(sh);
    a->tiles_dirty = true;
    *scene_id = id;
    return FW_OK;
}

// The arena's tuning for a caller's fw_tuning (NULL: the defaults), checked
// against the compiled kernel shapes.
// Owned nodes per thread of the compiled shape that runs a tuning: the smallest whose
// threads x NPT holds tile_nodes (so (512, 2048) with tile_nodes <= 512 takes the
// one-node-per-lane kernel, with fewer registers).
int shape_npt(const fw_tuning& t) {
    int npt = 0;
#define FW_NPT1(B, S, N, P, NS) \
    if (t.block == B && t.tile_slots == S && t.tile_nodes <= P * B && (npt == 0 || P < npt)) npt = P;
    FW_SHAPES(FW_NPT1)
#undef FW_NPT1
    return npt;
}

int tuning_of(const fw_tuning* tuning, fw_tuning* out) {
    fw_tuning t{};
    t.tile_nodes = 512;
    t.tile_slots = 1024;
    t.block = 256;
    t.grid_per_cu = 0;
    if (tuning) {
        int npt_max = 0;
#define FW_NPT(B, S, N, P, NS) \
    if (tuning->block == B && tuning->tile_slots == S) npt_max = std::max(npt_max, P);
        FW_SHAPES(FW_NPT)
#undef FW_NPT
        if (!npt_max)
            return fail(FW_E_INVALID, "unsupported (block, tile_slots); compiled shapes: (256,1024) (256,768) "
                                      "(512,1024) (256,512) (512,2048) (1024,2048)");
        if (tuning->tile_nodes < 1 || tuning->tile_nodes > npt_max * tuning->block)
            return fail(FW_E_INVALID, "tile_nodes must be in [1, " + std::to_string(npt_max * tuning->block) + "]");
        if (tuning->grid_per_cu < 0) return fail(FW_E_INVALID, "grid_per_cu must be >= 0");
        t = *tuning;
    }
    *out = t;
    return FW_OK;
}

}  // namespace

// ------------------------------------------------------------------ C ABI

extern "C" {

const char* fw_last_error(void) { return g_err_static ? g_err_static : g_err.c_str(); }

int32_t fw_abi_version(void) { return FW_ABI_VERSION; }

int32_t fw_tiling_cache_stats(int64_t* hits, int64_t* misses, int64_t* bytes, int32_t* entries) {
    std::lock_guard<std::mutex> lock(g_tc_mu);
    if (hits) *hits = g_tc_hits;
    if (misses) *misses = g_tc_misses;
    if (bytes) *bytes = g_tc_bytes;
    if (entries) *entries = (int32_t)g_tc.size();
    return FW_OK;
}

int32_t fw_tiling_cache_clear(void) {
    std::lock_guard<std::mutex> lock(g_tc_mu);
    g_tc.clear();
    g_tc_bytes = 0;
    g_tc_hits = 0;
    g_tc_misses = 0;
    return FW_OK;
}

int32_t fw_arena_create(int32_t device, int64_t capacity_bytes, const fw_tuning* tuning, fw_arena** out) {
    if (!out) return fail(FW_E_INVALID, "out is null");
    *out = nullptr;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(FW_E_INVALID, "device " + std::to_string(device) + " not present");
    HIPCHK(hipSetDevice(device));
    auto a = std::make_unique<fw_arena>();
    a->device = device;
    {
        int rc = tuning_of(tuning, &a->tune);
        if (rc) return rc;
    }
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    a->n_cu = prop.multiProcessorCount;
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    if (capacity_bytes <= 0) capacity_bytes = (int64_t)((double)fr * 0.8);
    capacity_bytes = align_up(capacity_bytes);
    HIPCHK(hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking));
    HIPCHK(hipMalloc(&a->base, capacity_bytes));
    a->capacity = capacity_bytes;
    a->free_list[0] = capacity_bytes;
    a->max_scenes = 16384;
    HIPCHK(hipMalloc(&a->d_scenes, sizeof(SceneDev) * (size_t)a->max_scenes));
    HIPCHK(hipMalloc(&a->d_ctl, sizeof(ArenaCtl)));
    ArenaCtl c0{0, 0, 0};
    HIPCHK(hipMemcpy(a->d_ctl, &c0, sizeof(c0), hipMemcpyHostToDevice));
    *out = a.release();
    return FW_OK;
}

int32_t fw_arena_destroy(fw_arena* a) {
    if (!a) return FW_OK;
    (void)hipSetDevice(a->device);
    (void)hipStreamSynchronize(a->stream);
    drop_graphs(a);
    for (auto& p : a->busy_pending) {
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }
    for (auto e : a->busy_pool) (void)hipEventDestroy(e);
    for (auto& s : a->staging) {
        (void)hipHostFree(s->ptr);
        (void)hipEventDestroy(s->ev);
    }
    (void)hipFree(a->base);
    (void)hipFree(a->d_scenes);
    (void)hipFree(a->d_tiles);
    (void)hipFree(a->d_ctl);
    (void)hipFree(a->d_flex);
    (void)hipFree(a->d_flex_items);
    (void)hipFree(a->d_flex_sched);
    (void)hipFree(a->d_rigid);
    (void)hipStreamDestroy(a->stream);
    delete a;
    return FW_OK;
}

int32_t fw_arena_get_info(fw_arena* a, fw_arena_info* out) {
    if (!a || !out) return fail(FW_E_INVALID, "null argument");
    std::memset(out, 0, sizeof(*out));
    out->capacity_bytes = a->capacity;
    out->used_bytes = a->used;
    out->global_step = a->clock;
    double algo = 0;
    for (auto& sp : a->scenes) {
        if (!sp || !sp->live) continue;
        out->n_scenes++;
        out->n_tiles += (int32_t)sp->tiles.size();
        out->n_nodes += sp->n_nodes;
        out->n_elems += sp->n_elems;
        if (sp->mat_type == FW_MAT_MISES) out->n_elems_j2 += sp->n_elems;
        out->halo_slots += sp->tiling.halo_slots;
        const double be = sp->mat_type == FW_MAT_MISES ? 64.0 : (sp->mat_type == FW_MAT_SPRING ? 12.0 : 32.0);
        algo += (double)sp->n_elems * be + (double)sp->n_nodes * 40.0;
    }
    out->algo_bytes_per_substep = algo;
    out->n_cu = a->n_cu;
    HIPCHK(hipSetDevice(a->device));
    HIPCHK(fw_substep_occupancy(a->tune.block, a->tune.tile_slots, shape_npt(a->tune), &out->tiles_per_cu));
    return FW_OK;
}

int32_t fw_scene_footprint(const fw_arena* ca, const fw_scene_desc* d, int64_t* bytes) {
    if (!ca || !bytes) return fail(FW_E_INVALID, "null argument");
    int rc = validate(d);
    if (rc) return rc;
    fw_arena* a = const_cast<fw_arena*>(ca);  // only the tiling cache changes
    std::unique_ptr<Tiling> T;
    rc = tiling_for(a, d, &T);
    if (rc) return rc;
    *bytes = layout_of(d, *T).total;
    cache_keep(a, std::move(T));
    return FW_OK;
}

int32_t fw_scene_prepare(const fw_scene_desc* d, const fw_tuning* tuning, fw_prepared** out) {
    if (!out) return fail(FW_E_INVALID, "null argument");
    *out = nullptr;
    fw_tuning tune;
    int rc = tuning_of(tuning, &tune);
    if (rc) return rc;
    std::unique_ptr<fw_prepared> P;
    rc = prepare_scene(d, tune, nullptr, &P);
    if (rc) return rc;
    *out = P.release();
    return FW_OK;
}

int32_t fw_prepared_footprint(const fw_prepared* p, int64_t* bytes) {
    if (!p || !bytes) return fail(FW_E_INVALID, "null argument");
    if (p->consumed) return fail(FW_E_INVALID, "prepared scene was already added");
    *bytes = p->L.total;
    return FW_OK;
}

int32_t fw_prepared_stage(fw_prepared* p) {
    if (!p) return fail(FW_E_INVALID, "null argument");
    if (p->consumed) return fail(FW_E_INVALID, "prepared scene was already added");
    if (p->pin) return FW_OK;
    PinnedBuf* b = pinned_get(p->L.host_bytes);
    if (!b) return FW_OK;  // no page-locked memory: the commit copies through the arena's staging
    std::memcpy(b->ptr, p->img.data(), (size_t)p->L.host_bytes);
    p->pin = b;
    std::vector<char>().swap(p->img);
    return FW_OK;
}

int32_t fw_prepared_free(fw_prepared* p) {
    if (p && p->pin) pinned_put(p->pin);
    delete p;
    return FW_OK;
}

int32_t fw_scene_add_prepared(fw_arena* a, fw_prepared* p, int32_t* scene_id) {
    if (!a || !p || !scene_id) return fail(FW_E_INVALID, "null argument");
    if (p->consumed) return fail(FW_E_INVALID, "prepared scene was already added");
    if (p->tune.tile_nodes != a->tune.tile_nodes || p->tune.tile_slots != a->tune.tile_slots ||
        p->tune.block != a->tune.block)
        return fail(FW_E_INVALID, "prepared scene was tiled for another kernel shape than this arena's");
    HIPCHK(hipSetDevice(a->device));
    return commit_scene(a, p, scene_id);
}

int32_t fw_scene_add(fw_arena* a, const fw_scene_desc* d, int32_t* scene_id) {
    if (!a || !scene_id) return fail(FW_E_INVALID, "null argument");
    HIPCHK(hipSetDevice(a->device));
    std::unique_ptr<fw_prepared> P;
    int rc = prepare_scene(d, a->tune, a, &P);
    if (rc) return rc;
    return commit_scene(a, P.get(), scene_id);
}

int32_t fw_scene_retire(fw_arena* a, int32_t id, float* x_out, float* v_out) {
    if (!a || !x_out) return fail(FW_E_INVALID, "null argument");
    SceneHost* s = get_scene(a, id);
    if (!s) return fail(FW_E_NOT_FOUND, "no scene " + std::to_string(id));
    if (s->flex_group >= 0)
        return fail(FW_E_INVALID, "scene " + std::to_string(id) + " belongs to FEM-FEM group " +
                                      std::to_string(s->flex_group) + "; remove the group first");
    HIPCHK(hipSetDevice(a->device));
    int64_t done;
    int buf;
    bool warmed;
    scene_state(a, s, &done, &buf, &warmed);
    const int64_t n = s->n_nodes, bytes = 8 * n;
    Staging* st = nullptr;
    int rc = get_staging(a, 2 * bytes, &st);
    if (rc) return rc;
    // the copies follow every substep launched so far on the stream, and precede the
    // removal below and any later upload into the freed block
    HIPCHK(hipMemcpyAsync(st->ptr, s->dev.pos[buf], bytes, hipMemcpyDeviceToHost, a->stream));
    if (v_out) HIPCHK(hipMemcpyAsync((char*)st->ptr + bytes, s->dev.vel, bytes, hipMemcpyDeviceToHost, a->stream));
    HIPCHK(hipEventRecord(st->ev, a->stream));
    st->busy = true;
    st->held = true;
    PendingRetire pr;
    pr.st = st;
    pr.n_nodes = n;
    pr.perm = s->tiling.node_perm;
    pr.x_out = x_out;
    pr.v_out = v_out;
    a->retiring.push_back(std::move(pr));
    return fw_scene_remove(a, id);
}

int32_t fw_retire_wait(fw_arena* a) {
    if (!a) return fail(FW_E_INVALID, "null arena");
    HIPCHK(hipSetDevice(a->device));
    for (auto& pr : a->retiring) {
        HIPCHK(hipEventSynchronize(pr.st->ev));
        const float2* x = (const float2*)pr.st->ptr;
        const float2* v = x + pr.n_nodes;
        for (int64_t i = 0; i < pr.n_nodes; ++i) {
            const int64_t o = pr.perm[i];
            pr.x_out[2 * o] = x[i].x;
            pr.x_out[2 * o + 1] = x[i].y;
            if (pr.v_out) {
                pr.v_out[2 * o] = v[i].x;
                pr.v_out[2 * o + 1] = v[i].y;
            }
        }
        pr.st->held = false;
    }
    a->retiring.clear();
    return FW_OK;
}

int32_t fw_scene_remove(fw_arena* a, int32_t id) {
    if (!a) return fail(FW_E_INVALID, "null arena");
    SceneHost* s = get_scene(a, id);
    if (!s) return fail(FW_E_NOT_FOUND, "no scene " + std::to_string(id));
    if (s->flex_group >= 0)
        return fail(FW_E_INVALID, "scene " + std::to_string(id) + " belongs to FEM-FEM group " +
                                      std::to_string(s->flex_group) + "; remove the group first");
    // No synchronisation: launches already in flight may still read this scene's
    // table entry and arrays, but every later write to them -- the live=0 flag
    // below, a reused block's upload, a reused id's entry -- is a copy on the
    // same stream and so lands only after those launches have finished.
    HIPCHK(hipSetDevice(a->device));
    const int32_t dead = 0;
    int rc = upload(a, &a->d_scenes[id].live, &dead, sizeof(int32_t));
    if (rc) return rc;
    arena_free(a, s->block_off, s->block_bytes);
    a->scenes[id].reset();
    a->free_ids.push_back(id);
    a->tiles_dirty = true;
    return FW_OK;
}

int32_t fw_scene_set_rigid(fw_arena* a, int32_t id, int32_t gi, const fw_ground_desc* g) {
    if (!a || !g) return fail(FW_E_INVALID, "null argument");
    SceneHost* s = get_scene(a, id);
    if (!s) return fail(FW_E_NOT_FOUND, "no scene " + std::to_string(id));
    if (gi < 0 || gi >= s->dev.hot.n_grounds ||
        (s->dev.ground[gi].kind != FW_GROUND_RIGID && s->dev.ground[gi].kind != FW_GROUND_RIGID_MESH_DEV))
        return fail(FW_E_INVALID, "ground " + std::to_string(gi) + " of scene " + std::to_string(id) +
                                      " is not a rigid record");
    if (g->kind != FW_GROUND_RIGID) return fail(FW_E_INVALID, "descriptor kind must be FW_GROUND_RIGID");
    int rc = validate_rigid(*g);
    if (rc) return rc;
    GroundDev G = s->dev.ground[gi];
    if (g->rigid_type != G.nx || g->rigid_tied != G.ny)
        return fail(FW_E_INVALID, "a rigid record keeps its type and tied flag");
    pack_aabb(G, *g);
    pack_rigid(G, *g);
    HIPCHK(hipSetDevice(a->device));
    if (G.kind == FW_GROUND_RIGID_MESH_DEV) {
        // the header always, the arrays mesh_update selects; stream-ordered behind
        // the substeps launched so far
        const MeshHost& mh = s->mesh[gi];
        if (g->mesh_nodes != mh.nodes || g->mesh_edges != mh.edges || g->mesh_cap_reg != mh.cap_reg ||
            g->mesh_cap_cells != mh.cap_cells || g->mesh_cap_pairs != mh.cap_pairs)
            return fail(FW_E_INVALID, "a mesh rigid record keeps its node and edge counts and hash capacities");
        char* b = static_cast<char*>(const_cast<void*>(G.data));
        const MeshLayout& M = mh.ml;
        const MeshDev H = mesh_header(*g, M, mh.n_items);
        rc = upload(a, b + M.hdr, &H, (int64_t)sizeof(H));
        if (!rc && (g->mesh_update & 1)) rc = upload(a, b + M.coords, g->mesh_coords, 8 * (int64_t)g->mesh_nodes);
        if (!rc && (g->mesh_update & 2)) {
            rc = upload(a, b + M.elb, g->mesh_elb, 8 * (int64_t)g->mesh_edges);
            if (!rc) rc = upload(a, b + M.eub, g->mesh_eub, 8 * (int64_t)g->mesh_edges);
            if (!rc) rc = upload(a, b + M.margin, g->mesh_margin, 4 * (int64_t)g->mesh_edges);
        }
        if (!rc && (g->mesh_update & 4)) {
            rc = upload(a, b + M.reg, g->mesh_reg, 8 * (int64_t)g->mesh_nreg);
            if (!rc) rc = upload(a, b + M.cell, g->mesh_cell, 8 * (int64_t)g->mesh_tc);
            if (!rc) rc = upload(a, b + M.sorted, g->mesh_sorted, 4 * (int64_t)g->mesh_npairs);
        }
        if (rc) return rc;
    }
    rc = upload(a, &a->d_scenes[id].ground[gi], &G, (int64_t)sizeof(GroundDev));
    if (rc) return rc;
    s->dev.ground[gi] = G;
    return FW_OK;
}

int32_t fw_scene_progress(fw_arena* a, int32_t id, int64_t* steps_done, int32_t* finished) {
    SceneHost* s = a ? get_scene(a, id) : nullptr;
    if (!s) return fail(FW_E_NOT_FOUND, "no scene " + std::to_string(id));
    int64_t done;
    int buf;
    bool warmed;
    scene_state(a, s, &done, &buf, &warmed);
    if (steps_done) *steps_done = done;
    if (finished) *finished = (s->horizon != INT32_MAX && done >= s->horizon) ? 1 : 0;
    return FW_OK;
}

// GPU-busy accounting: an event pair around each stepping call, when enabled.
static int busy_event(fw_arena* a, hipEvent_t* e) {
    if (!a->busy_pool.empty()) {
        *e = a->busy_pool.back();
        a->busy_pool.pop_back();
        return FW_OK;
    }
    HIPCHK(hipEventCreate(e));
    return FW_OK;
}

static int busy_begin(fw_arena* a, hipEvent_t* e0) {
    if (!a->track_busy) return FW_OK;
    int rc = busy_event(a, e0);
    if (rc) return rc;
    HIPCHK(hipEventRecord(*e0, a->stream));
    return FW_OK;
}

static int busy_end(fw_arena* a, hipEvent_t e0) {
    if (!a->track_busy) return FW_OK;
    hipEvent_t e1;
    int rc = busy_event(a, &e1);
    if (rc) return rc;
    HIPCHK(hipEventRecord(e1, a->stream));
    a->busy_pending.push_back({e0, e1});
    // resolve completed pairs now and then, so the pending list stays short
    if (a->busy_pending.size() >= 64) {
        size_t k = 0;
        for (; k < a->busy_pending.size(); ++k) {
            auto& p = a->busy_pending[k];
            if (hipEventQuery(p.second) != hipSuccess) break;
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, p.first, p.second));
            a->busy_ms += ms;
            a->busy_calls += 1;
            a->busy_pool.push_back(p.first);
            a->busy_pool.push_back(p.second);
        }
        a->busy_pending.erase(a->busy_pending.begin(), a->busy_pending.begin() + (long)k);
    }
    return FW_OK;
}

int32_t fw_arena_track_busy(fw_arena* a, int32_t enable) {
    if (!a) return fail(FW_E_INVALID, "null arena");
    a->track_busy = enable != 0;
    return FW_OK;
}

int32_t fw_wait_launches(fw_arena* a, int32_t keep) {
    if (!a || keep < 0) return fail(FW_E_INVALID, "bad arguments");
    if (!a->track_busy) return fail(FW_E_INVALID, "fw_wait_launches needs busy tracking (fw_arena_track_busy)");
    HIPCHK(hipSetDevice(a->device));
    const size_t n = a->busy_pending.size();
    if (n > (size_t)keep) HIPCHK(hipEventSynchronize(a->busy_pending[n - 1 - (size_t)keep].second));
    return FW_OK;
}

int32_t fw_arena_busy(fw_arena* a, double* busy_ms, int64_t* calls) {
    if (!a || !busy_ms) return fail(FW_E_INVALID, "bad arguments");
    HIPCHK(hipSetDevice(a->device));
    for (auto& p : a->busy_pending) {
        HIPCHK(hipEventSynchronize(p.second));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, p.first, p.second));
        a->busy_ms += ms;
        a->busy_calls += 1;
        a->busy_pool.push_back(p.first);
        a->busy_pool.push_back(p.second);
    }
    a->busy_pending.clear();
    *busy_ms = a->busy_ms;
    if (calls) *calls = a->busy_calls;
    return FW_OK;
}

// The substep launches of one substep over the whole tile table: one per element
// kind present (triangle tiles, then spring tiles), each covering its kind's tiles
// as the control block counts them at run time.
static hipError_t launch_all(fw_arena* a, int k, float dt_override, float damping_override, uint32_t f, int grid_tri,
                      int grid_spr, char* flex_scr = nullptr, bool sched_hash = false) {
    if (sched_hash && a->n_flex_live > 0) {  // scheduled hash rebuilds before this substep's contact pass
        hipError_t e = fw_launch_flex_hash(a->d_flex, a->n_flex_live, a->d_scenes, a->d_ctl, -1, k, 1, a->stream);
        if (e != hipSuccess) return e;
    }
    if (a->n_rigid_live > 0) {  // FEM-rigid forces before the substep that adds them
        hipError_t e = fw_launch_rigid_pass(a->d_scenes, a->d_rigid, a->n_rigid_live, a->d_ctl, k, f, a->stream);
        if (e != hipSuccess) return e;
    }
    if (a->n_flex_live > 0) {  // FEM-FEM contact lists before the substep that adds them
        hipError_t e = fw_launch_flex_contact(a->d_flex, a->n_flex_live, a->d_flex_items, a->n_flex_items, a->d_scenes,
                                              a->d_ctl, k, f, flex_scr, a->stream);
        if (e != hipSuccess) return e;
    }
    if (grid_tri > 0) {
        hipError_t e = fw_launch_substep(a->d_scenes, a->d_tiles, a->d_ctl, k, 0, -1, dt_override, damping_override, f,
                                         grid_tri, a->tune.block, a->tune.tile_slots, shape_npt(a->tune), 0, a->stream);
        if (e != hipSuccess) return e;
    }
    if (grid_spr > 0)
        return fw_launch_substep(a->d_scenes, a->d_tiles, a->d_ctl, k, 0, -1, dt_override, damping_override, f,
                                 grid_spr, a->tune.block, a->tune.tile_slots, shape_npt(a->tune), 1, a->stream);
    return hipSuccess;
}

// Whether some group has a hash rebuild scheduled before global step `step`.
static bool sched_at(const fw_arena* a, int64_t step) {
    for (const auto& F : a->flex)
        if (F)
            for (int r = 0; r < FW_FLEX_SCHED; ++r)
                if (F->sched[r] == step) return true;
    return false;
}

// After n substeps from the current clock: the scheduled rebuilds they ran are done.
static void sched_consume(fw_arena* a, int32_t n) {
    for (auto& F : a->flex)
        if (F)
            for (int r = 0; r < FW_FLEX_SCHED; ++r)
                if (F->sched[r] != INT32_MIN && F->sched[r] < a->clock + n) {
                    F->hash_built = true;
                    F->sched[r] = INT32_MIN;
                }
}

static int step_impl(fw_arena* a, int32_t n, float dt_override, float damping_override, uint32_t flags) {
    int rc = ensure_tiles(a);
    if (rc) return rc;
    rc = ensure_flex(a);
    if (rc) return rc;
    if ((flags & FW_STEP_PER_SCENE) && a->n_flex_live > 0)
        return fail(FW_E_UNSUPPORTED, "FW_STEP_PER_SCENE with FEM-FEM contact groups resident");
    const int nt = (int)a->tile_table.size();
    const int n_tri = a->n_tri_tiles, n_spr = nt - n_tri;
    for (int k = 0; k < n; ++k) {
        uint32_t f = flags & ~FW_STEP_PER_SCENE;
        if (k != n - 1) f &= ~FW_STEP_DIAG;
        if (nt > 0) {
            if (flags & FW_STEP_PER_SCENE) {
                HIPCHK(fw_launch_rigid_pass(a->d_scenes, a->d_rigid, a->n_rigid_live, a->d_ctl, k, f, a->stream));
                for (auto& sp : a->scenes) {
                    if (!sp || !sp->live) continue;
                    const int cnt = (int)sp->tiles.size();
                    HIPCHK(fw_launch_substep(a->d_scenes, a->d_tiles, a->d_ctl, k, sp->tile_base, cnt, dt_override, damping_override, f,
                                             launch_grid(a, cnt), a->tune.block, a->tune.tile_slots, shape_npt(a->tune),
                                             sp->mat_type == FW_MAT_SPRING ? 1 : 0, a->stream));
                }
            } else {
                HIPCHK(launch_all(a, k, dt_override, damping_override, f, n_tri ? launch_grid(a, n_tri) : 0,
                                  n_spr ? launch_grid(a, n_spr) : 0, nullptr, sched_at(a, a->clock + k)));
            }
        }
    }
    HIPCHK(fw_launch_advance_clock(a->d_ctl, n, a->stream));
    sched_consume(a, n);
    a->clock += n;
    return FW_OK;
}

int32_t fw_step(fw_arena* a, int32_t n, float dt_override, float damping_override, uint32_t flags) {
    if (!a || n < 0) return fail(FW_E_INVALID, "bad arguments");
    HIPCHK(hipSetDevice(a->device));
    if (n == 0) {  // no substep: only upload the tables that scene / group changes left pending
        int rc = ensure_tiles(a);
        return rc ? rc : ensure_flex(a);
    }
    hipEvent_t e0 = nullptr;
    int rc = ensure_tiles(a);  // table uploads go ahead of the busy bracket
    if (rc) return rc;
    rc = busy_begin(a, &e0);
    if (rc) return rc;
    rc = step_impl(a, n, dt_override, damping_override, flags);
    if (rc) return rc;
    return busy_end(a, e0);
}

#define HIPCHK_NOALLOC(expr)                                                \
    do {                                                                    \
        hipError_t e_ = (expr);                                             \
        if (e_ != hipSuccess) return fail_static(FW_E_HIP, hipGetErrorString(e_)); \
    } while (0)

int32_t fw_step_noalloc(fw_arena* a, int32_t n, float dt_override, float damping_override, uint32_t flags) {
    if (!a || n < 0) return fail_static(FW_E_INVALID, "fw_step_noalloc: bad arguments");
    if (n == 0) return FW_OK;
    if (a->tiles_dirty || a->flex_dirty)
        return fail_static(FW_E_INVALID, "fw_step_noalloc: scenes or groups changed since the last fw_step; call "
                                         "fw_step(a, 0, ...) once to upload the tables");
    if ((flags & FW_STEP_PER_SCENE) && a->n_flex_live > 0)
        return fail_static(FW_E_UNSUPPORTED, "FW_STEP_PER_SCENE with FEM-FEM contact groups resident");
    HIPCHK_NOALLOC(hipSetDevice(a->device));
    char* scr = nullptr;
    if (a->n_flex_live > 0) {
        if (a->flex_scratch_bytes > (int64_t)FW_NOALLOC_SCRATCH_BYTES)
            return fail_static(FW_E_CAPACITY, "fw_step_noalloc: the contact groups need more per-substep scratch "
                                              "than FW_NOALLOC_SCRATCH_BYTES");
        HIPCHK_NOALLOC(fw_flex_noalloc_scratch(a->device, &scr));
        HIPCHK_NOALLOC(fw_launch_flex_scratch_init(a->d_flex, a->n_flex_live, scr, a->stream));
    }
    const int nt = (int)a->tile_table.size();
    const int n_tri = a->n_tri_tiles, n_spr = nt - n_tri;
    for (int k = 0; k < n; ++k) {
        uint32_t f = flags & ~FW_STEP_PER_SCENE;
        if (k != n - 1) f &= ~FW_STEP_DIAG;
        if (nt <= 0) continue;
        if (flags & FW_STEP_PER_SCENE) {
            HIPCHK_NOALLOC(fw_launch_rigid_pass(a->d_scenes, a->d_rigid, a->n_rigid_live, a->d_ctl, k, f, a->stream));
            for (auto& sp : a->scenes) {
                if (!sp || !sp->live) continue;
                const int cnt = (int)sp->tiles.size();
                HIPCHK_NOALLOC(fw_launch_substep(a->d_scenes, a->d_tiles, a->d_ctl, k, sp->tile_base, cnt, dt_override,
                                                 damping_override, f, launch_grid(a, cnt), a->tune.block,
                                                 a->tune.tile_slots, shape_npt(a->tune),
                                                 sp->mat_type == FW_MAT_SPRING ? 1 : 0, a->stream));
            }
        } else {
            HIPCHK_NOALLOC(launch_all(a, k, dt_override, damping_override, f, n_tri ? launch_grid(a, n_tri) : 0,
                                      n_spr ? launch_grid(a, n_spr) : 0, scr, sched_at(a, a->clock + k)));
        }
    }
    HIPCHK_NOALLOC(fw_launch_advance_clock(a->d_ctl, n, a->stream));
    sched_consume(a, n);
    a->clock += n;
    return FW_OK;
}

int32_t fw_tiling_check(const fw_scene_desc* d, int32_t tile_nodes, int32_t tile_slots, int64_t* stats) {
    // Host-only: build the tiling a scene would get and verify it, without touching the GPU.
    //   every element owned by exactly one tile, owned ranges contiguous and disjoint;
    //   slots (owned + halo) and LDS nodes (owned + halo) within capacity; each slot
    //   record's local corner ids name the element's nodes; every tile's node CSR lists
    //   exactly the node's (element, corner) incidences in ascending caller element id;
    //   every halo copy is pushed exactly once, by its node's owner.
    int rc = validate(d);
    if (rc) return rc;
    Tiling T;
    const int node_cap = node_cap_of_slots(tile_slots, d->mat_type == FW_MAT_SPRING);
    if (node_cap == 0) return fail(FW_E_INVALID, "tile_slots must be one of the compiled shapes (512, 768, 1024, 2048)");
    rc = build_tiling(d, tile_nodes, tile_slots, node_cap, T);
    if (rc) return rc;
    const int N = d->n_nodes, E = d->n_elems;
    const int K = corners_of(d);
    std::vector<int> owned(E, 0);
    int64_t max_slots = 0, entries = 0;
    for (const TileDev& td : T.tiles) {
        if (td.n_own + td.n_halo > tile_slots || td.n_nodes > tile_nodes || td.n_nodes + td.n_hnodes > node_cap)
            return fail(FW_E_INVALID, "tile over capacity");
        max_slots = std::max<int64_t>(max_slots, td.n_own + td.n_halo);
        for (int i = td.elem0; i < td.elem0 + td.n_own; ++i) owned[T.elem_perm[i]]++;
        // storage node id of each local id
        std::vector<int32_t> local(td.n_nodes + td.n_hnodes);
        for (int i = 0; i < td.n_nodes; ++i) local[i] = td.node0 + i;
        for (int h = 0; h < td.n_hnodes; ++h) {
            local[td.n_nodes + h] = T.hnode[td.hnode_off + h];
            if (local[td.n_nodes + h] >= td.node0 && local[td.n_nodes + h] < td.node0 + td.n_nodes)
                return fail(FW_E_INVALID, "halo node list holds an owned node");
        }
        // invert the slot records: LDS position -> (caller element, corner)
        const int nslots = td.n_own + td.n_halo;
        const uint16_t* cp = T.csr_ptr.data() + td.ptr_off;
        std::vector<int64_t> at(cp[td.n_nodes], -1);
        int prev_halo = -1;
        for (int sl = 0; sl < nslots; ++sl) {
            const uint32_t* r = T.rec.data() + 2 * ((size_t)td.rec_off + sl);
            int eint = td.elem0 + sl;
            if (sl >= td.n_own) {
                eint = T.hel[(size_t)td.hel_off + sl - td.n_own];
                if (eint <= prev_halo || (eint >= td.elem0 && eint < td.elem0 + td.n_own))
                    return fail(FW_E_INVALID, "halo slots not ascending or name an owned element");
                prev_halo = eint;
            }
            const int ecaller = T.elem_perm[eint];
            if (K == 3 && (r[1] >> 31)) return fail(FW_E_INVALID, "slot record has stray bits");
            if (K == 2 && std::memcmp(&r[1], d->rest_len + ecaller, 4) != 0)
                return fail(FW_E_INVALID, "spring slot record holds the wrong rest length");
            const uint32_t words[3] = {r[0] & 0xFFFFu, r[0] >> 16, r[1] & 0xFFFFu};
            for (int c = 0; c < K; ++c) {
                const uint32_t lidc = words[c] & FW_REC_ID_MASK;
                if (lidc >= local.size() || T.node_perm[local[lidc]] != d->conn[K * ecaller + c])
                    return fail(FW_E_INVALID, "slot record names the wrong node");
                if ((int)lidc >= td.n_nodes) continue;  // halo node: its owner tile assembles it
                const uint32_t rk = (words[c] >> FW_REC_ID_BITS) | (K == 3 ? ((r[1] >> (16 + 5 * c)) & 31u) << 5 : 0u);
                const uint32_t pos = (uint32_t)cp[lidc] + rk;
                if (pos >= (uint32_t)cp[lidc + 1]) return fail(FW_E_INVALID, "corner rank beyond its node's list");
                if (pos >= at.size() || at[pos] != -1) return fail(FW_E_INVALID, "slot destinations collide");
                at[pos] = K * (int64_t)ecaller + c;
            }
        }
        for (int i = 0; i < td.n_nodes; ++i) {
            const int p = td.node0 + i;
            const int n = T.node_perm[p];
            std::vector<int64_t> want, got;
            for (int e = 0; e < E; ++e)  // brute force: fine for the small meshes this is used on
                for (int c = 0; c < K; ++c)
                    if (d->conn[K * e + c] == n) want.push_back(K * (int64_t)e + c);
            for (int j = cp[i]; j < cp[i + 1]; ++j) got.push_back(at[j]);
            entries += (int64_t)got.size();
            if (want != got) return fail(FW_E_INVALID, "incidence order of node " + std::to_string(n) + " is wrong");
        }
    }
    // push lists: each halo copy written exactly once, by the tile owning its node,
    // from that node's local id
    std::vector<int> written(T.hnode.size(), 0);
    for (const TileDev& td : T.tiles) {
        for (int q = 0; q < td.n_push; ++q) {
            const int lo = T.push[2 * ((size_t)td.push_off + q)], j = T.push[2 * ((size_t)td.push_off + q) + 1];
            if (lo < 0 || lo >= td.n_nodes || j < 0 || j >= (int)T.hnode.size())
                return fail(FW_E_INVALID, "push entry out of range");
            if (T.hnode[j] != td.node0 + lo) return fail(FW_E_INVALID, "push entry copies the wrong node");
            written[j]++;
        }
    }
    for (size_t j = 0; j < written.size(); ++j)
        if (written[j] != 1) return fail(FW_E_INVALID, "halo copy " + std::to_string(j) + " written " + std::to_string(written[j]) + " times");
    for (int e = 0; e < E; ++e)
        if (owned[e] != 1) return fail(FW_E_INVALID, "element " + std::to_string(e) + " owned " + std::to_string(owned[e]) + " times");
    for (int i = 0; i < N; ++i)
        if (T.node_perm[T.node_inv[i]] != i) return fail(FW_E_INVALID, "node permutation is not a bijection");
    if (stats) {
        stats[0] = (int64_t)T.tiles.size();
        stats[1] = T.halo_slots;
        stats[2] = max_slots;
        stats[3] = entries;
    }
    return FW_OK;
}

int32_t fw_scene_reset_gate(fw_arena* a, int32_t id) {
    if (!a) return fail(FW_E_INVALID, "null arena");
    SceneHost* s = get_scene(a, id);
    if (!s) return fail(FW_E_NOT_FOUND, "no scene " + std::to_string(id));
    HIPCHK(hipSetDevice(a->device));
    // next launch runs local step clock - s0
    s->dev.hot.gate_base = (int32_t)(a->clock - s->s0);
    return upload(a, &a->d_scenes[id].hot.gate_base, &s->dev.hot.gate_base, sizeof(int32_t));
}

// The instantiated graph of n substeps with `flags` for the current tables, captured
// on first use.
static int graph_for(fw_arena* a, int32_t n, uint32_t flags, hipGraphExec_t* out) {
    if (!a || n <= 0) return fail(FW_E_INVALID, "bad arguments");
    if (flags & FW_STEP_PER_SCENE) return fail(FW_E_INVALID, "graph replay covers the batched launch only");
    HIPCHK(hipSetDevice(a->device));
    int rc = ensure_tiles(a);
    if (rc) return rc;
    rc = ensure_flex(a);
    if (rc) return rc;
    const int nt = (int)a->tile_table.size();
    const int grid = launch_grid(a, std::max(nt, 1));
    // the element kinds present decide which launches a frame holds; the graph's grids
    // cover the table's capacity, so scenes can come and go without a new capture
    const int kinds = (a->n_tri_tiles > 0 || nt == 0 ? 1 : 0) | (nt > a->n_tri_tiles ? 2 : 0);
    if (grid > a->graph_grid || kinds != a->graph_kinds) drop_graphs(a);
    hipGraphExec_t exec = nullptr;
    for (auto& g : a->graphs)
        if (g.n == n && g.flags == flags) exec = g.exec;
    if (!exec) {
        if (a->graphs.size() >= 8) {  // keep the cache small: the oldest entry goes
            (void)hipGraphExecDestroy(a->graphs.front().exec);
            a->graphs.erase(a->graphs.begin());
        }
        const int cap_grid = std::max(grid, launch_grid(a, a->tile_cap));
        hipGraph_t g;
        std::lock_guard<std::mutex> cap(g_capture_mu);
        HIPCHK(hipStreamBeginCapture(a->stream, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < n; ++k) {
            uint32_t f = flags;
            if (k != n - 1) f &= ~FW_STEP_DIAG;
            hipError_t e = launch_all(a, k, -1.0f, -1.0f, f, (kinds & 1) ? cap_grid : 0, (kinds & 2) ? cap_grid : 0,
                                      nullptr, a->n_flex_live > 0);
            if (e != hipSuccess) {
                (void)hipStreamEndCapture(a->stream, &g);
                return fail(FW_E_HIP, std::string("capture launch: ") + hipGetErrorString(e));
            }
        }
        hipError_t e = fw_launch_advance_clock(a->d_ctl, n, a->stream);
        hipError_t e2 = hipStreamEndCapture(a->stream, &g);
        if (e != hipSuccess || e2 != hipSuccess) return fail(FW_E_HIP, "graph capture failed");
        hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) return fail(FW_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
        if (a->graphs.empty()) {
            a->graph_grid = cap_grid;
            a->graph_kinds = kinds;
        }
        a->graphs.push_back({exec, n, flags});
    }
    *out = exec;
    return FW_OK;
}

int32_t fw_step_graph_prepare(fw_arena* a, int32_t n, uint32_t flags) {
    hipGraphExec_t exec = nullptr;
    return graph_for(a, n, flags, &exec);
}

int32_t fw_step_graph(fw_arena* a, int32_t n, uint32_t flags) {
    hipGraphExec_t exec = nullptr;
    int rc = graph_for(a, n, flags, &exec);
    if (rc) return rc;
    hipEvent_t e0 = nullptr;
    rc = busy_begin(a, &e0);
    if (rc) return rc;
    HIPCHK(hipGraphLaunch(exec, a->stream));
    rc = busy_end(a, e0);
    if (rc) return rc;
    sched_consume(a, n);
    a->clock += n;
    return FW_OK;
}

int32_t fw_sync(fw_arena* a) {
    if (!a) return fail(FW_E_INVALID, "null arena");
    HIPCHK(hipSetDevice(a->device));
    HIPCHK(hipStreamSynchronize(a->stream));
    return FW_OK;
}

static int field_geometry(const SceneHost* s, int field, bool* per_node, int* width, int* elsize) {
    *per_node = true;
    *width = 2;
    *elsize = 4;
    switch (field) {
        case FW_FIELD_X: case FW_FIELD_V: case FW_FIELD_U: case FW_FIELD_FORCE: break;
        case FW_FIELD_MASS: *width = 1; break;
        case FW_FIELD_CONTACT: *width = 1; *elsize = 1; break;
        case FW_FIELD_STRESS: case FW_FIELD_STRAIN: case FW_FIELD_EPS_P: *per_node = false; *width = 3; break;
        case FW_FIELD_EQPS: case FW_FIELD_AREA: *per_node = false; *width = 1; break;
        case FW_FIELD_JINV: *per_node = false; *width = 4; break;
        case FW_FIELD_YIELD: *per_node = false; *width = 1; *elsize = 1; break;
        case FW_FIELD_AABB: *per_node = false; *width = 4; break;
        default: return fail(FW_E_INVALID, "unknown field " + std::to_string(field));
    }
    (void)s;
    return FW_OK;
}

int32_t fw_scene_read(fw_arena* a, int32_t id, int32_t field, void* out, int64_t out_bytes) {
    if (!a || !out) return fail(FW_E_INVALID, "null argument");
    SceneHost* s = get_scene(a, id);
    if (!s) return fail(FW_E_NOT_FOUND, "no scene " + std::to_string(id));
    HIPCHK(hipSetDevice(a->device));
    bool per_node;
    int width, elsize;
    int rc = field_geometry(s, field, &per_node, &width, &elsize);
    if (rc) return rc;
    const SceneDev& D = s->dev;
    int64_t done;
    int buf;
    bool warmed;
    scene_state(a, s, &done, &buf, &warmed);
    HIPCHK(hipStreamSynchronize(a->stream));
    if (field == FW_FIELD_AABB) {
        if (out_bytes < 16) return fail(FW_E_INVALID, "output buffer too small");
        // the AABB at the current x was written by the last completed local step, in the
        // sub-slots of one buffer
        const int kk = s->dev.hot.aabb_k;
        std::vector<int32_t> enc((size_t)3 * kk * FW_AABB_SUB_INTS);
        HIPCHK(hipMemcpy(enc.data(), s->dev.hot.aabb, enc.size() * 4, hipMemcpyDeviceToHost));
        const int slot = warmed ? (int)(((done % 3) + 3) % 3) : 0;
        int32_t r[4] = {fw_f2o(1e30f), fw_f2o(1e30f), fw_f2o(-1e30f), fw_f2o(-1e30f)};
        for (int j = 0; j < kk; ++j) {
            const int32_t* e = enc.data() + ((size_t)slot * kk + j) * FW_AABB_SUB_INTS;
            r[0] = std::min(r[0], e[0]);
            r[1] = std::min(r[1], e[1]);
            r[2] = std::max(r[2], e[2]);
            r[3] = std::max(r[3], e[3]);
        }
        float* o = (float*)out;
        for (int k = 0; k < 4; ++k) o[k] = fw_o2f(r[k]);
        return FW_OK;
    }
    const int64_t count = per_node ? s->n_nodes : s->n_elems;
    const int64_t need = count * width * elsize;
    if (out_bytes < need) return fail(FW_E_INVALID, "output buffer too small: need " + std::to_string(need));
    const void* src = nullptr;
    int dev_width = width;  // components stored per item on the device
    switch (field) {
        case FW_FIELD_X: src = D.pos[buf]; break;
        case FW_FIELD_V: src = D.vel; break;
        case FW_FIELD_U: src = D.disp; break;
        case FW_FIELD_FORCE: src = D.force; break;
        case FW_FIELD_MASS: src = D.mass; break;
        case FW_FIELD_CONTACT: src = D.cflag; break;
        case FW_FIELD_STRESS: src = D.stress; dev_width = 4; break;
        case FW_FIELD_STRAIN: src = D.strain; dev_width = 4; break;
        case FW_FIELD_EPS_P: src = D.j2[buf]; dev_width = 4; break;
        case FW_FIELD_EQPS: src = D.j2[buf]; dev_width = 4; break;
        case FW_FIELD_AREA: src = D.area; break;
        case FW_FIELD_JINV: src = D.jinv; break;
        case FW_FIELD_YIELD: src = D.yflag; break;
    }
    if (!src) {  // field not stored for this scene (no displacement tracking / not J2): zeros
        std::memset(out, 0, need);
        return FW_OK;
    }
    std::vector<char> tmp((size_t)(count * dev_width * elsize));
    HIPCHK(hipMemcpy(tmp.data(), src, tmp.size(), hipMemcpyDeviceToHost));
    const std::vector<int32_t>& perm = per_node ? s->tiling.node_perm : s->tiling.elem_perm;
    char* o = (char*)out;
    const int comp0 = field == FW_FIELD_EQPS ? 3 : 0;
    for (int64_t i = 0; i < count; ++i) {
        const char* srow = tmp.data() + i * dev_width * elsize + comp0 * elsize;
        std::memcpy(o + (int64_t)perm[i] * width * elsize, srow, (size_t)width * elsize);
    }
    return FW_OK;
}

int32_t fw_scene_write(fw_arena* a, int32_t id, int32_t field, const void* in, int64_t in_bytes) {
    if (!a || !in) return fail(FW_E_INVALID, "null argument");
    SceneHost* s = get_scene(a, id);
    if (!s) return fail(FW_E_NOT_FOUND, "no scene " + std::to_string(id));
    if (field != FW_FIELD_X && field != FW_FIELD_V)
        return fail(FW_E_UNSUPPORTED, "only positions and velocities can be written");
    HIPCHK(hipSetDevice(a->device));
    const int64_t need = (int64_t)s->n_nodes * 8;
    if (in_bytes < need) return fail(FW_E_INVALID, "input buffer too small");
    int64_t done;
    int buf;
    bool warmed;
    scene_state(a, s, &done, &buf, &warmed);
    std::vector<float2> tmp(s->n_nodes);
    const float* src = (const float*)in;
    for (int i = 0; i < s->n_nodes; ++i) {
        int o = s->tiling.node_perm[i];
        tmp[i] = make_float2(src[2 * o], src[2 * o + 1]);
    }
    HIPCHK(hipStreamSynchronize(a->stream));
    void* dst = field == FW_FIELD_X ? (void*)s->dev.pos[buf] : (void*)s->dev.vel;
    HIPCHK(hipMemcpy(dst, tmp.data(), need, hipMemcpyHostToDevice));
    // the halo copies the next substep reads hold the same positions (velocities of
    // spring scenes: their force reads both ends' velocities)
    float2* hdst = field == FW_FIELD_X ? s->dev.hpos[buf] : s->dev.hvel[buf];
    if (hdst && !s->tiling.hnode.empty()) {
        std::vector<float2> hp(s->tiling.hnode.size());
        for (size_t j = 0; j < hp.size(); ++j) hp[j] = tmp[s->tiling.hnode[j]];
        HIPCHK(hipMemcpy(hdst, hp.data(), 8 * hp.size(), hipMemcpyHostToDevice));
    }
    return FW_OK;
}

int32_t fw_scene_mass_scaling(fw_arena* a, int32_t id, float dt_target) {
    if (!a) return fail(FW_E_INVALID, "null arena");
    SceneHost* s = get_scene(a, id);
    if (!s) return fail(FW_E_NOT_FOUND, "no scene " + std::to_string(id));
    if (!(dt_target > 0.0f)) return fail(FW_E_INVALID, "dt_target must be > 0");
    HIPCHK(hipSetDevice(a->device));
    int64_t done;
    int buf;
    bool warmed;
    scene_state(a, s, &done, &buf, &warmed);
    // enqueued behind every substep launched so far, so it sees the current positions
    HIPCHK(fw_launch_mass_scaling(a->d_scenes, s->tiles_dev, (int)s->tiles.size(), buf, dt_target, a->tune.block,
                                  a->tune.tile_slots, a->stream));
    return FW_OK;
}

int32_t fw_time_steps(fw_arena* a, int32_t n, int32_t mode, double* total_ms, double* kernel_ms) {
    if (!a || n <= 0 || !total_ms || !kernel_ms) return fail(FW_E_INVALID, "bad arguments");
    HIPCHK(hipSetDevice(a->device));
    int rc = ensure_tiles(a);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(a->stream));
    hipEvent_t t0, t1;
    HIPCHK(hipEventCreate(&t0));
    HIPCHK(hipEventCreate(&t1));
    rc = ensure_flex(a);
    if (rc) return rc;
    double ksum = 0.0;
    if (mode == 0) {
        // per-launch events bracket each substep kernel on the launch stream
        std::vector<hipEvent_t> ev(2 * (size_t)n);
        for (auto& e : ev) HIPCHK(hipEventCreate(&e));
        const int nt = (int)a->tile_table.size();
        const int n_tri = a->n_tri_tiles, n_spr = nt - n_tri;
        HIPCHK(hipEventRecord(t0, a->stream));
        for (int k = 0; k < n; ++k) {
            HIPCHK(hipEventRecord(ev[2 * k], a->stream));
            HIPCHK(launch_all(a, k, -1.0f, -1.0f, 0, n_tri ? launch_grid(a, n_tri) : 0, n_spr ? launch_grid(a, n_spr) : 0));
            HIPCHK(hipEventRecord(ev[2 * k + 1], a->stream));
        }
        HIPCHK(fw_launch_advance_clock(a->d_ctl, n, a->stream));
        a->clock += n;
        HIPCHK(hipEventRecord(t1, a->stream));
        HIPCHK(hipEventSynchronize(t1));
        for (int k = 0; k < n; ++k) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
            ksum += ms;
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        *kernel_ms = ksum / n;
    } else {
        HIPCHK(hipEventRecord(t0, a->stream));
        if (mode == 1)
            rc = fw_step_graph(a, n, 0);
        else
            rc = step_impl(a, n, -1.0f, -1.0f, FW_STEP_PER_SCENE);
        if (rc) return rc;
        HIPCHK(hipEventRecord(t1, a->stream));
        HIPCHK(hipEventSynchronize(t1));
    }
    float tot = 0;
    HIPCHK(hipEventElapsedTime(&tot, t0, t1));
    *total_ms = tot;
    if (mode != 0) *kernel_ms = tot / n;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    return FW_OK;
}

int32_t fw_flex_add(fw_arena* a, const fw_flex_desc* d, int32_t* group_id) {
    if (!a || !d || !group_id) return fail(FW_E_INVALID, "bad arguments");
    const int B = d->n_bodies, S = d->n_segs, W = d->n_items, P = d->n_pairs;
    if (B < 2 || S < 1 || W < 0 || P < 1) return fail(FW_E_INVALID, "a group needs two bodies, segments and pairs");
    if (!d->scene_ids || !d->body_seg_off || !d->seg_node || !d->pair_body || !d->body_bnodes_off || !d->body_bnodes ||
        (W > 0 && (!d->item_body || !d->item_node || !d->item_target || !d->item_pair || !d->item_penalty ||
                   !d->item_friction)))
        return fail(FW_E_INVALID, "group arrays are required");
    if (S >= (1 << 29)) return fail(FW_E_UNSUPPORTED, "a group holds at most 2^29 - 1 boundary segments");
    if (d->max_cells < 1 || d->max_query < 1 || d->max_pairs < 1) return fail(FW_E_INVALID, "hash limits must be >= 1");
    HIPCHK(hipSetDevice(a->device));

    std::vector<SceneHost*> sc(B);
    for (int b = 0; b < B; ++b) {
        sc[b] = get_scene(a, d->scene_ids[b]);
        if (!sc[b]) return fail(FW_E_NOT_FOUND, "no scene " + std::to_string(d->scene_ids[b]));
        if (sc[b]->mat_type == FW_MAT_SPRING) return fail(FW_E_INVALID, "FEM-FEM groups hold FEM scenes only");
        if (sc[b]->flex_group >= 0) return fail(FW_E_INVALID, "scene already belongs to a FEM-FEM group");
        if (sc[b]->s0 != sc[0]->s0 || sc[b]->horizon != sc[0]->horizon)
            return fail(FW_E_INVALID, "a group's scenes must be added before the same substep, with the same horizon");
        for (int c = 0; c < b; ++c)
            if (d->scene_ids[c] == d->scene_ids[b]) return fail(FW_E_INVALID, "a scene appears twice in the group");
    }
    if (d->body_seg_off[0] != 0 || d->body_seg_off[B] != S || d->body_bnodes_off[0] != 0)
        return fail(FW_E_INVALID, "segment / boundary-node offsets must start at 0 and end at the counts");
    std::vector<int32_t> seg_body(S);
    for (int b = 0; b < B; ++b) {
        if (d->body_seg_off[b + 1] < d->body_seg_off[b] || d->body_bnodes_off[b + 1] < d->body_bnodes_off[b])
            return fail(FW_E_INVALID, "offsets must be non-decreasing");
        for (int k = d->body_seg_off[b]; k < d->body_seg_off[b + 1]; ++k) seg_body[k] = b;
    }
    std::vector<char> in_pair(B, 0);
    for (int q = 0; q < P; ++q) {
        const int pa = d->pair_body[2 * q], pb = d->pair_body[2 * q + 1];
        if (pa < 0 || pa >= B || pb < 0 || pb >= B || pa == pb) return fail(FW_E_INVALID, "bad pair bodies");
        in_pair[pa] = in_pair[pb] = 1;
    }
    // slots: the boundary nodes of the bodies in some pair
    std::vector<std::vector<int32_t>> slot_of(B);
    std::vector<int32_t> body_slot_off(B + 1, 0), slot_node, slot_body;
    for (int b = 0; b < B; ++b) {
        body_slot_off[b] = (int32_t)slot_node.size();
        if (!in_pair[b]) continue;
        const int N = sc[b]->n_nodes;
        slot_of[b].assign(N, -1);
        for (int k = d->body_bnodes_off[b]; k < d->body_bnodes_off[b + 1]; ++k) {
            const int o = d->body_bnodes[k];
            if (o < 0 || o >= N) return fail(FW_E_INVALID, "boundary node outside its body");
            if (slot_of[b][o] >= 0) continue;
            slot_of[b][o] = (int32_t)slot_node.size();
            slot_node.push_back(sc[b]->tiling.node_inv[o]);
            slot_body.push_back(b);
        }
    }
    body_slot_off[B] = (int32_t)slot_node.size();
    const int NS = (int)slot_node.size();
    std::vector<int2> seg_node(S), seg_slot(S);
    for (int k = 0; k < S; ++k) {
        const int b = seg_body[k], N = sc[b]->n_nodes;
        const int o0 = d->seg_node[2 * k], o1 = d->seg_node[2 * k + 1];
        if (o0 < 0 || o0 >= N || o1 < 0 || o1 >= N || o0 == o1) return fail(FW_E_INVALID, "bad segment nodes");
        seg_node[k] = make_int2(sc[b]->tiling.node_inv[o0], sc[b]->tiling.node_inv[o1]);
        seg_slot[k] = make_int2(-1, -1);
        if (in_pair[b]) {
            seg_slot[k] = make_int2(slot_of[b][o0], slot_of[b][o1]);
            if (seg_slot[k].x < 0 || seg_slot[k].y < 0) return fail(FW_E_INVALID, "segment node is not a boundary node");
        }
    }
    std::vector<int32_t> item_node(W), item_slot(W), need(B, 0), item_tied(W, 0);
    int n_tied = 0;
    for (int w = 0; w < W; ++w) {
        if (d->item_tied) {
            if (d->item_tied[w] != 0 && d->item_tied[w] != 1) return fail(FW_E_INVALID, "item_tied entries are 0 or 1");
            item_tied[w] = d->item_tied[w];
            n_tied += item_tied[w];
        }
        const int b = d->item_body[w], t = d->item_target[w], q = d->item_pair[w];
        if (b < 0 || b >= B || t < 0 || t >= B || b == t || q < 0 || q >= P || !in_pair[b] || !in_pair[t])
            return fail(FW_E_INVALID, "bad work item " + std::to_string(w));
        const int o = d->item_node[w];
        if (o < 0 || o >= sc[b]->n_nodes || slot_of[b][o] < 0)
            return fail(FW_E_INVALID, "work item node is not a boundary node");
        item_node[w] = sc[b]->tiling.node_inv[o];
        item_slot[w] = slot_of[b][o];
        need[b] += 1;
        need[t] += 2;
    }
    std::vector<char*> body_flex(B, nullptr);
    std::vector<int32_t> body_nn(B);
    for (int b = 0; b < B; ++b) {
        body_nn[b] = sc[b]->n_nodes;
        if (!in_pair[b]) continue;
        for (int g = 0; g < sc[b]->dev.hot.n_grounds; ++g) {
            const GroundDev& G = sc[b]->dev.ground[g];
            if (G.kind != FW_GROUND_FLEX_DEV) continue;
            if (G.ny < need[b])
                return fail(FW_E_INVALID, "scene " + std::to_string(d->scene_ids[b]) + " FEM-FEM record holds " +
                                              std::to_string(G.ny) + " contributions, the group needs " +
                                              std::to_string(need[b]));
            body_flex[b] = (char*)G.data;
        }
        if (!body_flex[b])
            return fail(FW_E_INVALID, "scene " + std::to_string(d->scene_ids[b]) + " has no FEM-FEM ground record");
    }

    // one block: tables (uploaded), then hash and scratch (cleared)
    int64_t o = 0;
    auto take = [&](int64_t bytes) {
        int64_t r = o;
        o += align_up(std::max<int64_t>(bytes, 4));
        return r;
    };
    const int64_t o_bodies = take((int64_t)sizeof(FlexBody) * B);
    const int64_t o_scene = take(4 * B), o_soff = take(4 * (B + 1)), o_sloff = take(4 * (B + 1)), o_flex = take(8 * B),
                  o_nn = take(4 * B), o_snode = take(4 * NS), o_sbody = take(4 * NS), o_segb = take(4 * S),
                  o_segn = take(8 * S), o_segs = take(8 * S), o_ib = take(4 * W), o_in = take(4 * W),
                  o_is = take(4 * W), o_it = take(4 * W), o_ip = take(4 * W), o_ipen = take(4 * W),
                  o_ifr = take(4 * W), o_cache = take(4 * W), o_pair = take(8 * P), o_tied = take(4 * W),
                  o_tseg = take(4 * W);
    const int64_t host_bytes = o;
    const int64_t o_grid = take(sizeof(FlexGrid)), o_box = take(16 * S), o_scell = take(16 * S), o_bcell = take(16 * B), o_segp = take(4 * S),
                  o_crng = take(8 * (int64_t)d->max_cells),
                  o_cs = take(4 * (int64_t)d->max_cells), o_ce = take(4 * (int64_t)d->max_cells),
                  o_sorted = take(4 * (int64_t)d->max_pairs), o_cent = take(4 * (int64_t)d->max_pairs),
                  o_ri = take(8 * W), o_rf = take(16 * W),
                  o_slst = take(4 * NS), o_slcur = take(4 * NS), o_ek = take(12 * (int64_t)W), o_ef = take(24 * (int64_t)W),
                  o_twg = take(16 * (int64_t)W), o_ltot = take(4);
    const int64_t total = o;
    const int64_t off = arena_alloc(a, total);
    if (off < 0) return fail(FW_E_CAPACITY, "arena full: FEM-FEM group needs " + std::to_string(total) + " bytes");
    char* blk = a->base + off;
    std::vector<char> img((size_t)host_bytes, 0);
    auto put = [&](int64_t at, const void* src, int64_t bytes) {
        if (bytes > 0) std::memcpy(img.data() + at, src, (size_t)bytes);
    };
    std::vector<char*> flex_dev(B);
    for (int b = 0; b < B; ++b) flex_dev[b] = body_flex[b];
    std::vector<FlexBody> bodies(B);
    for (int b = 0; b < B; ++b) {
        const SceneDev& SD = sc[b]->dev;
        bodies[b] = FlexBody{{SD.pos[0], SD.pos[1]}, SD.vel, SD.hot.aabb, SD.hot.aabb_k, 0};
    }
    put(o_bodies, bodies.data(), (int64_t)sizeof(FlexBody) * B);
    put(o_scene, d->scene_ids, 4 * B);
    put(o_soff, d->body_seg_off, 4 * (B + 1));
    put(o_sloff, body_slot_off.data(), 4 * (B + 1));
    put(o_flex, flex_dev.data(), 8 * B);
    put(o_nn, body_nn.data(), 4 * B);
    put(o_snode, slot_node.data(), 4 * NS);
    put(o_sbody, slot_body.data(), 4 * NS);
    put(o_segb, seg_body.data(), 4 * S);
    put(o_segn, seg_node.data(), 8 * S);
    put(o_segs, seg_slot.data(), 8 * S);
    put(o_ib, d->item_body, 4 * W);
    put(o_in, item_node.data(), 4 * W);
    put(o_is, item_slot.data(), 4 * W);
    put(o_it, d->item_target, 4 * W);
    put(o_ip, d->item_pair, 4 * W);
    put(o_ipen, d->item_penalty, 4 * W);
    put(o_ifr, d->item_friction, 4 * W);
    {
        std::vector<int32_t> minus(W, -1);
        put(o_cache, minus.data(), 4 * W);
        put(o_tseg, minus.data(), 4 * W);
    }
    put(o_tied, item_tied.data(), 4 * W);
    put(o_pair, d->pair_body, 8 * P);
    int rc = upload(a, blk, img.data(), host_bytes);
    if (rc) {
        arena_free(a, off, total);
        return rc;
    }
    HIPCHK(hipMemsetAsync(blk + host_bytes, 0, (size_t)(total - host_bytes), a->stream));

    auto F = std::make_unique<FlexHost>();
    F->block_off = off;
    F->block_bytes = total;
    F->scenes.assign(d->scene_ids, d->scene_ids + B);
    F->n_items = W;
    for (int r = 0; r < FW_FLEX_SCHED; ++r) F->sched[r] = INT32_MIN;
    F->max_cells = d->max_cells;
    F->max_pairs = d->max_pairs;
    F->n_tied = n_tied;
    FlexDev& D = F->dev;
    D.n_bodies = B;
    D.n_segs = S;
    D.n_items = W;
    D.n_pairs = P;
    D.n_slots = NS;
    D.gated = d->gated ? 1 : 0;
    D.max_cells = d->max_cells;
    D.max_query = d->max_query;
    D.max_pairs = d->max_pairs;
    D.n_tied = n_tied;
    D.bodies = (const FlexBody*)(blk + o_bodies);
    D.s0 = sc[0]->s0;
    D.horizon = sc[0]->horizon;
    D.body_scene = (const int32_t*)(blk + o_scene);
    D.body_seg_off = (const int32_t*)(blk + o_soff);
    D.body_slot_off = (const int32_t*)(blk + o_sloff);
    D.body_flex = (char* const*)(blk + o_flex);
    D.body_nn = (const int32_t*)(blk + o_nn);
    D.slot_node = (const int32_t*)(blk + o_snode);
    D.slot_body = (const int32_t*)(blk + o_sbody);
    D.seg_body = (const int32_t*)(blk + o_segb);
    D.seg_node = (const int2*)(blk + o_segn);
    D.seg_slot = (const int2*)(blk + o_segs);
    D.item_body = (const int32_t*)(blk + o_ib);
    D.item_node = (const int32_t*)(blk + o_in);
    D.item_slot = (const int32_t*)(blk + o_is);
    D.item_target = (const int32_t*)(blk + o_it);
    D.item_pair = (const int32_t*)(blk + o_ip);
    D.item_penalty = (const float*)(blk + o_ipen);
    D.item_friction = (const float*)(blk + o_ifr);
    D.cache = (int32_t*)(blk + o_cache);
    D.pair_body = (const int2*)(blk + o_pair);
    D.item_tied = (const int32_t*)(blk + o_tied);
    D.tie_seg = (int32_t*)(blk + o_tseg);
    D.tie_wg = (float4*)(blk + o_twg);
    D.grid = (FlexGrid*)(blk + o_grid);
    D.seg_box = (float4*)(blk + o_box);
    D.seg_cells = (int4*)(blk + o_scell);
    D.body_cells = (int4*)(blk + o_bcell);
    D.cell_rng = (int2*)(blk + o_crng);
    D.seg_pairs = (int32_t*)(blk + o_segp);
    D.cell_start = (int32_t*)(blk + o_cs);
    D.cell_end = (int32_t*)(blk + o_ce);
    D.sorted = (int32_t*)(blk + o_sorted);
    D.cell_ent = (int32_t*)(blk + o_cent);
    D.res_i = (int2*)(blk + o_ri);
    D.res_f = (float4*)(blk + o_rf);
    D.slot_start = (int32_t*)(blk + o_slst);
    D.slot_cur = (int32_t*)(blk + o_slcur);
    D.ent_key = (int32_t*)(blk + o_ek);
    D.ent_f = (float2*)(blk + o_ef);
    D.list_total = (int32_t*)(blk + o_ltot);

    int id = (int)a->flex.size();
    for (int g = 0; g < (int)a->flex.size(); ++g)
        if (!a->flex[g]) {
            id = g;
            break;
        }
    for (int b = 0; b < B; ++b) sc[b]->flex_group = id;
    if (id == (int)a->flex.size())
        a->flex.push_back(std::move(F));
    else
        a->flex[id] = std::move(F);
    a->flex_dirty = true;
    *group_id = id;
    return FW_OK;
}

int32_t fw_flex_rebuild(fw_arena* a, int32_t gid) {
    if (!a) return fail(FW_E_INVALID, "null arena");
    if (gid != -1 && (gid < 0 || gid >= (int)a->flex.size() || !a->flex[gid]))
        return fail(FW_E_NOT_FOUND, "no FEM-FEM group " + std::to_string(gid));
    HIPCHK(hipSetDevice(a->device));
    int rc = ensure_tiles(a);
    if (rc) return rc;
    rc = ensure_flex(a);
    if (rc) return rc;
    for (size_t g = 0; g < a->flex.size(); ++g) {
        if (!a->flex[g] || (gid >= 0 && (int)g != gid)) continue;
        const SceneHost* s0 = get_scene(a, a->flex[g]->scenes[0]);
        if (!s0 || a->clock < s0->s0)
            return fail(FW_E_INVALID, "group " + std::to_string(g) + " has not run its warm-up substep yet");
    }
    HIPCHK(fw_launch_flex_hash(a->d_flex, a->n_flex_live, a->d_scenes, a->d_ctl, gid >= 0 ? a->flex_row[gid] : -1, 0, 0,
                               a->stream));
    for (size_t g = 0; g < a->flex.size(); ++g)
        if (a->flex[g] && (gid < 0 || (int)g == gid)) a->flex[g]->hash_built = true;
    return FW_OK;
}

int32_t fw_flex_schedule_rebuild(fw_arena* a, int32_t gid, int32_t at) {
    if (!a) return fail(FW_E_INVALID, "null arena");
    if (at < 0) return fail(FW_E_INVALID, "at must be >= 0 (substeps from the next one)");
    if (gid != -1 && (gid < 0 || gid >= (int)a->flex.size() || !a->flex[gid]))
        return fail(FW_E_NOT_FOUND, "no FEM-FEM group " + std::to_string(gid));
    HIPCHK(hipSetDevice(a->device));
    const int64_t step = a->clock + (int64_t)at;
    if (step >= INT32_MAX) return fail(FW_E_INVALID, "scheduled step out of range");
    // check every group first, so that a refusal changes nothing
    for (size_t g = 0; g < a->flex.size(); ++g) {
        if (!a->flex[g] || (gid >= 0 && (int)g != gid)) continue;
        const FlexHost& F = *a->flex[g];
        const SceneHost* s0 = get_scene(a, F.scenes[0]);
        if (!s0 || step < s0->s0)
            return fail(FW_E_INVALID, "group " + std::to_string(g) + ": a rebuild needs the warm-up substep before it");
        bool dup = false, room = false;
        for (int r = 0; r < FW_FLEX_SCHED; ++r) {
            dup |= F.sched[r] == step;
            room |= F.sched[r] == INT32_MIN;
        }
        if (!dup && !room)
            return fail(FW_E_CAPACITY, "group " + std::to_string(g) + " already holds " +
                                           std::to_string(FW_FLEX_SCHED) + " pending rebuilds");
    }
    for (size_t g = 0; g < a->flex.size(); ++g) {
        if (!a->flex[g] || (gid >= 0 && (int)g != gid)) continue;
        FlexHost& F = *a->flex[g];
        int slot = -1;
        bool dup = false;
        for (int r = 0; r < FW_FLEX_SCHED; ++r) {
            if (F.sched[r] == step) dup = true;
            if (F.sched[r] == INT32_MIN && slot < 0) slot = r;
        }
        if (!dup) F.sched[slot] = (int32_t)step;
    }
    int rc = ensure_flex(a);  // table rows of the live groups
    if (rc) return rc;
    return upload_sched(a, a->n_flex_live);
}

int32_t fw_flex_init_ties(fw_arena* a, int32_t gid) {
    if (!a || gid < 0 || gid >= (int)a->flex.size() || !a->flex[gid]) return fail(FW_E_NOT_FOUND, "no FEM-FEM group");
    HIPCHK(hipSetDevice(a->device));
    int rc = ensure_tiles(a);
    if (rc) return rc;
    rc = ensure_flex(a);
    if (rc) return rc;
    const FlexHost& F = *a->flex[gid];
    const SceneHost* s0 = get_scene(a, F.scenes[0]);
    if (!s0 || a->clock < s0->s0)
        return fail(FW_E_INVALID, "group " + std::to_string(gid) + " has not run its warm-up substep yet");
    if (F.n_tied == 0) return FW_OK;
    if (!F.hash_built)
        return fail(FW_E_INVALID, "group " + std::to_string(gid) + " has no spatial hash yet: call fw_flex_rebuild first");
    HIPCHK(fw_launch_flex_ties(a->d_flex, a->flex_row[gid], F.n_items, a->d_scenes, a->d_ctl, a->stream));
    return FW_OK;
}

int32_t fw_flex_remove(fw_arena* a, int32_t gid) {
    if (!a || gid < 0 || gid >= (int)a->flex.size() || !a->flex[gid]) return fail(FW_E_NOT_FOUND, "no FEM-FEM group");
    HIPCHK(hipSetDevice(a->device));
    FlexHost& F = *a->flex[gid];
    // later substeps must not add the last contributions again: clear the ranges
    for (int32_t sid : F.scenes) {
        SceneHost* s = get_scene(a, sid);
        if (!s) continue;
        s->flex_group = -1;
        for (int g = 0; g < s->dev.hot.n_grounds; ++g)
            if (s->dev.ground[g].kind == FW_GROUND_FLEX_DEV)
                HIPCHK(hipMemsetAsync(const_cast<void*>(s->dev.ground[g].data), 0, 8 * (size_t)s->n_nodes, a->stream));
    }
    arena_free(a, F.block_off, F.block_bytes);
    a->flex[gid].reset();
    a->flex_dirty = true;
    return FW_OK;
}

int32_t fw_flex_read(fw_arena* a, int32_t gid, int32_t field, void* out, int64_t out_bytes) {
    if (!a || !out || gid < 0 || gid >= (int)a->flex.size() || !a->flex[gid])
        return fail(FW_E_NOT_FOUND, "no FEM-FEM group");
    const FlexHost& F = *a->flex[gid];
    const void* src = nullptr;
    int64_t bytes = 0;
    switch (field) {
        case FW_FLEX_CACHE: src = F.dev.cache; bytes = 4 * (int64_t)F.n_items; break;
        case FW_FLEX_GRID: src = F.dev.grid; bytes = sizeof(FlexGrid); break;
        case FW_FLEX_CELL_START: src = F.dev.cell_start; bytes = 4 * (int64_t)F.max_cells; break;
        case FW_FLEX_CELL_END: src = F.dev.cell_end; bytes = 4 * (int64_t)F.max_cells; break;
        case FW_FLEX_SORTED: src = F.dev.sorted; bytes = 4 * (int64_t)F.max_pairs; break;
        case FW_FLEX_TIE_SEG: src = F.dev.tie_seg; bytes = 4 * (int64_t)F.n_items; break;
        case FW_FLEX_TIE_STATE: src = F.dev.tie_wg; bytes = 16 * (int64_t)F.n_items; break;
        default: return fail(FW_E_INVALID, "unknown FEM-FEM field " + std::to_string(field));
    }
    if (out_bytes < bytes) return fail(FW_E_INVALID, "output buffer too small: need " + std::to_string(bytes));
    HIPCHK(hipSetDevice(a->device));
    HIPCHK(hipStreamSynchronize(a->stream));
    HIPCHK(hipMemcpy(out, src, (size_t)bytes, hipMemcpyDeviceToHost));
    return FW_OK;
}

int32_t fw_bandwidth_probe(int32_t device, int64_t bytes, int32_t reps, double* gbs, double* kernel_ms) {
    if (bytes < 4096 || reps < 1 || !gbs || !kernel_ms) return fail(FW_E_INVALID, "bad arguments");
    HIPCHK(hipSetDevice(device));
    const int64_t n4 = bytes / 16;
    void *src = nullptr, *dst = nullptr;
    HIPCHK(hipMalloc(&src, n4 * 16));
    HIPCHK(hipMalloc(&dst, n4 * 16));
    HIPCHK(hipMemset(src, 0, n4 * 16));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    const int grid = prop.multiProcessorCount * 8;
    hipEvent_t t0, t1;
    HIPCHK(hipEventCreate(&t0));
    HIPCHK(hipEventCreate(&t1));
    HIPCHK(fw_launch_copy_probe(src, dst, n4, grid, nullptr));  // first touch
    HIPCHK(hipEventRecord(t0, nullptr));
    for (int r = 0; r < reps; ++r) HIPCHK(fw_launch_copy_probe(src, dst, n4, grid, nullptr));
    HIPCHK(hipEventRecord(t1, nullptr));
    HIPCHK(hipEventSynchronize(t1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, t0, t1));
    *kernel_ms = ms / reps;
    *gbs = 2.0 * (double)(n4 * 16) / (*kernel_ms * 1e-3) / 1e9;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    (void)hipFree(src);
    (void)hipFree(dst);
    return FW_OK;
}

int32_t fw_bandwidth_ceiling(int32_t device, int64_t bytes, int32_t reps, double* gbs, double* kernel_ms,
                             int32_t* grid_per_cu, int32_t* ilp) {
    if (bytes < 4096 || reps < 1 || !gbs || !kernel_ms || !grid_per_cu || !ilp) return fail(FW_E_INVALID, "bad arguments");
    HIPCHK(hipSetDevice(device));
    const int64_t n4 = bytes / 16;
    void *src = nullptr, *dst = nullptr;
    HIPCHK(hipMalloc(&src, n4 * 16));
    HIPCHK(hipMalloc(&dst, n4 * 16));
    HIPCHK(hipMemset(src, 0, n4 * 16));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    hipEvent_t t0, t1;
    HIPCHK(hipEventCreate(&t0));
    HIPCHK(hipEventCreate(&t1));
    *gbs = 0.0;
    *kernel_ms = 0.0;
    int rc = FW_OK;
    for (int g : {2, 4, 8, 16, 32}) {
        for (int u : {1, 2, 4, 8}) {
            const int grid = prop.multiProcessorCount * g;
            hipError_t e = fw_launch_stream_copy(src, dst, n4, grid, u, nullptr);  // first touch / warm
            if (e == hipSuccess) e = hipEventRecord(t0, nullptr);
            for (int r = 0; r < reps && e == hipSuccess; ++r) e = fw_launch_stream_copy(src, dst, n4, grid, u, nullptr);
            if (e == hipSuccess) e = hipEventRecord(t1, nullptr);
            if (e == hipSuccess) e = hipEventSynchronize(t1);
            float ms = 0;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
            if (e != hipSuccess) {
                rc = fail(FW_E_HIP, std::string("stream copy: ") + hipGetErrorString(e));
                break;
            }
            const double k = ms / reps, rate = 2.0 * (double)(n4 * 16) / (k * 1e-3) / 1e9;
            if (rate > *gbs) {
                *gbs = rate;
                *kernel_ms = k;
                *grid_per_cu = g;
                *ilp = u;
            }
        }
        if (rc) break;
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    (void)hipFree(src);
    (void)hipFree(dst);
    return rc;
}

}  // extern "C"
