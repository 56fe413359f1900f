/*
 * flex_oracle.c -- serial float32 restatement of FlatWorld's batched FEM-FEM
 * penalty contact with the FEM spatial hash.  TEST INFRASTRUCTURE ONLY (see
 * fem_oracle.h).
 *
 * Every reference kernel below is run serially in ascending thread id.  Line
 * references are to the reference package (flatworld/...).
 */
#include "fem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* wp.dot / wp.length of vec2: a zero-initialised sum of products */
static float dot2f(float ax, float ay, float bx, float by) { return (0.0f + ax * bx) + ay * by; }
static float len2f(float x, float y) { return sqrtf(dot2f(x, y, x, y)); }
static float minf_(float a, float b) { return a < b ? a : b; }
static float maxf_(float a, float b) { return a > b ? a : b; }
static int32_t mini_(int32_t a, int32_t b) { return a < b ? a : b; }
static int32_t maxi_(int32_t a, int32_t b) { return a > b ? a : b; }

/* contact_detection.py:15-61 */
int32_t ora_point_edge(const float p[2], const float a[2], const float b[2], float* pen, float normal[2], float w[2]) {
    const float ex = b[0] - a[0], ey = b[1] - a[1];
    const float len = len2f(ex, ey);
    int32_t inside = 0;
    *pen = 1e9f;
    normal[0] = normal[1] = 0.0f;
    w[0] = w[1] = 0.0f;
    if (len > 1e-9f) {
        const float dx = ex / len, dy = ey / len;
        normal[0] = dy;
        normal[1] = -dx;
        const float tx = p[0] - a[0], ty = p[1] - a[1];
        float t = dot2f(tx, ty, dx, dy);
        t = t / len;
        if (0.0f <= t && t <= 1.0f) {
            inside = 1;
            *pen = dot2f(tx, ty, normal[0], normal[1]);
            w[0] = 1.0f - t;
            w[1] = t;
        }
    }
    return inside;
}

/* detectPointToMeshBoundary (contact_detection.py:64-90): the penetration when
 * the point projects inside the segment within `limit`, else 1e9; the weights
 * come back either way. */
static float detect(const float p[2], const float a[2], const float b[2], float limit, float nrm[2], float w[2]) {
    float pen, n[2];
    const int32_t inside = ora_point_edge(p, a, b, &pen, n, w);
    if (fabsf(pen) < limit && inside) {
        nrm[0] = n[0];
        nrm[1] = n[1];
        return pen;
    }
    nrm[0] = nrm[1] = 0.0f;
    return 1e9f;
}

static const float* seg_xy(const ora_flex* f, ora_scene* const* bodies, int32_t s, int k) {
    return bodies[f->seg_body[s]]->x + 2 * f->seg_node[2 * s + k];
}

/* Cell range of a box, clamped to the grid (spatialmanager.py:160-175, 276-291). */
static void cell_range(const ora_flex* f, float lbx, float lby, float ubx, float uby, int32_t r[4]) {
    const float dlx = lbx - f->glb[0], dly = lby - f->glb[1];
    const float dux = ubx - f->glb[0], duy = uby - f->glb[1];
    int32_t lx = (int32_t)floorf(dlx / f->gs[0]);
    int32_t ly = (int32_t)floorf(dly / f->gs[1]);
    int32_t ux = (int32_t)floorf(dux / f->gs[0]);
    int32_t uy = (int32_t)floorf(duy / f->gs[1]);
    lx = mini_(maxi_(lx, 0), f->cn[0] - 1);
    ly = mini_(maxi_(ly, 0), f->cn[1] - 1);
    ux = mini_(maxi_(ux, 0), f->cn[0] - 1);
    uy = mini_(maxi_(uy, 0), f->cn[1] - 1);
    r[0] = lx;
    r[1] = ly;
    r[2] = ux;
    r[3] = uy;
}

/* femspringmanager.py:421-487 (populate, velocity buffer 0) and
 * spatialmanager.py:30-57, 202-385, 613-712 (reset, add_element, build). */
void ora_flex_hash_build(ora_flex* f, ora_scene* const* bodies) {
    float glb0 = 1e9f, glb1 = 1e9f, gub0 = -1e9f, gub1 = -1e9f;
    for (int32_t b = 0; b < f->n_bodies; ++b) {
        const float* bb = bodies[b]->aabb;
        glb0 = minf_(glb0, bb[0]);
        glb1 = minf_(glb1, bb[1]);
        gub0 = maxf_(gub0, bb[2]);
        gub1 = maxf_(gub1, bb[3]);
    }
    const float extent = maxf_(gub0 - glb0, gub1 - glb1);
    const float expand = extent * 0.01f + 0.0f;
    f->glb[0] = glb0 - expand;
    f->glb[1] = glb1 - expand;
    f->gub[0] = gub0 + expand;
    f->gub[1] = gub1 + expand;

    /* element boxes (two nodes; the third connectivity entry is -1) and the
       smallest box diagonal */
    const int32_t ne = f->n_segs;
    float* box = (float*)malloc(sizeof(float) * 4 * (size_t)(ne > 0 ? ne : 1));
    float est = 1e30f;
    for (int32_t s = 0; s < ne; ++s) {
        float lb0 = 1e30f, lb1 = 1e30f, ub0 = -1e30f, ub1 = -1e30f;
        for (int k = 0; k < 2; ++k) {
            const float* c = seg_xy(f, bodies, s, k);
            lb0 = minf_(lb0, c[0]);
            lb1 = minf_(lb1, c[1]);
            ub0 = maxf_(ub0, c[0]);
            ub1 = maxf_(ub1, c[1]);
        }
        box[4 * s] = lb0 - 0.0f;
        box[4 * s + 1] = lb1 - 0.0f;
        box[4 * s + 2] = ub0 + 0.0f;
        box[4 * s + 3] = ub1 + 0.0f;
        est = minf_(est, len2f(box[4 * s + 2] - box[4 * s], box[4 * s + 3] - box[4 * s + 1]));
    }

    /* _build_grid_kernel (spatialmanager.py:217-245); pow(max_cells, 0.5) is the
       correctly rounded square root */
    float cell = est * 1.5f;
    const float ex = f->gub[0] - f->glb[0], ey = f->gub[1] - f->glb[1];
    float mpd = maxf_(sqrtf((float)f->max_cells), 1.0f);
    cell = maxf_(cell, ex / mpd);
    cell = maxf_(cell, ey / mpd);
    const float cs = cell + 1e-6f;
    const int32_t nx = (int32_t)maxf_(ceilf(ex / cs), 1.0f);
    const int32_t ny = (int32_t)maxf_(ceilf(ey / cs), 1.0f);
    f->cn[0] = nx;
    f->cn[1] = ny;
    f->gs[0] = ex / (float)nx;
    f->gs[1] = ey / (float)ny;
    f->total_cells = mini_(nx * ny, f->max_cells);
    const int32_t tc = f->total_cells;
    if (ne == 0) {
        free(box);
        return;
    }

    /* expand, count, prefix, scatter: serially, a cell's entries are in ascending
       segment order */
    int32_t* count = (int32_t*)calloc((size_t)(tc > 0 ? tc : 1), sizeof(int32_t));
    int32_t np_ = 0;
    for (int pass = 0; pass < 2; ++pass) {
        np_ = 0;
        if (pass == 1) {
            int32_t acc = 0;
            for (int32_t c = 0; c < tc; ++c) {
                f->cell_start[c] = acc;
                acc += count[c];
                count[c] = f->cell_start[c];
            }
        }
        for (int32_t s = 0; s < ne; ++s) {
            int32_t r[4];
            cell_range(f, box[4 * s], box[4 * s + 1], box[4 * s + 2], box[4 * s + 3], r);
            for (int32_t ix = r[0]; ix <= r[2]; ++ix)
                for (int32_t iy = r[1]; iy <= r[3]; ++iy) {
                    const int32_t cid = ix + iy * nx;
                    if (cid < tc) {
                        if (np_ < f->max_pairs) {
                            if (pass == 0)
                                count[cid]++;
                            else
                                f->sorted_seg[count[cid]++] = s;
                        }
                        np_++;
                    }
                }
        }
    }
    for (int32_t c = 0; c < tc; ++c) f->cell_end[c] = count[c];
    f->n_hash_pairs = mini_(np_, f->max_pairs);
    free(count);
    free(box);
}

/* spatialmanager.py:131-199 */
int32_t ora_flex_query(const ora_flex* f, const float p[2], float buffer, int32_t target, int32_t* out) {
    int32_t n = 0;
    const int32_t tc = f->total_cells;
    if (tc <= 0) return 0;
    int32_t r[4];
    cell_range(f, p[0] - buffer, p[1] - buffer, p[0] + buffer, p[1] + buffer, r);
    for (int32_t ix = r[0]; ix <= r[2]; ++ix)
        for (int32_t iy = r[1]; iy <= r[3]; ++iy) {
            const int32_t cid = ix + iy * f->cn[0];
            if (cid < 0 || cid >= tc) continue;
            for (int32_t q = f->cell_start[cid]; q < f->cell_end[cid]; ++q) {
                const int32_t s = f->sorted_seg[q];
                if (s < 0) continue;
                if (f->seg_body[s] != target && target != -1) continue;
                int32_t already = 0;
                for (int32_t k = 0; k < n; ++k)
                    if (out[k] == s) already = 1;
                if (!already && n < f->max_query) out[n++] = s;
            }
        }
    return n;
}

/* Pair activation through the BVH: inclusive overlap of the two bodies' AABBs,
 * both widened by 0.1 * extent on full rebuilds (counter % 100 == 0); the same
 * rule as ora_gate (flatworld/explicitloop.py:477-487, bvh.py). */
int32_t ora_gate_pair(const float alb[2], const float aub[2], const float blb[2], const float bub[2],
                      int32_t counter) {
    float a0[2] = {alb[0], alb[1]}, a1[2] = {aub[0], aub[1]};
    float b0[2] = {blb[0], blb[1]}, b1[2] = {bub[0], bub[1]};
    if (counter % 100 == 0) {
        for (int k = 0; k < 2; ++k) {
            const float ba = 0.1f * (a1[k] - a0[k]);
            const float la = a0[k] - ba, ua = a1[k] + ba;
            const float bb = 0.1f * (b1[k] - b0[k]);
            const float lb = b0[k] - bb, ub = b1[k] + bb;
            a0[k] = la;
            a1[k] = ua;
            b0[k] = lb;
            b1[k] = ub;
        }
    }
    return a0[0] <= b1[0] && a1[0] >= b0[0] && a0[1] <= b1[1] && a1[1] >= b0[1];
}

typedef struct {
    int32_t hit, seg;
    float tf[2], w[2];
} item_result;

/* _fem_fem_tie_elem_normal (mixedcontact.py:150-163): the edge's left normal
 * (-t.y, t.x), the opposite of the contact normal, (1, 0) for a degenerate edge. */
static void tie_normal(const float a[2], const float c[2], float n[2]) {
    const float ex = c[0] - a[0], ey = c[1] - a[1];
    const float elen = len2f(ex, ey);
    float tx = 1.0f, ty = 0.0f;
    if (elen > 1e-10f) {
        tx = ex / elen;
        ty = ey / elen;
    }
    n[0] = -ty;
    n[1] = tx;
}

/* _eval_fixed_tie_point (mixedcontact.py:136-147) */
static void tie_point(const float a[2], const float c[2], const float w[2], float cp[2]) {
    cp[0] = (0.0f + a[0] * w[0]) + c[0] * w[1];
    cp[1] = (0.0f + a[1] * w[0]) + c[1] * w[1];
}

void ora_flex_ties_init(ora_flex* f, ora_scene* const* bodies) {
    const float qbuf = maxf_(maxf_(1e-4f, maxf_(f->gs[0], f->gs[1])), 0.001f);
    int32_t* list = (int32_t*)malloc(sizeof(int32_t) * (size_t)(f->max_query > 0 ? f->max_query : 1));
    for (int32_t w = 0; w < f->n_items; ++w) {
        if (!f->item_tied[w] || f->tie_elem[w] >= 0) continue;
        const ora_scene* B = bodies[f->item_body[w]];
        const int32_t nid = f->item_node[w];
        const float p[2] = {B->x[2 * nid], B->x[2 * nid + 1]};
        float best = 1e9f, bw[2] = {0.0f, 0.0f};
        int32_t be = -1;
        const int32_t cnt = ora_flex_query(f, p, qbuf, f->item_target[w], list);
        for (int32_t j = 0; j < cnt; ++j) {
            const int32_t s = list[j];
            float n[2], wt[2];
            const float pen = detect(p, seg_xy(f, bodies, s, 0), seg_xy(f, bodies, s, 1), qbuf, n, wt);
            if (fabsf(pen) < fabsf(best) && fabsf(pen) < qbuf) {
                best = pen;
                bw[0] = wt[0];
                bw[1] = wt[1];
                be = s;
            }
        }
        if (be >= 0) {
            const float* a = seg_xy(f, bodies, be, 0);
            const float* c = seg_xy(f, bodies, be, 1);
            float cp[2], n[2];
            tie_point(a, c, bw, cp);
            tie_normal(a, c, n);
            f->tie_elem[w] = be;
            f->tie_w[2 * w] = bw[0];
            f->tie_w[2 * w + 1] = bw[1];
            f->tie_gap[w] = dot2f(p[0] - cp[0], p[1] - cp[1], n[0], n[1]);
        }
    }
    free(list);
}

/* _batched_flexflex_contact_kernel_sh (mixedcontact.py:886-1025) for every work
 * item, then the node-major contribution lists. */
void ora_flex_contacts(ora_flex* f, ora_scene* const* bodies, int32_t counter) {
    const int32_t W = f->n_items;
    item_result* res = (item_result*)calloc((size_t)(W > 0 ? W : 1), sizeof(item_result));
    int32_t* active = (int32_t*)calloc((size_t)(f->n_pairs > 0 ? f->n_pairs : 1), sizeof(int32_t));
    int32_t* list = (int32_t*)malloc(sizeof(int32_t) * (size_t)(f->max_query > 0 ? f->max_query : 1));
    for (int32_t q = 0; q < f->n_pairs; ++q) {
        if (!f->gated) {
            active[q] = 1;
            continue;
        }
        const float* a = bodies[f->pair_body[2 * q]]->aabb;
        const float* b = bodies[f->pair_body[2 * q + 1]]->aabb;
        active[q] = ora_gate_pair(a, a + 2, b, b + 2, counter);
    }
    const float min_buf = maxf_(f->gs[0], f->gs[1]);
    const float qbuf = maxf_(maxf_(1e-4f, min_buf), 0.016f);

    /* tie springs: every resolved tied item, no pair gate (mixedcontact.py:773-811) */
    item_result* tie = (item_result*)calloc((size_t)(W > 0 ? W : 1), sizeof(item_result));
    for (int32_t w = 0; w < W; ++w) {
        if (!f->item_tied[w] || f->tie_elem[w] < 0) continue;
        const ora_scene* B = bodies[f->item_body[w]];
        const int32_t nid = f->item_node[w], s = f->tie_elem[w];
        const float* a = seg_xy(f, bodies, s, 0);
        const float* c = seg_xy(f, bodies, s, 1);
        float cp[2], n[2];
        tie_point(a, c, f->tie_w + 2 * w, cp);
        tie_normal(a, c, n);
        const float gap = f->tie_gap[w], k = f->item_penalty[w];
        const float tx = cp[0] + gap * n[0], ty = cp[1] + gap * n[1];
        tie[w].hit = 1;
        tie[w].seg = s;
        tie[w].tf[0] = k * (tx - B->x[2 * nid]);
        tie[w].tf[1] = k * (ty - B->x[2 * nid + 1]);
        tie[w].w[0] = f->tie_w[2 * w];
        tie[w].w[1] = f->tie_w[2 * w + 1];
    }

    for (int32_t w = 0; w < W; ++w) {
        if (f->item_tied[w] || !active[f->item_pair[w]]) continue;
        const ora_scene* B = bodies[f->item_body[w]];
        const int32_t nid = f->item_node[w];
        const float p[2] = {B->x[2 * nid], B->x[2 * nid + 1]};
        const float vel[2] = {B->v[2 * nid], B->v[2 * nid + 1]};
        const int32_t tb = f->item_target[w];
        const int32_t be_off = f->body_seg_off[tb], be_cnt = f->body_seg_off[tb + 1] - be_off;
        const float penalty = f->item_penalty[w], friction = f->item_friction[w];

        float best_pen = 1e9f, bn[2] = {0.0f, 0.0f}, bw[2] = {0.0f, 0.0f};
        int32_t found = 0, bs = -1;
        const int32_t cached = f->cache[w];
        if (cached >= be_off && cached < be_off + be_cnt) {
            const float* a = seg_xy(f, bodies, cached, 0);
            const float* c = seg_xy(f, bodies, cached, 1);
            float lb0 = 1e30f, lb1 = 1e30f, ub0 = -1e30f, ub1 = -1e30f;
            lb0 = minf_(lb0, a[0]);
            lb1 = minf_(lb1, a[1]);
            ub0 = maxf_(ub0, a[0]);
            ub1 = maxf_(ub1, a[1]);
            lb0 = minf_(lb0, c[0]);
            lb1 = minf_(lb1, c[1]);
            ub0 = maxf_(ub0, c[0]);
            ub1 = maxf_(ub1, c[1]);
            const float margin = maxf_(1e-4f, 1e-2f * len2f(ub0 - lb0, ub1 - lb1));
            lb0 = lb0 - margin;
            lb1 = lb1 - margin;
            ub0 = ub0 + margin;
            ub1 = ub1 + margin;
            /* _point_outside_aabb (mixedcontact.py:113-119) */
            const int32_t outside = minf_(p[0] - lb0, p[1] - lb1) < 0.0f || minf_(ub0 - p[0], ub1 - p[1]) < 0.0f;
            if (!outside) {
                float n[2], wt[2];
                const float pen = detect(p, a, c, margin, n, wt);
                if (pen < 1e-4f && fabsf(pen) < margin) {
                    best_pen = pen;
                    bn[0] = n[0];
                    bn[1] = n[1];
                    bw[0] = wt[0];
                    bw[1] = wt[1];
                    bs = cached;
                    found = 1;
                }
            }
        }
        if (!found) {
            const int32_t cnt = ora_flex_query(f, p, qbuf, tb, list);
            for (int32_t j = 0; j < cnt; ++j) {
                const int32_t s = list[j];
                float n[2], wt[2];
                const float pen = detect(p, seg_xy(f, bodies, s, 0), seg_xy(f, bodies, s, 1), qbuf, n, wt);
                if (pen < best_pen && pen < 1e-4f && fabsf(pen) < qbuf) {
                    best_pen = pen;
                    bn[0] = n[0];
                    bn[1] = n[1];
                    bw[0] = wt[0];
                    bw[1] = wt[1];
                    bs = s;
                    found = 1;
                    f->cache[w] = s;
                }
            }
        }
        if (!found) f->cache[w] = -1;
        if (found && best_pen < 0.0f) {
            const float nfx = ((-bn[0]) * penalty) * best_pen;
            const float nfy = ((-bn[1]) * penalty) * best_pen;
            float tx = nfx, ty = nfy;
            if (friction > 1e-9f) {
                const ora_scene* T = bodies[tb];
                const int32_t n0 = f->seg_node[2 * bs], n1 = f->seg_node[2 * bs + 1];
                const float svx = T->v[2 * n0] * bw[0] + T->v[2 * n1] * bw[1];
                const float svy = T->v[2 * n0 + 1] * bw[0] + T->v[2 * n1 + 1] * bw[1];
                const float rvx = vel[0] - svx, rvy = vel[1] - svy;
                const float d = dot2f(rvx, rvy, bn[0], bn[1]);
                const float tgx = rvx - d * bn[0], tgy = rvy - d * bn[1];
                const float tlen = len2f(tgx, tgy);
                if (tlen > 1e-9f) {
                    const float fdx = (-tgx) / tlen, fdy = (-tgy) / tlen;
                    const float nl = len2f(nfx, nfy);
                    tx = tx + (fdx * friction) * nl;
                    ty = ty + (fdy * friction) * nl;
                }
            }
            res[w].hit = 1;
            res[w].seg = bs;
            res[w].tf[0] = tx;
            res[w].tf[1] = ty;
            res[w].w[0] = bw[0];
            res[w].w[1] = bw[1];
        }
    }

    /* node-major lists: the node's own force, then -(tf * w0) and -(tf * w1) on the
       segment's nodes, in ascending work item (the serial order of the atomics) */
    for (int32_t b = 0; b < f->n_bodies; ++b) memset(f->c_off[b], 0, sizeof(int32_t) * (size_t)(bodies[b]->n_nodes + 1));
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            for (int32_t b = 0; b < f->n_bodies; ++b) {
                int32_t* off = f->c_off[b];
                int32_t acc = 0;
                for (int32_t n = 0; n <= bodies[b]->n_nodes; ++n) {
                    const int32_t c = off[n];
                    off[n] = acc;
                    acc += c;
                }
            }
        }
        for (int32_t r = 0; r < 2 * W; ++r) {
            const item_result* R = r < W ? tie + r : res + (r - W);  /* ties first, then contacts */
            const int32_t w = r < W ? r : r - W;
            if (!R->hit) continue;
            const int32_t s = R->seg;
            const int32_t tgt[3][2] = {{f->item_body[w], f->item_node[w]},
                                       {f->seg_body[s], f->seg_node[2 * s]},
                                       {f->seg_body[s], f->seg_node[2 * s + 1]}};
            const float tx = R->tf[0], ty = R->tf[1];
            const float val[3][2] = {{tx, ty},
                                     {-(tx * R->w[0]), -(ty * R->w[0])},
                                     {-(tx * R->w[1]), -(ty * R->w[1])}};
            for (int k = 0; k < 3; ++k) {
                int32_t* off = f->c_off[tgt[k][0]];
                const int32_t n = tgt[k][1];
                if (pass == 0) {
                    off[n + 1]++;
                } else {
                    const int32_t j = off[n + 1]++;
                    f->c_f[tgt[k][0]][2 * j] = val[k][0];
                    f->c_f[tgt[k][0]][2 * j + 1] = val[k][1];
                }
            }
        }
    }
    free(list);
    free(active);
    free(res);
    free(tie);
}
