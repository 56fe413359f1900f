/*
 * fem_oracle.c -- serial float32 restatement of FlatWorld's explicit-FEM
 * substep.  TEST INFRASTRUCTURE ONLY (see fem_oracle.h).
 *
 * Line references are to the reference package (flatworld/...).
 */
#include "fem_oracle.h"

#include <math.h>
#include <string.h>

/* ------------------------------------------------------------------ 2x2 algebra */

/* Row-major 2x2: m[0]=m00 m[1]=m01 m[2]=m10 m[3]=m11. */
static void mat_mul(const float a[4], const float b[4], float c[4]) {
    c[0] = a[0] * b[0] + a[1] * b[2];
    c[1] = a[0] * b[1] + a[1] * b[3];
    c[2] = a[2] * b[0] + a[3] * b[2];
    c[3] = a[2] * b[1] + a[3] * b[3];
}

static float mat_det(const float m[4]) { return m[0] * m[3] - m[1] * m[2]; }

/* Textbook adjugate inverse, adj(m) * (1/det). */
static void mat_inv(const float m[4], float r[4]) {
    float inv_det = 1.0f / mat_det(m);
    r[0] = m[3] * inv_det;
    r[1] = -m[1] * inv_det;
    r[2] = -m[2] * inv_det;
    r[3] = m[0] * inv_det;
}

/* C = F^T F */
static void mat_ftf(const float F[4], float C[4]) {
    C[0] = F[0] * F[0] + F[2] * F[2];
    C[1] = F[0] * F[1] + F[2] * F[3];
    C[2] = F[1] * F[0] + F[3] * F[2];
    C[3] = F[1] * F[1] + F[3] * F[3];
}

/* Natural log of a float from IEEE double operations only, rounded to float
 * once.  The reference evaluates wp.log in float32 (warp/native/builtin.h,
 * logf); any faithful float log fits the restated algorithm, and this one is
 * chosen because the HIP kernel evaluates the identical operation sequence,
 * which makes the Neo-Hookean stress reproducible bit for bit:
 *   x = m * 2^e, m in [sqrt(1/2), sqrt(2)),  log m = 2 atanh(s),
 *   s = (m - 1) / (m + 1) = f / (2 + f),  atanh(s)/s = sum s^2k / (2k + 1), k <= 6.
 * |s| <= 0.1716, so the dropped tail is below 2e-12 relative. */
static float log_f32(float x)
{
    if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : NAN;
    if (isinf(x)) return x;
    int e = 0;
    if (x < 1.17549435e-38f) { x = x * 8388608.0f; e = -23; }
    uint32_t b;
    memcpy(&b, &x, 4);
    e += (int)((b >> 23) & 0xFFu) - 127;
    uint32_t mb = (b & 0x007FFFFFu) | 0x3F800000u;
    float m;
    memcpy(&m, &mb, 4);
    if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
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

/* Exported for the pinning tests of the log above. */
float ora_log_f32(float x) { return log_f32(x); }

/* float -> int32 with the device conversion's semantics: NaN -> 0, saturating
 * at the int32 range.  Identical to a C cast for finite in-range values; the
 * reference leaves the other cases undefined. */
static int cvt_i32(float v) {
    if (v != v) return 0;
    if (v >= 2147483648.0f) return 2147483647;
    if (v <= -2147483648.0f) return (int)(-2147483647 - 1);
    return (int)v;
}

/* ------------------------------------------------------------------ materials */

/* flatworld/materials/materialfunctions.py:10-29 (getStress2D). */
void ora_stress2d(int32_t type, float E, float nu, const float F[4], float stress[3], float strain[3]) {
    float lam = E * nu / ((1.0f + nu) * (1.0f - 2.0f * nu));
    float mu = E / (2.0f * (1.0f + nu));
    float C[4];
    mat_ftf(F, C);
    /* strain = 0.5 * (F^T F - I) */
    float e00 = 0.5f * (C[0] - 1.0f);
    float e01 = 0.5f * (C[1] - 0.0f);
    float e11 = 0.5f * (C[3] - 1.0f);
    strain[0] = e00;
    strain[1] = e11;
    strain[2] = e01;
    stress[0] = 0.0f;
    stress[1] = 0.0f;
    stress[2] = 0.0f;
    if (type == ORA_ELASTIC) {
        /* (lam * tr(E)) * I + 2 mu E */
        float ltr = lam * (e00 + e11);
        float tmu = 2.0f * mu;
        stress[0] = ltr * 1.0f + tmu * e00;
        stress[1] = ltr * 1.0f + tmu * e11;
        stress[2] = ltr * 0.0f + tmu * e01;
    } else if (type == ORA_NEOHOOKEAN) {
        /* (lam ln J - mu) C^-1 + mu I */
        float Ci[4];
        mat_inv(C, Ci);
        float J = mat_det(F);
        float c = lam * log_f32(J) - mu;
        stress[0] = c * Ci[0] + mu * 1.0f;
        stress[1] = c * Ci[3] + mu * 1.0f;
        stress[2] = c * Ci[1] + mu * 0.0f;
    }
}

/* flatworld/materials/materialfunctions.py:35-101 (misesReturnMap2D).
 * Returns 1 when the trial state is above yield (f_trial > 0). */
int32_t ora_mises2d(float E, float nu, float sy, float H, const float F[4], const float eps_p_old[3],
                    float eqps_old, float stress[3], float eps_p_new[3], float* eqps_new, float strain[3]) {
    float lam = E * nu / ((1.0f + nu) * (1.0f - 2.0f * nu));
    float mu = E / (2.0f * (1.0f + nu));
    float C[4];
    mat_ftf(F, C);
    float et0 = 0.5f * (C[0] - 1.0f);
    float et1 = 0.5f * (C[3] - 1.0f);
    float et2 = 0.5f * (C[1] - 0.0f);
    strain[0] = et0;
    strain[1] = et1;
    strain[2] = et2;
    float ee0 = et0 - eps_p_old[0];
    float ee1 = et1 - eps_p_old[1];
    float ee2 = et2 - eps_p_old[2];
    float tr = ee0 + ee1;
    float s0 = lam * tr + 2.0f * mu * ee0;
    float s1 = lam * tr + 2.0f * mu * ee1;
    float s2 = 2.0f * mu * ee2;
    float p = (s0 + s1) / 3.0f;
    float d0 = s0 - p;
    float d1 = s1 - p;
    float d2 = s2;
    float vm = sqrtf(d0 * d0 + d1 * d1 - d0 * d1 + 3.0f * d2 * d2);
    float f_trial = vm - (sy + H * eqps_old);
    eps_p_new[0] = eps_p_old[0];
    eps_p_new[1] = eps_p_old[1];
    eps_p_new[2] = eps_p_old[2];
    *eqps_new = eqps_old;
    stress[0] = s0;
    stress[1] = s1;
    stress[2] = s2;
    if (f_trial > 0.0f) {
        float denom = 3.0f * mu + H;
        float dgamma = f_trial / denom;
        *eqps_new = eqps_old + dgamma;
        float scale = 1.0f - 3.0f * mu * dgamma / vm;
        stress[0] = d0 * scale + p;
        stress[1] = d1 * scale + p;
        stress[2] = d2 * scale;
        float n0 = d0 / vm, n1 = d1 / vm, n2 = d2 / vm;
        eps_p_new[0] = eps_p_old[0] + dgamma * n0;
        eps_p_new[1] = eps_p_old[1] + dgamma * n1;
        eps_p_new[2] = eps_p_old[2] + dgamma * n2;
        return 1;
    }
    return 0;
}

/* ------------------------------------------------------------------ element */

/* flatworld/femspringmanager.py:217-280 with get_b_matrix_2d (71-105).
 * f holds el_fint = area * B^T S for corners 0..2 (before the minus sign of
 * the scatter). */
void ora_element_force(const ora_scene* s, int32_t e, float f[6], float cauchy[3], float strain[3],
                       float eps_p_new[3], float* eqps_new, int32_t* yielded) {
    const int32_t n1 = s->conn[3 * e + 0], n2 = s->conn[3 * e + 1], n3 = s->conn[3 * e + 2];
    const float* x = s->x;
    float Jac[4] = {x[2 * n1] - x[2 * n3], x[2 * n2] - x[2 * n3], x[2 * n1 + 1] - x[2 * n3 + 1],
                    x[2 * n2 + 1] - x[2 * n3 + 1]};
    const float* Ji = s->jinv + 4 * e;
    float F[4];
    mat_mul(Jac, Ji, F);

    /* B matrix entries (3x6), flatworld/femspringmanager.py:71-105 */
    float d00 = Ji[0], d01 = Ji[1], d10 = Ji[2], d11 = Ji[3];
    float d20 = -Ji[0] - Ji[2];
    float d21 = -Ji[1] - Ji[3];
    float B[3][6];
    B[0][0] = d00 * F[0];
    B[0][1] = d00 * F[2];
    B[1][0] = d01 * F[1];
    B[1][1] = d01 * F[3];
    B[2][0] = d01 * F[0] + d00 * F[1];
    B[2][1] = d00 * F[3] + d01 * F[2];
    B[0][2] = d10 * F[0];
    B[0][3] = d10 * F[2];
    B[1][2] = d11 * F[1];
    B[1][3] = d11 * F[3];
    B[2][2] = d11 * F[0] + d10 * F[1];
    B[2][3] = d10 * F[3] + d11 * F[2];
    B[0][4] = d20 * F[0];
    B[0][5] = d20 * F[2];
    B[1][4] = d21 * F[1];
    B[1][5] = d21 * F[3];
    B[2][4] = d21 * F[0] + d20 * F[1];
    B[2][5] = d20 * F[3] + d21 * F[2];

    float sig[3];
    *yielded = 0;
    if (s->mat_type == ORA_MISES) {
        *yielded = ora_mises2d(s->E, s->nu, s->sigma_y, s->H, F, s->eps_p + 3 * e, s->eqps[e], sig, eps_p_new,
                               eqps_new, strain);
    } else {
        ora_stress2d(s->mat_type, s->E, s->nu, F, sig, strain);
    }

    /* PK2 -> Cauchy (diagnostic), flatworld/femspringmanager.py:267-272 */
    float S[4] = {sig[0], sig[2], sig[2], sig[1]};
    float J = mat_det(F);
    float FS[4], FSFt[4], Ft[4] = {F[0], F[2], F[1], F[3]};
    mat_mul(F, S, FS);
    mat_mul(FS, Ft, FSFt);
    float jd = J > 1e-12f ? J : 1e-12f;
    cauchy[0] = FSFt[0] / jd;
    cauchy[1] = FSFt[3] / jd;
    cauchy[2] = FSFt[1] / jd;

    float a = s->area[e];
    for (int k = 0; k < 6; ++k) {
        float t = B[0][k] * sig[0] + B[1][k] * sig[1] + B[2][k] * sig[2];
        f[k] = t * a;
    }
}

/* ------------------------------------------------------------------ springs */

/* Warp's dot of two vec2: a zero-initialised sum over the components. */
static float dot2(float ax, float ay, float bx, float by) { return (0.0f + ax * bx) + ay * by; }

/* flatworld/femspringmanager.py:283-312: xab = x_a - x_b, lnew = |xab|,
 * dir = xab / lnew, f = k (lnew / l0 - 1) dir + c (v_ab . dir) dir. */
void ora_spring_force(const ora_scene* s, int32_t e, float f[2]) {
    const int32_t ia = s->conn[2 * e], ib = s->conn[2 * e + 1];
    const float xab0 = s->x[2 * ia] - s->x[2 * ib];
    const float xab1 = s->x[2 * ia + 1] - s->x[2 * ib + 1];
    const float lnew = sqrtf(dot2(xab0, xab1, xab0, xab1));
    const float d0 = xab0 / lnew, d1 = xab1 / lnew;
    const float a = s->E * (lnew / s->rest_len[e] - 1.0f);
    float f0 = a * d0, f1 = a * d1;
    const float vab0 = s->v[2 * ia] - s->v[2 * ib];
    const float vab1 = s->v[2 * ia + 1] - s->v[2 * ib + 1];
    const float b = s->spring_c * dot2(vab0, vab1, d0, d1);
    f0 = f0 + b * d0;
    f1 = f1 + b * d1;
    f[0] = f0;
    f[1] = f1;
}

/* flatworld/femspringmanager.py:141-201, one thread: FEM elements first, then springs. */
float ora_mass_scaling(ora_scene* s, float dt_target) {
    const float safety = 0.4f;
    float added = 0.0f;
    const float* x = s->x;
    if (s->mat_type != ORA_SPRING) {
        for (int32_t e = 0; e < s->n_elems; ++e) {
            const int32_t n1 = s->conn[3 * e], n2 = s->conn[3 * e + 1], n3 = s->conn[3 * e + 2];
            const float a0 = x[2 * n2] - x[2 * n1], a1 = x[2 * n2 + 1] - x[2 * n1 + 1];
            const float b0 = x[2 * n3] - x[2 * n2], b1 = x[2 * n3 + 1] - x[2 * n2 + 1];
            const float c0 = x[2 * n1] - x[2 * n3], c1 = x[2 * n1 + 1] - x[2 * n3 + 1];
            const float e01 = sqrtf(dot2(a0, a1, a0, a1));
            const float e12 = sqrtf(dot2(b0, b1, b0, b1));
            const float e20 = sqrtf(dot2(c0, c1, c0, c1));
            const float m12 = e12 < e20 ? e12 : e20; /* wp.min */
            const float L = e01 < m12 ? e01 : m12;
            const float c_e = sqrtf(s->E / s->rho);
            const float dt_elem = safety * L / c_e;
            if (dt_elem < dt_target && L > 1e-15f) {
                const float r = dt_target / (safety * L);
                const float rho_new = s->E * r * r;
                const float dm = (rho_new - s->rho) * s->area[e] / 3.0f;
                s->mass[n1] = s->mass[n1] + dm;
                s->mass[n2] = s->mass[n2] + dm;
                s->mass[n3] = s->mass[n3] + dm;
                added = added + (rho_new - s->rho) * s->area[e];
            }
        }
    } else {
        for (int32_t e = 0; e < s->n_elems; ++e) {
            const float m_min = 4.0f * s->E * dt_target * dt_target;
            for (int c = 0; c < 2; ++c) {
                const int32_t nid = s->conn[2 * e + c];
                if (s->mass[nid] < m_min) {
                    const float dm = m_min - s->mass[nid];
                    s->mass[nid] = m_min;
                    added = added + dm;
                }
            }
        }
    }
    return added;
}

/* ------------------------------------------------------------------ grounds */

/* flatworld/mixedcontact.py:166-217 (_nearest_on_curve_2d). */
static float hf_signed(const ora_ground* g, float x, float z, float n[2]) {
    const int nx = g->hf_nx;
    float span = g->hf_ub_x - g->hf_lb_x;
    float u = 0.0f;
    if (span > 1e-12f) u = (x - g->hf_lb_x) / span;
    u = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
    float sgrid = u * (float)(nx - 1);
    int i0 = cvt_i32(floorf(sgrid));
    int i1 = i0 + 1;
    if (i1 >= nx) i1 = nx - 1;
    if (i0 < 0) i0 = 0;
    float t = sgrid - (float)i0;
    float h = g->heights[i0] + (g->heights[i1] - g->heights[i0]) * t;
    int ir = cvt_i32(roundf(sgrid));
    if (ir < 0) ir = 0;
    if (ir >= nx) ir = nx - 1;
    int j0 = ir - 1, j1 = ir + 1;
    if (j0 < 0) j0 = 0;
    if (j1 >= nx) j1 = nx - 1;
    float dxw = span / (float)(nx - 1);
    float denom = dxw * (float)(j1 - j0);
    float dhdx = 0.0f;
    if (denom > 1e-6f) dhdx = (g->heights[j1] - g->heights[j0]) / denom;
    float n0 = -dhdx, n1 = 1.0f;
    if (g->reverse == 1) {
        n0 = -n0;
        n1 = -n1;
    }
    float nlen = sqrtf(n0 * n0 + n1 * n1);
    if (nlen > 1e-12f) {
        n0 = n0 / nlen;
        n1 = n1 / nlen;
    }
    n[0] = n0;
    n[1] = n1;
    return (x - x) * n0 + (z - h) * n1;
}

/* flatworld/mixedcontact.py:220-302 (_voxel_signed_distance_2d). */
static float voxel_signed(const ora_ground* g, const float p[2], float n[2]) {
    const int nx = g->vox_nx, ny = g->vox_ny;
    const int32_t* occ = g->occ;
    float best_d = 1e9f;
    n[0] = 0.0f;
    n[1] = 1.0f;
    int i = cvt_i32(floorf((p[0] - g->vox_lb[0]) / g->vox_dx));
    int j = cvt_i32(floorf((p[1] - g->vox_lb[1]) / g->vox_dz));
    if (0 <= i && i < nx && 0 <= j && j < ny && occ[i * ny + j] == 1) {
        float dnx = 1e9f, dpx = 1e9f, dny = 1e9f, dpy = 1e9f;
        for (int step = 0; step < nx; ++step) {
            int idx = i - 1 - step;
            if (idx < 0 || occ[idx * ny + j] == 0) {
                float bx = g->vox_lb[0] + (float)(idx + 1) * g->vox_dx;
                dnx = p[0] - bx;
                break;
            }
        }
        for (int step = 0; step < nx; ++step) {
            int idx = i + 1 + step;
            if (idx >= nx || occ[idx * ny + j] == 0) {
                float bx = g->vox_lb[0] + (float)idx * g->vox_dx;
                dpx = bx - p[0];
                break;
            }
        }
        for (int step = 0; step < ny; ++step) {
            int idx = j - 1 - step;
            if (idx < 0 || occ[i * ny + idx] == 0) {
                float by = g->vox_lb[1] + (float)(idx + 1) * g->vox_dz;
                dny = p[1] - by;
                break;
            }
        }
        for (int step = 0; step < ny; ++step) {
            int idx = j + 1 + step;
            if (idx >= ny || occ[i * ny + idx] == 0) {
                float by = g->vox_lb[1] + (float)idx * g->vox_dz;
                dpy = by - p[1];
                break;
            }
        }
        float mv = dnx;
        int face = 0;
        if (dpx < mv) { mv = dpx; face = 1; }
        if (dny < mv) { mv = dny; face = 2; }
        if (dpy < mv) { mv = dpy; face = 3; }
        if (face == 0) { best_d = -dnx; n[0] = -1.0f; n[1] = 0.0f; }
        else if (face == 1) { best_d = -dpx; n[0] = 1.0f; n[1] = 0.0f; }
        else if (face == 2) { best_d = -dny; n[0] = 0.0f; n[1] = -1.0f; }
        else { best_d = -dpy; n[0] = 0.0f; n[1] = 1.0f; }
    }
    return best_d;
}

/* Force of one ground on one boundary node; returns 1 when it pushes.
 * Plane: flatworld/mixedcontact.py:660-684 (no damping).
 * Height field: flatworld/mixedcontact.py:1566-1602.
 * Voxel: flatworld/mixedcontact.py:1605-1643. */
static int32_t rigid_force(const ora_ground* g, const float p[2], const float vel[2], float force[2]);

int32_t ora_ground_force(const ora_ground* g, const float p[2], const float vel[2], float dt, float force[2]) {
    const float k = g->penalty;
    const float damping_coeff = 0.5f;
    force[0] = 0.0f;
    force[1] = 0.0f;
    if (g->kind == ORA_PLANE) {
        float pen = (p[0] - g->point[0]) * g->normal[0] + (p[1] - g->point[1]) * g->normal[1];
        if (pen < 0.0f) {
            force[0] = -((g->normal[0] * k) * pen);
            force[1] = -((g->normal[1] * k) * pen);
            return 1;
        }
        return 0;
    }
    if (g->kind == ORA_HEIGHTFIELD) {
        float n[2];
        float sd = hf_signed(g, p[0], p[1], n);
        if (sd < 0.0f) {
            float c0 = ((-n[0]) * k) * sd;
            float c1 = ((-n[1]) * k) * sd;
            float vn = vel[0] * n[0] + vel[1] * n[1];
            if (vn < 0.0f) {
                c0 = c0 + (((((-n[0]) * damping_coeff) * k) * vn) * dt);
                c1 = c1 + (((((-n[1]) * damping_coeff) * k) * vn) * dt);
            }
            force[0] = c0;
            force[1] = c1;
            return 1;
        }
        return 0;
    }
    if (g->kind == ORA_RIGID) return rigid_force(g, p, vel, force);
    if (g->kind == ORA_VOXEL) {
        float n[2];
        float sd = voxel_signed(g, p, n);
        if (sd < 0.0f) {
            float pen = -sd;
            float c0 = (n[0] * k) * pen;
            float c1 = (n[1] * k) * pen;
            float vn = vel[0] * n[0] + vel[1] * n[1];
            if (vn < 0.0f) {
                c0 = c0 + (((((-n[0]) * damping_coeff) * k) * vn) * dt);
                c1 = c1 + (((((-n[1]) * damping_coeff) * k) * vn) * dt);
            }
            force[0] = c0;
            force[1] = c1;
            return 1;
        }
        return 0;
    }
    return 0;
}

/* ---------------------------------------------------------------- rigid primitives
 * Warp's vector helpers as the reference kernels use them: dot and length are a
 * zero-initialised sum of products, a mat22-vec2 product is (m_i0 v_0) + (m_i1 v_1),
 * vec / scalar and vec * scalar act per component. */
static float rdot(float ax, float ay, float bx, float by) { return (0.0f + ax * bx) + ay * by; }
static float rlen(float x, float y) { return sqrtf(rdot(x, y, x, y)); }
static float rmaxf(float a, float b) { return a > b ? a : b; }
static float rminf(float a, float b) { return a < b ? a : b; }
static float rmaxf_(float a, float b) { return a > b ? a : b; }

/* sphereSDFQuery (flatworld/utils.py:34-40) */
static float sphere_sdf(const float p[2], const float c[2], float radius, float n[2]) {
    const float dx = p[0] - c[0], dy = p[1] - c[1];
    const float nlen = rlen(dx, dy);
    const float pen = nlen - radius;
    const float d = rmaxf(nlen, 1e-9f);
    n[0] = dx / d;
    n[1] = dy / d;
    return pen;
}

/* obbSDFQuery (flatworld/utils.py:57-131): rot row-major [[r0, r1], [r2, r3]] */
static float obb_sdf(const float p[2], const float c[2], const float ext[2], const float rot[4], float n[2]) {
    float best = -1e9f;
    float nl0 = 0.0f, nl1 = 0.0f;
    const float h0 = 0.5f * ext[0], h1 = 0.5f * ext[1];
    const float q0 = p[0] - c[0], q1 = p[1] - c[1];
    /* transpose(rot) @ q */
    const float l0 = rot[0] * q0 + rot[2] * q1;
    const float l1 = rot[1] * q0 + rot[3] * q1;
    const float sd0 = l0 - h0, sd1 = -l0 - h0, sd2 = l1 - h1, sd3 = -l1 - h1;
    if (sd0 > best) { best = sd0; nl0 = 1.0f; nl1 = 0.0f; }
    if (sd1 > best) { best = sd1; nl0 = -1.0f; nl1 = 0.0f; }
    if (sd2 > best) { best = sd2; nl0 = 0.0f; nl1 = 1.0f; }
    if (sd3 > best) { best = sd3; nl0 = 0.0f; nl1 = -1.0f; }
    if (sd0 > 0.0f || sd1 > 0.0f || sd2 > 0.0f || sd3 > 0.0f) {
        float cx = l0, cy = l1;
        if (l0 < -h0) cx = -h0;
        else if (l0 > h0) cx = h0;
        if (l1 < -h1) cy = -h1;
        else if (l1 > h1) cy = h1;
        const float dx = l0 - cx, dy = l1 - cy;
        const float dist = rlen(dx, dy);
        if (dist > 1e-9f) {
            nl0 = dx / dist;
            nl1 = dy / dist;
        }
        best = dist;
    }
    n[0] = rot[0] * nl0 + rot[1] * nl1;
    n[1] = rot[2] * nl0 + rot[3] * nl1;
    return best;
}

/* capsuleSDFQuery with calMinDisNode2Segment (flatworld/utils.py:43-55, 133-145):
 * segment from lc to 2 origin - lc */
static float capsule_sdf(const float p[2], const float origin[2], const float lc[2], float radius, float n[2]) {
    const float qx = 2.0f * origin[0] - lc[0], qy = 2.0f * origin[1] - lc[1];
    const float apx = p[0] - lc[0], apy = p[1] - lc[1];
    const float abx = qx - lc[0], aby = qy - lc[1];
    const float ab_sq = rdot(abx, aby, abx, aby);
    float t = rdot(apx, apy, abx, aby) / ab_sq;
    t = rmaxf(0.0f, rminf(1.0f, t));
    const float clx = lc[0] + t * abx, cly = lc[1] + t * aby;
    const float dis = rlen(p[0] - clx, p[1] - cly);
    const float pen = dis - radius;
    const float nvx = p[0] - clx, nvy = p[1] - cly;
    const float nlen = rlen(nvx, nvy);
    const float d = rmaxf(nlen, 1e-9f);
    n[0] = nvx / d;
    n[1] = nvy / d;
    return pen;
}

float ora_rigid_detect(const ora_ground* g, const float p[2], float normal[2], float cpoint[2]) {
    float pen = 1e9f;
    normal[0] = 0.0f;
    normal[1] = 0.0f;
    cpoint[0] = p[0];
    cpoint[1] = p[1];
    if (g->rigid_type == ORA_RIGID_BALL) {
        pen = sphere_sdf(p, g->rigid_center, g->rigid_radius, normal);
    } else if (g->rigid_type == ORA_RIGID_BOX) {
        pen = obb_sdf(p, g->rigid_center, g->rigid_prim, g->rigid_rot, normal);
    } else if (g->rigid_type == ORA_RIGID_CAPSULE) {
        const float* r = g->rigid_rot;
        const float lc[2] = {g->rigid_center[0] + (r[0] * g->rigid_prim[0] + r[1] * g->rigid_prim[1]),
                             g->rigid_center[1] + (r[2] * g->rigid_prim[0] + r[3] * g->rigid_prim[1])};
        pen = capsule_sdf(p, g->rigid_center, lc, g->rigid_radius, normal);
    } else {
        return pen;
    }
    if (pen < 0.0f) {
        cpoint[0] = p[0] - normal[0] * pen;
        cpoint[1] = p[1] - normal[1] * pen;
    }
    return pen;
}

void ora_rigid_aabb(int32_t type, const float c[2], const float r[4], const float e[2], float radius, float lb[2],
                    float ub[2]) {
    if (type == ORA_RIGID_BALL) {
        lb[0] = c[0] - radius;
        lb[1] = c[1] - radius;
        ub[0] = c[0] + radius;
        ub[1] = c[1] + radius;
    } else if (type == ORA_RIGID_BOX) {
        /* corners centre + 0.5 rot (+-e0, +-e1); the first one as -0.5 * (rot @ e) */
        float cx[4], cy[4];
        const float s0[4] = {1.0f, 1.0f, 1.0f, -1.0f}, s1[4] = {1.0f, -1.0f, 1.0f, 1.0f};
        for (int k = 0; k < 4; ++k) {
            const float e0 = s0[k] * e[0], e1 = s1[k] * e[1];
            const float f = k == 0 ? -0.5f : 0.5f;
            cx[k] = c[0] + f * (r[0] * e0 + r[1] * e1);
            cy[k] = c[1] + f * (r[2] * e0 + r[3] * e1);
        }
        lb[0] = rminf(rminf(cx[0], cx[1]), rminf(cx[2], cx[3]));
        lb[1] = rminf(rminf(cy[0], cy[1]), rminf(cy[2], cy[3]));
        ub[0] = rmaxf(rmaxf(cx[0], cx[1]), rmaxf(cx[2], cx[3]));
        ub[1] = rmaxf(rmaxf(cy[0], cy[1]), rmaxf(cy[2], cy[3]));
    } else if (type == ORA_RIGID_CAPSULE) {
        const float alen = rlen(e[0], e[1]);
        const float a0 = e[0] / alen, a1 = e[1] / alen;
        const float d0 = r[0] * a0 + r[1] * a1, d1 = r[2] * a0 + r[3] * a1;
        const float p00 = c[0] - d0 * alen, p01 = c[1] - d1 * alen;
        const float p10 = c[0] + d0 * alen, p11 = c[1] + d1 * alen;
        const float perp0 = radius * sqrtf(rmaxf(0.0f, 1.0f - d0 * d0)), axial0 = radius * fabsf(d0);
        const float perp1 = radius * sqrtf(rmaxf(0.0f, 1.0f - d1 * d1)), axial1 = radius * fabsf(d1);
        lb[0] = (rminf(p00, p10) - perp0) - axial0;
        lb[1] = (rminf(p01, p11) - perp1) - axial1;
        ub[0] = (rmaxf(p00, p10) + perp0) + axial0;
        ub[1] = (rmaxf(p01, p11) + perp1) + axial1;
    }
}

/* _batched_rigid_prim_kernel (flatworld/mixedcontact.py:1025-1117), the FEM node's
 * part: nothing outside the rigid's AABB; a tied node is pulled to its contact point
 * while within the margin (half the AABB diagonal); otherwise a penalty force on
 * penetration deeper than 1e-4, with Coulomb friction against the rigid's surface
 * velocity.  Returns 1 when a force is applied. */
static int32_t rigid_force(const ora_ground* g, const float p[2], const float vel[2], float force[2]) {
    const float* lb = g->aabb_lo;
    const float* ub = g->aabb_hi;
    if (rminf(p[0] - lb[0], p[1] - lb[1]) < 0.0f || rminf(ub[0] - p[0], ub[1] - p[1]) < 0.0f) return 0;
    float n[2], cp[2];
    const float pen = ora_rigid_detect(g, p, n, cp);
    const float margin = rlen(ub[0] - lb[0], ub[1] - lb[1]) * 0.5f;
    const float k = g->penalty;
    if (g->rigid_tied == 1 && pen < margin) {
        force[0] = k * (cp[0] - p[0]);
        force[1] = k * (cp[1] - p[1]);
        return 1;
    }
    if (!(pen < margin)) return 0;
    const float penetration = pen - 1e-4f;
    if (!(penetration < 0.0f)) return 0;
    const float nf0 = ((-n[0]) * k) * penetration, nf1 = ((-n[1]) * k) * penetration;
    float t0 = nf0, t1 = nf1;
    const float mu = g->rigid_friction;
    if (mu > 1e-9f) {
        const float rx = cp[0] - g->rigid_center[0], ry = cp[1] - g->rigid_center[1];
        const float w = g->rigid_omega;
        const float sv0 = g->rigid_vel[0] + (-w * ry), sv1 = g->rigid_vel[1] + w * rx;
        const float rv0 = vel[0] - sv0, rv1 = vel[1] - sv1;
        const float vn = rdot(rv0, rv1, n[0], n[1]);
        const float tv0 = rv0 - vn * n[0], tv1 = rv1 - vn * n[1];
        const float tlen = rlen(tv0, tv1);
        if (tlen > 1e-9f) {
            const float fd0 = -tv0 / tlen, fd1 = -tv1 / tlen;
            const float nflen = rlen(nf0, nf1);
            t0 = t0 + (fd0 * mu) * nflen;
            t1 = t1 + (fd1 * mu) * nflen;
        }
    }
    force[0] = t0;
    force[1] = t1;
    return 1;
}

/* ---------------------------------------------------------------- rigid meshes
 * _batched_rigid_mesh_kernel_sh with its workset filter, and the mesh tie kernels
 * (flatworld/mixedcontact.py:519-588, 1119-1401, 1402-1565), serially in work-item
 * order, for the FEM node's side (the reaction impulses on the rigid are not
 * formed).  The loop's mesh rigids and their spatial hash come from the host. */

static int32_t rm_outside(const float p[2], float lb0, float lb1, float ub0, float ub1) {
    /* _point_outside_aabb (mixedcontact.py:113-119) */
    return rminf(p[0] - lb0, p[1] - lb1) < 0.0f || rminf(ub0 - p[0], ub1 - p[1]) < 0.0f;
}

static float rm_dist_sq(const float p[2], float lb0, float lb1, float ub0, float ub1) {
    /* _point_aabb_distance_sq (mixedcontact.py:122-132) */
    float d2 = 0.0f, v = 0.0f;
    if (p[0] < lb0) v = lb0 - p[0];
    else if (p[0] > ub0) v = p[0] - ub0;
    d2 = d2 + v * v;
    v = 0.0f;
    if (p[1] < lb1) v = lb1 - p[1];
    else if (p[1] > ub1) v = p[1] - ub1;
    d2 = d2 + v * v;
    return d2;
}

typedef struct {
    float pen, nrm[2], w[2];
    int32_t elem, did, found, has_cand;
} rm_best;

/* detectPointToMeshBoundary on edge e (contact_detection.py:64-90). */
static float rm_detect(const ora_ground* g, const float p[2], int32_t e, float limit, float nrm[2], float w[2]) {
    const float* a = g->mesh_coords + 2 * g->mesh_conn[2 * e];
    const float* b = g->mesh_coords + 2 * g->mesh_conn[2 * e + 1];
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

/* _process_rigid_sh_cell (mixedcontact.py:382-467): every edge of rigid `target` in
 * cell cid, each improvement written to the item's cache. */
static void rm_cell(const ora_ground* g, int32_t cid, const float p[2], int32_t target, int32_t item, rm_best* b) {
    if (!(cid >= 0 && cid < g->mesh_tc)) return;
    const float best_abs = fabsf(b->pen);
    if (b->found == 1 && best_abs < 1e8f) {
        const int32_t nx = g->mesh_cn[0], ix = cid % nx, iy = cid / nx;
        const float l0 = g->mesh_glb[0] + (float)ix * g->mesh_gs[0], l1 = g->mesh_glb[1] + (float)iy * g->mesh_gs[1];
        const float u0 = l0 + g->mesh_gs[0], u1 = l1 + g->mesh_gs[1];
        if (rm_dist_sq(p, l0, l1, u0, u1) > best_abs * best_abs) return;
    }
    for (int32_t q = g->mesh_cell[2 * cid]; q < g->mesh_cell[2 * cid + 1]; ++q) {
        const int32_t ei = g->mesh_sorted[q];
        if (ei < 0) continue;
        const int32_t did = g->mesh_reg[2 * ei], e = g->mesh_reg[2 * ei + 1];
        if (did != target) continue;
        b->has_cand = 1;
        float l0 = 1e30f, l1 = 1e30f, u0 = -1e30f, u1 = -1e30f;
        for (int vj = 0; vj < 2; ++vj) {
            const float* c = g->mesh_coords + 2 * g->mesh_conn[2 * e + vj];
            l0 = rminf(l0, c[0]);
            l1 = rminf(l1, c[1]);
            u0 = c[0] > u0 ? c[0] : u0;
            u1 = c[1] > u1 ? c[1] : u1;
        }
        const float margin = rmaxf_(1e-4f, g->mesh_margin[e]);
        l0 = l0 - margin;
        l1 = l1 - margin;
        u0 = u0 + margin;
        u1 = u1 + margin;
        const float best_abs2 = fabsf(b->pen);
        if (b->found == 1 && best_abs2 < 1e8f && rm_dist_sq(p, l0, l1, u0, u1) > best_abs2 * best_abs2) continue;
        if (!rm_outside(p, l0, l1, u0, u1)) {
            float nrm[2], w[2];
            const float pen = rm_detect(g, p, e, margin, nrm, w);
            if (fabsf(pen) < fabsf(b->pen)) {
                b->pen = pen;
                b->nrm[0] = nrm[0];
                b->nrm[1] = nrm[1];
                b->w[0] = w[0];
                b->w[1] = w[1];
                b->elem = e;
                b->did = did;
                b->found = 1;
                g->st_cache[2 * item] = e;
                g->st_cache[2 * item + 1] = did;
            }
        }
    }
}

/* Query cells of p with buffer qb, clamped (mixedcontact.py:1226-1243). */
static void rm_cells(const ora_ground* g, const float p[2], float qb, int32_t r[4]) {
    const float d0 = (p[0] - qb) - g->mesh_glb[0], d1 = (p[1] - qb) - g->mesh_glb[1];
    const float e0 = (p[0] + qb) - g->mesh_glb[0], e1 = (p[1] + qb) - g->mesh_glb[1];
    int32_t lx = (int32_t)floorf(d0 / g->mesh_gs[0]), ly = (int32_t)floorf(d1 / g->mesh_gs[1]);
    int32_t ux = (int32_t)floorf(e0 / g->mesh_gs[0]), uy = (int32_t)floorf(e1 / g->mesh_gs[1]);
    const int32_t cx = g->mesh_cn[0] - 1, cy = g->mesh_cn[1] - 1;
    lx = lx > 0 ? lx : 0;
    ly = ly > 0 ? ly : 0;
    lx = lx < cx ? lx : cx;
    ly = ly < cy ? ly : cy;
    ux = ux > 0 ? ux : 0;
    uy = uy > 0 ? uy : 0;
    ux = ux < cx ? ux : cx;
    uy = uy < cy ? uy : cy;
    r[0] = lx;
    r[1] = ly;
    r[2] = ux;
    r[3] = uy;
}

static void rm_clear(const ora_ground* g, int32_t k) {
    /* _clear_rigid_mesh_friction_state (mixedcontact.py:347-361) */
    g->st_pelem[k] = -1;
    g->st_pvalid[k] = 0;
    g->st_pw[2 * k] = 0.0f;
    g->st_pw[2 * k + 1] = 0.0f;
    g->st_pf[2 * k] = 0.0f;
    g->st_pf[2 * k + 1] = 0.0f;
    g->st_ppen[k] = 0.0f;
}

static void rm_surface(const ora_ground* g, int32_t e, const float w[2], float out[2]) {
    /* mesh_boundary_coords[c0] * w0 + mesh_boundary_coords[c1] * w1 */
    const float* a = g->mesh_coords + 2 * g->mesh_conn[2 * e];
    const float* b = g->mesh_coords + 2 * g->mesh_conn[2 * e + 1];
    out[0] = a[0] * w[0] + b[0] * w[1];
    out[1] = a[1] * w[0] + b[1] * w[1];
}

/* Force of the mesh record on work item k (node n at p); returns 1 when it applies one. */
static int32_t rigid_mesh_item(const ora_ground* g, int32_t k, const float p[2], int32_t counter, int32_t active,
                               float force[2]) {
    const float penalty = g->penalty;
    force[0] = force[1] = 0.0f;
    if (g->rigid_tied == 1) {
        /* _initialize_fem_rigid_mesh_ties_once_kernel at the first contact substep, then
           _apply_fem_rigid_mesh_ties_fixed_kernel every substep, whatever the gate */
        if (counter == 0 && g->st_tres[k] == 0 && g->mesh_tc > 0) {
            const float qb = rmaxf_(1e-4f, rmaxf_(g->mesh_gs[0], g->mesh_gs[1]));
            int32_t r[4];
            rm_cells(g, p, qb, r);
            rm_best b = {1e9f, {0.0f, 0.0f}, {0.0f, 0.0f}, -1, -1, 0, 0};
            for (int32_t ix = r[0]; ix <= r[2]; ++ix)
                for (int32_t iy = r[1]; iy <= r[3]; ++iy) rm_cell(g, ix + iy * g->mesh_cn[0], p, g->mesh_did, k, &b);
            if (b.found == 1) {
                g->st_tres[k] = 1;
                g->st_cache[2 * k] = b.elem;
                g->st_cache[2 * k + 1] = b.did;
                g->st_tw[2 * k] = b.w[0];
                g->st_tw[2 * k + 1] = b.w[1];
            }
        }
        if (g->st_tres[k] == 0) return 0;
        const int32_t e = g->st_cache[2 * k];
        const float* a = g->mesh_coords + 2 * g->mesh_conn[2 * e];
        const float* c = g->mesh_coords + 2 * g->mesh_conn[2 * e + 1];
        /* _eval_fixed_tie_point: (0, 0) + a * w0 + c * w1 */
        const float cp0 = (0.0f + a[0] * g->st_tw[2 * k]) + c[0] * g->st_tw[2 * k + 1];
        const float cp1 = (0.0f + a[1] * g->st_tw[2 * k]) + c[1] * g->st_tw[2 * k + 1];
        force[0] = penalty * (cp0 - p[0]);
        force[1] = penalty * (cp1 - p[1]);
        return 1;
    }
    if (counter == g->mesh_reset_at) g->st_skip[k] = -1;  /* explicitloop.py:543-553 */
    /* _build_active_rigid_mesh_workset_kernel (mixedcontact.py:519-588) */
    int32_t keep = active != 0 && g->st_skip[k] != 1 &&
                   !rm_outside(p, g->aabb_lo[0] - 1e-4f, g->aabb_lo[1] - 1e-4f, g->aabb_hi[0] + 1e-4f,
                               g->aabb_hi[1] + 1e-4f);
    if (!keep) {
        rm_clear(g, k);
        return 0;
    }
    const int32_t target = g->mesh_did;
    const float qbs = rmaxf_(1e-4f, rmaxf_(g->mesh_gs[0], g->mesh_gs[1]));
    rm_best b = {1e9f, {0.0f, 0.0f}, {0.0f, 0.0f}, -1, 0, 0, 0};
    const int32_t cel = g->st_cache[2 * k], cdid = g->st_cache[2 * k + 1];
    if (cel >= 0 && cdid == target) {
        const float margin = rmaxf_(1e-4f, g->mesh_margin[cel]);
        const float l0 = g->mesh_elb[2 * cel] - margin, l1 = g->mesh_elb[2 * cel + 1] - margin;
        const float u0 = g->mesh_eub[2 * cel] + margin, u1 = g->mesh_eub[2 * cel + 1] + margin;
        if (!rm_outside(p, l0, l1, u0, u1)) {
            float nrm[2], w[2];
            const float pen = rm_detect(g, p, cel, margin, nrm, w);
            if (fabsf(pen) < fabsf(b.pen)) {
                b.pen = pen;
                b.nrm[0] = nrm[0];
                b.nrm[1] = nrm[1];
                b.w[0] = w[0];
                b.w[1] = w[1];
                b.elem = cel;
                b.did = cdid;
                b.found = 1;
            }
        }
    }
    const int32_t run_full = !(g->mesh_subcycle > 1 && counter % g->mesh_subcycle != 0);
    if (b.found == 0 && (g->st_skip[k] == 1 || run_full == 0)) {
        rm_clear(g, k);
        return 0;
    }
    if (b.found == 0) {
        if (g->mesh_tc > 0) {
            int32_t r[4];
            rm_cells(g, p, qbs, r);
            int32_t local_first = 0, h[4] = {0, 0, 0, 0};
            if (cel >= 0 && cdid == target) {
                local_first = 1;
                /* _point_to_cell_coord of the cached edge's box centre (mixedcontact.py:322-344) */
                const float c0 = 0.5f * (g->mesh_elb[2 * cel] + g->mesh_eub[2 * cel]);
                const float c1 = 0.5f * (g->mesh_elb[2 * cel + 1] + g->mesh_eub[2 * cel + 1]);
                int32_t rx = (int32_t)floorf((c0 - g->mesh_glb[0]) / g->mesh_gs[0]);
                int32_t ry = (int32_t)floorf((c1 - g->mesh_glb[1]) / g->mesh_gs[1]);
                if (rx < 0) rx = 0;
                if (ry < 0) ry = 0;
                if (rx > g->mesh_cn[0] - 1) rx = g->mesh_cn[0] - 1;
                if (ry > g->mesh_cn[1] - 1) ry = g->mesh_cn[1] - 1;
                h[0] = rx > r[0] ? rx : r[0];
                h[1] = ry > r[1] ? ry : r[1];
                h[2] = rx < r[2] ? rx : r[2];
                h[3] = ry < r[3] ? ry : r[3];
                for (int32_t ix = h[0]; ix <= h[2]; ++ix)
                    for (int32_t iy = h[1]; iy <= h[3]; ++iy) rm_cell(g, ix + iy * g->mesh_cn[0], p, target, k, &b);
            }
            for (int32_t ix = r[0]; ix <= r[2]; ++ix)
                for (int32_t iy = r[1]; iy <= r[3]; ++iy) {
                    if (local_first && ix >= h[0] && ix <= h[2] && iy >= h[1] && iy <= h[3]) continue;
                    rm_cell(g, ix + iy * g->mesh_cn[0], p, target, k, &b);
                }
        }
        if (b.has_cand == 0) g->st_skip[k] = 1;
    }
    if (!(b.found == 1 && b.pen < 0.0f)) {
        rm_clear(g, k);
        return 0;
    }
    const float nf0 = ((-b.nrm[0]) * penalty) * b.pen, nf1 = ((-b.nrm[1]) * penalty) * b.pen;
    float t0 = nf0, t1 = nf1;
    const float mu = g->rigid_friction;
    if (mu > 1e-9f) {
        const float Fy = mu * rlen(nf0, nf1);
        const float kt = penalty * 1.0f;
        float pf0 = 0.0f, pf1 = 0.0f, s0 = 0.0f, s1 = 0.0f;
        if (g->st_pvalid[k] == 1) {
            float cur[2], prev[2];
            rm_surface(g, b.elem, b.w, cur);
            rm_surface(g, g->st_pelem[k], g->st_pw + 2 * k, prev);
            s0 = cur[0] - prev[0];
            s1 = cur[1] - prev[1];
            const float sd = rdot(s0, s1, b.nrm[0], b.nrm[1]);
            s0 = s0 - sd * b.nrm[0];
            s1 = s1 - sd * b.nrm[1];
            pf0 = g->st_pf[2 * k];
            pf1 = g->st_pf[2 * k + 1];
            const float pd = rdot(pf0, pf1, b.nrm[0], b.nrm[1]);
            pf0 = pf0 - pd * b.nrm[0];
            pf1 = pf1 - pd * b.nrm[1];
        }
        const float ft0 = pf0 + kt * s0, ft1 = pf1 + kt * s1;
        const float fn = rlen(ft0, ft1);
        float fr0, fr1;
        if (fn <= Fy) {
            fr0 = ft0;
            fr1 = ft1;
        } else {
            fr0 = (ft0 / fn) * Fy;
            fr1 = (ft1 / fn) * Fy;
        }
        g->st_pelem[k] = b.elem;
        g->st_pvalid[k] = 1;
        g->st_pw[2 * k] = b.w[0];
        g->st_pw[2 * k + 1] = b.w[1];
        g->st_pf[2 * k] = fr0;
        g->st_pf[2 * k + 1] = fr1;
        g->st_ppen[k] = b.pen;
        t0 = t0 + fr0;
        t1 = t1 + fr1;
    } else {
        rm_clear(g, k);
    }
    force[0] = t0;
    force[1] = t1;
    return 1;
}

/* Pair activation: inclusive AABB overlap of the FEM boundary AABB and the
 * ground AABB, both widened by 0.1 * extent on full BVH rebuilds
 * (counter % 100 == 0) and raw on refits.
 * flatworld/explicitloop.py:477-487, flatworld/bvh.py:39-46, 321-342, 393-477. */
int32_t ora_gate(const float fem_lb[2], const float fem_ub[2], const ora_ground* g, int32_t counter) {
    float alb[2] = {fem_lb[0], fem_lb[1]}, aub[2] = {fem_ub[0], fem_ub[1]};
    float blb[2] = {g->aabb_lo[0], g->aabb_lo[1]}, bub[2] = {g->aabb_hi[0], g->aabb_hi[1]};
    if (counter % 100 == 0) {
        for (int k = 0; k < 2; ++k) {
            float ba = 0.1f * (aub[k] - alb[k]);
            float la = alb[k] - ba, ua = aub[k] + ba;
            float bb = 0.1f * (bub[k] - blb[k]);
            float lb = blb[k] - bb, ub = bub[k] + bb;
            alb[k] = la;
            aub[k] = ua;
            blb[k] = lb;
            bub[k] = ub;
        }
    }
    return alb[0] <= bub[0] && aub[0] >= blb[0] && alb[1] <= bub[1] && aub[1] >= blb[1];
}

/* flatworld/femspringmanager.py:361-408 (one FEM domain). */
void ora_update_bbox(ora_scene* s) {
    float lb0 = 1e30f, lb1 = 1e30f, ub0 = -1e30f, ub1 = -1e30f;
    for (int32_t k = 0; k < s->n_bnodes; ++k) {
        int32_t n = s->bnodes[k];
        float cx = s->x[2 * n], cy = s->x[2 * n + 1];
        lb0 = lb0 < cx ? lb0 : cx;
        lb1 = lb1 < cy ? lb1 : cy;
        ub0 = ub0 > cx ? ub0 : cx;
        ub1 = ub1 > cy ? ub1 : cy;
    }
    s->aabb[0] = lb0;
    s->aabb[1] = lb1;
    s->aabb[2] = ub0;
    s->aabb[3] = ub1;
}

/* ------------------------------------------------------------------ driver */

/* flatworld/femspringmanager.py:113-138 */
void ora_prepare(ora_scene* s) {
    if (s->mat_type == ORA_SPRING) {
        /* np.full(nnodes, float(domain.mass), float32); springs have no area or J^-1 */
        for (int32_t i = 0; i < s->n_nodes; ++i) s->mass[i] = s->rho;
        ora_update_bbox(s);
        return;
    }
    for (int32_t i = 0; i < s->n_nodes; ++i) s->mass[i] = 0.0f;
    for (int32_t e = 0; e < s->n_elems; ++e) {
        const int32_t n1 = s->conn[3 * e], n2 = s->conn[3 * e + 1], n3 = s->conn[3 * e + 2];
        const float* x = s->x;
        float Jac[4] = {x[2 * n1] - x[2 * n3], x[2 * n2] - x[2 * n3], x[2 * n1 + 1] - x[2 * n3 + 1],
                        x[2 * n2 + 1] - x[2 * n3 + 1]};
        float vol = fabsf(Jac[0] * Jac[3] - Jac[1] * Jac[2]) * 0.5f;
        s->area[e] = vol;
        mat_inv(Jac, s->jinv + 4 * e);
        float nd_mass = vol * s->rho / 3.0f;
        s->mass[n1] += nd_mass;
        s->mass[n2] += nd_mass;
        s->mass[n3] += nd_mass;
    }
    ora_update_bbox(s);
}

void ora_substep(ora_scene* s, float dt, float damping, int32_t counter, int32_t with_contact) {
    const int32_t N = s->n_nodes;
    float* fx = s->fext;
    for (int32_t i = 0; i < 2 * N; ++i) fx[i] = 0.0f; /* _clear_fext_kernel of the previous substep */
    if (s->contact_flag) memset(s->contact_flag, 0, (size_t)N);

    /* ground forces at (x_n, v_n) in work-list order: pairs, then boundary nodes
       (flatworld/mixedcontact.py:1654-1754) */
    if (with_contact) {
        for (int32_t gi = 0; gi < s->n_grounds; ++gi) {
            const ora_ground* g = s->grounds + gi;
            if (g->kind == ORA_FLEX) {
                /* contributions in ascending work item, after the plane kernel */
                for (int32_t k = 0; k < s->n_bnodes; ++k) {
                    int32_t n = s->bnodes[k];
                    for (int32_t j = g->flex_off[n]; j < g->flex_off[n + 1]; ++j) {
                        fx[2 * n] += g->flex_f[2 * j];
                        fx[2 * n + 1] += g->flex_f[2 * j + 1];
                    }
                }
                continue;
            }
            if (g->kind == ORA_RIGID && g->rigid_type == ORA_RIGID_MESH) {
                /* the mesh kernels gate their non-tied items themselves and apply ties
                   whatever the gate (mixedcontact.py:1716-1742) */
                const int32_t act = ora_gate(s->aabb, s->aabb + 2, g, counter);
                for (int32_t k = 0; k < s->n_bnodes; ++k) {
                    int32_t n = s->bnodes[k];
                    float f[2];
                    if (rigid_mesh_item(g, k, s->x + 2 * n, counter, act, f)) {
                        fx[2 * n] += f[0];
                        fx[2 * n + 1] += f[1];
                        if (s->contact_flag) s->contact_flag[n] = 1;
                    }
                }
                continue;
            }
            if (!ora_gate(s->aabb, s->aabb + 2, g, counter)) continue;
            for (int32_t k = 0; k < s->n_bnodes; ++k) {
                int32_t n = s->bnodes[k];
                float f[2];
                if (ora_ground_force(g, s->x + 2 * n, s->v + 2 * n, dt, f)) {
                    fx[2 * n] += f[0];
                    fx[2 * n + 1] += f[1];
                    if (s->contact_flag) s->contact_flag[n] = 1;
                }
            }
        }
    }

    /* spring forces, ascending element id (flatworld/femspringmanager.py:969-983, 283-312) */
    if (s->mat_type == ORA_SPRING) {
        for (int32_t e = 0; e < s->n_elems; ++e) {
            float f[2];
            ora_spring_force(s, e, f);
            const int32_t ia = s->conn[2 * e], ib = s->conn[2 * e + 1];
            fx[2 * ia] += -f[0];
            fx[2 * ia + 1] += -f[1];
            fx[2 * ib] += f[0];
            fx[2 * ib + 1] += f[1];
        }
    }

    /* element forces, ascending element id (flatworld/femspringmanager.py:947-967) */
    for (int32_t e = 0; e < s->n_elems && s->mat_type != ORA_SPRING; ++e) {
        float f[6], cauchy[3], strain[3], epn[3], eqn = 0.0f;
        int32_t yielded = 0;
        ora_element_force(s, e, f, cauchy, strain, epn, &eqn, &yielded);
        if (s->mat_type == ORA_MISES) {
            s->eps_p[3 * e] = epn[0];
            s->eps_p[3 * e + 1] = epn[1];
            s->eps_p[3 * e + 2] = epn[2];
            s->eqps[e] = eqn;
        }
        if (s->yield_flag) s->yield_flag[e] = (uint8_t)yielded;
        if (s->stress) {
            s->stress[3 * e] = cauchy[0];
            s->stress[3 * e + 1] = cauchy[1];
            s->stress[3 * e + 2] = cauchy[2];
        }
        if (s->strain) {
            s->strain[3 * e] = strain[0];
            s->strain[3 * e + 1] = strain[1];
            s->strain[3 * e + 2] = strain[2];
        }
        for (int c = 0; c < 3; ++c) {
            int32_t n = s->conn[3 * e + c];
            fx[2 * n] += -f[2 * c];
            fx[2 * n + 1] += -f[2 * c + 1];
        }
    }

    /* symplectic Euler, flatworld/femspringmanager.py:315-358 */
    for (int32_t i = 0; i < N; ++i) {
        const int32_t bc = s->bc[i];
        const float m = s->mass[i];
        const float bx = s->bcv[2 * i], by = s->bcv[2 * i + 1];
        float f0 = fx[2 * i], f1 = fx[2 * i + 1];
        if (bc & 1) { /* UTYPE */
            f0 = 0.0f;
            f1 = 0.0f;
        } else if (bc & 4) { /* ATYPE */
            f0 = m * bx;
            f1 = m * by;
        } else if (bc & 16) { /* GRAVITY */
            f0 = f0 + m * bx;
            f1 = f1 + m * by;
        } else if (bc & 8) { /* FORCETYPE */
            f0 = f0 + bx;
            f1 = f1 + by;
        }
        fx[2 * i] = f0;
        fx[2 * i + 1] = f1;
        float rm = 1.0f / m;
        float a0 = 0.0f + (f0 - damping * s->v[2 * i]) * rm;
        float a1 = 0.0f + (f1 - damping * s->v[2 * i + 1]) * rm;
        float v0 = s->v[2 * i] + a0 * dt;
        float v1 = s->v[2 * i + 1] + a1 * dt;
        if (bc & 2) { /* VTYPE */
            v0 = bx;
            v1 = by;
        } else if (bc & 1) {
            v0 = 0.0f;
            v1 = 0.0f;
        }
        s->v[2 * i] = v0;
        s->v[2 * i + 1] = v1;
        float du0 = v0 * dt, du1 = v1 * dt;
        s->u[2 * i] = s->u[2 * i] + du0;
        s->u[2 * i + 1] = s->u[2 * i + 1] + du1;
        s->x[2 * i] = s->x[2 * i] + du0;
        s->x[2 * i + 1] = s->x[2 * i + 1] + du1;
    }

    ora_update_bbox(s);
}
