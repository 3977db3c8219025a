/*
 * CPU oracle for rigid contact generation — TEST INFRASTRUCTURE ONLY.
 *
 * Float32 restatement (compiled with -ffp-contract=off, IEEE sqrt/div) of the
 * reference collision path for primitive shapes:
 *   shape AABBs             sim/collision/pipeline.py:121-179
 *   explicit-pair AABB test sim/collision/broad_phase_common.py:18-35,
 *                           sim/collision/broad_phase_nxn.py:21-59
 *   primitive narrowphase   sim/collision/narrow_phase.py:141-577
 *   primitives              sim/collision/primitive.py:37-90 (segment helpers),
 *                           89-129 (plane-sphere, sphere-sphere), 132-166
 *                           (sphere-capsule), 169-264 (capsule-capsule),
 *                           267-319 (plane-capsule), 322-351 (plane-ellipsoid),
 *                           354-428 (plane-box), 431-501 (sphere-cylinder),
 *                           504-657 (plane-cylinder), 1150-1205 (sphere-box)
 *   gap check + writer      sim/collision/contact_data.py:36-56,
 *                           sim/collision/pipeline.py:51-118
 * Contacts are appended per world in candidate-pair order; a pair whose whole
 * batch does not fit into the world's capacity is skipped, like the
 * reference's overflow rule (narrow_phase.py:534-539).
 *
 * Only tests/, the smoke check and the CPU-baseline leg of bench.py may load
 * this library.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define GEO_PLANE 1
#define GEO_SPHERE 3
#define GEO_CAPSULE 4
#define GEO_ELLIPSOID 5
#define GEO_CYLINDER 6
#define GEO_BOX 7
#define GEO_CONE 9
#define CYL_FLAT_COS 0.9238795325112867f /* cos(22.5 deg), primitive.py:31-34 */
#define MAXVAL 1e10f
#define MINVAL 1e-15f

typedef struct { float x, y, z; } v3;
typedef struct { float x, y, z, w; } q4;

static v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3 scl(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static v3 cross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static float length(v3 a) { return sqrtf(dot(a, a)); }
static v3 normalize(v3 a) {
    float l = length(a);
    if (l > 0.0f) return mk(a.x / l, a.y / l, a.z / l);
    return mk(0.0f, 0.0f, 0.0f);
}
static float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }

static v3 qrot(q4 q, v3 v) {
    v3 qv = mk(q.x, q.y, q.z);
    v3 t = scl(cross(qv, v), 2.0f);
    return add(add(v, scl(t, q.w)), cross(qv, t));
}
static q4 qmul(q4 a, q4 b) {
    q4 r;
    r.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
    r.y = a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x;
    r.z = a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w;
    r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    return r;
}
typedef struct { v3 p; q4 q; } tf;
static tf load_tf(const float* a) {
    tf t;
    t.p = mk(a[0], a[1], a[2]);
    t.q.x = a[3]; t.q.y = a[4]; t.q.z = a[5]; t.q.w = a[6];
    return t;
}
static tf tf_mul(tf a, tf b) {
    tf r;
    r.p = add(a.p, qrot(a.q, b.p));
    r.q = qmul(a.q, b.q);
    return r;
}
static v3 tf_inv_point(tf a, v3 x) {
    q4 qi = a.q; qi.x = -qi.x; qi.y = -qi.y; qi.z = -qi.z;
    return qrot(qi, sub(x, a.p));
}
static v3 tf_inv_vector(tf a, v3 x) {
    q4 qi = a.q; qi.x = -qi.x; qi.y = -qi.y; qi.z = -qi.z;
    return qrot(qi, x);
}
/* R columns = rotated basis vectors; matvec R @ v and R^T @ v */
static void rot_cols(q4 q, v3 c[3]) {
    c[0] = qrot(q, mk(1.0f, 0.0f, 0.0f));
    c[1] = qrot(q, mk(0.0f, 1.0f, 0.0f));
    c[2] = qrot(q, mk(0.0f, 0.0f, 1.0f));
}
static v3 rmul(const v3 c[3], v3 v) { return add(add(scl(c[0], v.x), scl(c[1], v.y)), scl(c[2], v.z)); }
static v3 rtmul(const v3 c[3], v3 v) { return mk(dot(c[0], v), dot(c[1], v), dot(c[2], v)); }

/* ------------------------------------------------------------------ AABB */
static void shape_aabb(int type, v3 pos, q4 q, v3 scale, float radius, float eff_gap, v3* lo, v3* hi) {
    v3 h;
    if (type == GEO_PLANE && scale.x == 0.0f && scale.y == 0.0f) {
        h = mk(radius, radius, radius);
    } else if (type == GEO_SPHERE) {
        h = mk(scale.x, scale.x, scale.x);
    } else if (type == GEO_CAPSULE) {
        v3 a = qrot(q, mk(0.0f, 0.0f, 1.0f));
        h = mk(fabsf(a.x) * scale.y + scale.x, fabsf(a.y) * scale.y + scale.x, fabsf(a.z) * scale.y + scale.x);
    } else if (type == GEO_BOX) {
        v3 c[3];
        rot_cols(q, c);
        h = mk(fabsf(c[0].x) * scale.x + fabsf(c[1].x) * scale.y + fabsf(c[2].x) * scale.z,
               fabsf(c[0].y) * scale.x + fabsf(c[1].y) * scale.y + fabsf(c[2].y) * scale.z,
               fabsf(c[0].z) * scale.x + fabsf(c[1].z) * scale.y + fabsf(c[2].z) * scale.z);
    } else if (type == GEO_ELLIPSOID) {
        /* extent along world axis i: the support value sqrt(sum_k (R_ik s_k)^2) */
        v3 c[3];
        rot_cols(q, c);
        v3 a = scl(c[0], scale.x), b = scl(c[1], scale.y), d = scl(c[2], scale.z);
        h = mk(sqrtf(a.x * a.x + b.x * b.x + d.x * d.x), sqrtf(a.y * a.y + b.y * b.y + d.y * d.y),
               sqrtf(a.z * a.z + b.z * b.z + d.z * d.z));
    } else if (type == GEO_CYLINDER) {
        /* extent along world axis i: r sqrt(1 - a_i^2) + h |a_i| for the unit axis a */
        v3 a = qrot(q, mk(0.0f, 0.0f, 1.0f));
        h = mk(scale.x * sqrtf(fmaxf(1.0f - a.x * a.x, 0.0f)) + scale.y * fabsf(a.x),
               scale.x * sqrtf(fmaxf(1.0f - a.y * a.y, 0.0f)) + scale.y * fabsf(a.y),
               scale.x * sqrtf(fmaxf(1.0f - a.z * a.z, 0.0f)) + scale.y * fabsf(a.z));
    } else if (type == GEO_CONE) {
        /* apex a_w = +h along the axis, base rim radius r at -h: along world axis i the extremes
           are +-h a_i (apex / base centre) and -h a_i +- r sqrt(1 - a_i^2) (rim) */
        v3 a = qrot(q, mk(0.0f, 0.0f, 1.0f));
        float av[3] = {a.x, a.y, a.z}, lo3[3], hi3[3];
        for (int i = 0; i < 3; ++i) {
            const float rim = scale.x * sqrtf(fmaxf(1.0f - av[i] * av[i], 0.0f));
            const float ap = scale.y * av[i], bc = -scale.y * av[i];
            hi3[i] = fmaxf(ap, bc + rim);
            lo3[i] = fminf(ap, bc - rim);
        }
        *lo = sub(add(pos, mk(lo3[0], lo3[1], lo3[2])), mk(eff_gap, eff_gap, eff_gap));
        *hi = add(add(pos, mk(hi3[0], hi3[1], hi3[2])), mk(eff_gap, eff_gap, eff_gap));
        return;
    } else {
        h = mk(radius, radius, radius);
    }
    *lo = sub(sub(pos, h), mk(eff_gap, eff_gap, eff_gap));
    *hi = add(add(pos, h), mk(eff_gap, eff_gap, eff_gap));
}

static int aabb_overlap(v3 l1, v3 u1, v3 l2, v3 u2) {
    return l1.x <= u2.x && u1.x >= l2.x && l1.y <= u2.y && u1.y >= l2.y && l1.z <= u2.z && u1.z >= l2.z;
}

/* -------------------------------------------------------------- primitives */
static v3 closest_segment_point(v3 a, v3 b, v3 pt) {
    v3 ab = sub(b, a);
    float t = dot(sub(pt, a), ab) / (dot(ab, ab) + 1e-6f);
    return add(a, scl(ab, clampf(t, 0.0f, 1.0f)));
}

static void plane_sphere(v3 n, v3 ppos, v3 spos, float r, float* dist, v3* pos) {
    float d = dot(sub(spos, ppos), n) - r;
    *dist = d;
    *pos = sub(spos, scl(n, r + 0.5f * d));
}

static void sphere_sphere(v3 p1, float r1, v3 p2, float r2, float* dist, v3* pos, v3* n) {
    v3 dir = sub(p2, p1);
    float d = length(dir);
    v3 nn;
    if (d == 0.0f) nn = mk(1.0f, 0.0f, 0.0f);
    else nn = mk(dir.x / d, dir.y / d, dir.z / d);
    d = d - (r1 + r2);
    *pos = add(p1, scl(nn, r1 + 0.5f * d));
    *dist = d;
    *n = nn;
}

static void sphere_capsule(v3 sp, float sr, v3 cp, v3 axis, float cr, float hh, float* dist, v3* pos, v3* n) {
    v3 seg = scl(axis, hh);
    v3 pt = closest_segment_point(sub(cp, seg), add(cp, seg), sp);
    sphere_sphere(sp, sr, pt, cr, dist, pos, n);
}

static void capsule_capsule(v3 p1, v3 a1, float r1, float h1, v3 p2, v3 a2, float r2, float h2,
                            float dist[2], v3 pos[2], v3* normal) {
    dist[0] = MAXVAL; dist[1] = MAXVAL;
    v3 axis1 = scl(a1, h1);
    v3 axis2 = scl(a2, h2);
    v3 dif = sub(p1, p2);
    float ma = dot(axis1, axis1);
    float mb = -dot(axis1, axis2);
    float mc = dot(axis2, axis2);
    float u = -dot(axis1, dif);
    float v = dot(axis2, dif);
    float det = ma * mc - mb * mb;
    if (fabsf(det) >= MINVAL) {
        float inv = 1.0f / det;
        float x1 = (mc * u - mb * v) * inv;
        float x2 = (ma * v - mb * u) * inv;
        if (x1 > 1.0f) { x1 = 1.0f; x2 = (v - mb) / mc; }
        else if (x1 < -1.0f) { x1 = -1.0f; x2 = (v + mb) / mc; }
        if (x2 > 1.0f) { x2 = 1.0f; x1 = clampf((u - mb) / ma, -1.0f, 1.0f); }
        else if (x2 < -1.0f) { x2 = -1.0f; x1 = clampf((u + mb) / ma, -1.0f, 1.0f); }
        v3 v1 = add(p1, scl(axis1, x1));
        v3 v2 = add(p2, scl(axis2, x2));
        sphere_sphere(v1, r1, v2, r2, &dist[0], &pos[0], normal);
    } else {
        v3 v1 = add(p1, axis1);
        float x2 = clampf((v - mb) / mc, -1.0f, 1.0f);
        v3 v2 = add(p2, scl(axis2, x2));
        sphere_sphere(v1, r1, v2, r2, &dist[0], &pos[0], normal);
        v3 n2;
        v1 = sub(p1, axis1);
        x2 = clampf((v + mb) / mc, -1.0f, 1.0f);
        v2 = add(p2, scl(axis2, x2));
        sphere_sphere(v1, r1, v2, r2, &dist[1], &pos[1], &n2);
    }
}

static void plane_capsule(v3 n, v3 ppos, v3 cp, v3 axis, float r, float hh, float dist[2], v3 pos[2]) {
    v3 seg = scl(axis, hh);
    plane_sphere(n, ppos, add(cp, seg), r, &dist[0], &pos[0]);
    plane_sphere(n, ppos, sub(cp, seg), r, &dist[1], &pos[1]);
}

static void plane_ellipsoid(v3 n, v3 ppos, v3 epos, q4 eq, v3 size, float* dist, v3* pos) {
    v3 c[3];
    rot_cols(eq, c);
    v3 ln = rtmul(c, n);
    v3 sup = normalize(mk(ln.x * size.x, ln.y * size.y, ln.z * size.z));
    sup = mk(-sup.x, -sup.y, -sup.z);
    v3 p = add(epos, rmul(c, mk(sup.x * size.x, sup.y * size.y, sup.z * size.z)));
    float d = dot(n, sub(p, ppos));
    *dist = d;
    *pos = sub(p, scl(n, d * 0.5f));
}

static void sphere_cylinder(v3 sp, float sr, v3 cp, v3 axis, float cr, float hh, float* dist, v3* pos, v3* n) {
    v3 vec = sub(sp, cp);
    float x = dot(vec, axis);
    v3 a_proj = scl(axis, x);
    v3 p_proj = sub(vec, a_proj);
    float p_sqr = dot(p_proj, p_proj);
    int side = fabsf(x) < hh;
    int cap = p_sqr < cr * cr;
    if (side && cap) {
        float d_cap = hh - fabsf(x);
        float d_rad = cr - sqrtf(p_sqr);
        if (d_cap < d_rad) side = 0;
        else cap = 0;
    }
    if (side) {
        sphere_sphere(sp, sr, add(cp, a_proj), cr, dist, pos, n);
    } else if (cap) {
        v3 cpos, pn;
        if (x > 0.0f) { cpos = add(cp, scl(axis, hh)); pn = axis; }
        else { cpos = sub(cp, scl(axis, hh)); pn = mk(-axis.x, -axis.y, -axis.z); }
        plane_sphere(pn, cpos, sp, sr, dist, pos);
        *n = mk(-pn.x, -pn.y, -pn.z);
    } else {
        float l = sqrtf(p_sqr);
        float inv = 1.0f / (l != 0.0f ? l : 1e-15f);
        p_proj = scl(p_proj, cr * inv);
        v3 cap_off = scl(axis, (x < 0.0f ? -1.0f : 1.0f) * hh);
        sphere_sphere(sp, sr, add(add(cp, cap_off), p_proj), 0.0f, dist, pos, n);
    }
}

/* contacts of a cylinder on a plane: the deepest rim point first, then a fixed tripod on the near
   cap (near upright) or the two side generators and the cap-rim point facing the plane (rolling),
   each kept when it is farther than 1 % of the larger dimension from the deepest point */
static void plane_cylinder(v3 n, v3 ppos, v3 cc, v3 axis, float r, float hh, float dist[4], v3 pos[4]) {
    for (int i = 0; i < 4; ++i) { dist[i] = MAXVAL; pos[i] = mk(0.0f, 0.0f, 0.0f); }
    float dna = dot(n, axis);
    if (dna > 0.0f) { axis = mk(-axis.x, -axis.y, -axis.z); dna = -dna; }
    v3 cap = add(cc, scl(axis, hh));
    v3 pa = add(mk(-n.x, -n.y, -n.z), scl(axis, dna));
    float pa_sq = dot(pa, pa);
    int has_align = pa_sq > 1e-10f;
    if (has_align) pa = scl(pa, 1.0f / sqrtf(pa_sq));
    v3 ref = mk(1.0f, 0.0f, 0.0f);
    if (fabsf(dot(axis, ref)) > 0.9f) ref = mk(0.0f, 1.0f, 0.0f);
    v3 pf = normalize(sub(ref, scl(axis, dot(axis, ref))));
    float abs_dot = -dna;
    int flat = abs_dot >= CYL_FLAT_COS;
    v3 dperp = has_align ? pa : pf;
    v3 dpt = add(cap, scl(dperp, r));
    float dd = dot(sub(dpt, ppos), n);
    v3 dpos = sub(dpt, scl(n, dd * 0.5f));
    dist[0] = dd;
    pos[0] = dpos;
    int nc = 1;
    float thr = 0.01f * fmaxf(r, hh);
    float thr_sq = thr * thr;
    if (flat) {
        v3 u = scl(pf, r);
        v3 v = scl(cross(axis, pf), r);
        const float c120 = -0.5f, s120 = 0.8660254f;
        v3 pts[3];
        pts[0] = add(cap, u);
        pts[1] = add(add(cap, scl(u, c120)), scl(v, s120));
        pts[2] = sub(add(cap, scl(u, c120)), scl(v, s120));
        for (int k = 0; k < 3; ++k) {
            float d = dot(sub(pts[k], ppos), n);
            v3 p = sub(pts[k], scl(n, d * 0.5f));
            if (nc < 4 && dot(sub(p, dpos), sub(p, dpos)) > thr_sq) { dist[nc] = d; pos[nc] = p; ++nc; }
        }
    } else {
        v3 pr = has_align ? pa : pf;
        v3 u = scl(pr, r);
        v3 v = scl(cross(axis, pr), r);
        v3 pts[2];
        pts[0] = add(add(cc, scl(axis, hh)), u);
        pts[1] = add(sub(cc, scl(axis, hh)), u);
        for (int k = 0; k < 2; ++k) {
            float d = dot(sub(pts[k], ppos), n);
            v3 p = sub(pts[k], scl(n, d * 0.5f));
            if (nc < 4 && dot(sub(p, dpos), sub(p, dpos)) > thr_sq) { dist[nc] = d; pos[nc] = p; ++nc; }
        }
        v3 ppv = add(cap, v), pnv = sub(cap, v);
        float dpv = dot(sub(ppv, ppos), n), dnv = dot(sub(pnv, ppos), n);
        int use_p = dpv <= dnv;
        v3 pt = use_p ? ppv : pnv;
        float d = use_p ? dpv : dnv;
        v3 p = sub(pt, scl(n, d * 0.5f));
        if (nc < 4 && dot(sub(p, dpos), sub(p, dpos)) > thr_sq) { dist[nc] = d; pos[nc] = p; ++nc; }
    }
}

static void plane_box(v3 n, v3 ppos, v3 bpos, q4 bq, v3 size, float margin, float dist[4], v3 pos[4]) {
    v3 c[3];
    rot_cols(bq, c);
    float center_dist = dot(sub(bpos, ppos), n);
    for (int i = 0; i < 4; ++i) { dist[i] = MAXVAL; pos[i] = mk(0.0f, 0.0f, 0.0f); }
    int ncontact = 0, worst = 0;
    for (int i = 0; i < 8; ++i) {
        v3 corner = mk((i & 1) ? size.x : -size.x, (i & 2) ? size.y : -size.y, (i & 4) ? size.z : -size.z);
        corner = rmul(c, corner);
        float ldist = dot(n, corner);
        float cdist = center_dist + ldist;
        if (cdist > margin) continue;
        v3 cpos = sub(add(corner, bpos), scl(n, 0.5f * cdist));
        if (ncontact < 4) {
            dist[ncontact] = cdist;
            pos[ncontact] = cpos;
            if (ncontact == 0 || cdist > dist[worst]) worst = ncontact;
            ncontact++;
        } else if (cdist < dist[worst]) {
            dist[worst] = cdist;
            pos[worst] = cpos;
            worst = 0;
            if (dist[1] > dist[worst]) worst = 1;
            if (dist[2] > dist[worst]) worst = 2;
            if (dist[3] > dist[worst]) worst = 3;
        }
    }
}

static float vget(v3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
static void vset(v3* v, int i, float s) { if (i == 0) v->x = s; else if (i == 1) v->y = s; else v->z = s; }

static void sphere_box(v3 sp, float r, v3 bpos, q4 bq, v3 size, float* dist, v3* pos, v3* normal) {
    v3 c[3];
    rot_cols(bq, c);
    v3 center = rtmul(c, sub(sp, bpos));
    v3 clamped = mk(fmaxf(-size.x, fminf(size.x, center.x)), fmaxf(-size.y, fminf(size.y, center.y)),
                    fmaxf(-size.z, fminf(size.z, center.z)));
    v3 diff = sub(clamped, center);
    float d = length(diff);
    v3 cdir = diff;
    if (d != 0.0f) cdir = mk(diff.x / d, diff.y / d, diff.z / d);
    v3 lp;
    if (d <= 1e-6f) {
        float closest = 2.0f * (size.x + size.y + size.z);
        int k = 0;
        for (int i = 0; i < 6; ++i) {
            float sgn = (i % 2) ? 1.0f : -1.0f;
            float fd = fabsf(sgn * vget(size, i / 2) - vget(center, i / 2));
            if (closest > fd) { closest = fd; k = i; }
        }
        v3 nearest = mk(0.0f, 0.0f, 0.0f);
        vset(&nearest, k / 2, (k % 2) ? -1.0f : 1.0f);
        lp = add(center, scl(nearest, (r - closest) / 2.0f));
        *normal = rmul(c, nearest);
        *dist = -closest - r;
    } else {
        v3 deepest = add(center, scl(cdir, r));
        lp = scl(add(clamped, deepest), 0.5f);
        *normal = rmul(c, cdir);
        *dist = d - r;
    }
    *pos = add(bpos, rmul(c, lp));
}

#include "convex.h"

/* pairs (sorted types) the primitive chain leaves to the MPR / GJK path (narrow_phase.py:573-575) */
static int cvx_type(int t) {
    return t == GEO_SPHERE || t == GEO_CAPSULE || t == GEO_ELLIPSOID || t == GEO_CYLINDER || t == GEO_BOX ||
           t == GEO_CONE;
}
static int cvx_routed(int ta, int tb) { return (cvx_type(ta) && cvx_type(tb)) || (ta == GEO_PLANE && tb == GEO_CONE); }

/* Support point of a shape in its own frame (a test hook for the support maps). */
void sapx_oracle_support(int type, const float* scale, const float* d, float* out) {
    cvx_geom g = {type, mk(scale[0], scale[1], scale[2])};
    v3 p = cvx_support(&g, mk(d[0], d[1], d[2]));
    out[0] = p.x; out[1] = p.y; out[2] = p.z;
}

/* ------------------------------------------------------------ gap check */
static void contact_points(v3 center, v3 n_in, float d, float ra, float rb, v3* xa, v3* xb, v3* n) {
    *n = normalize(n_in);
    *xa = sub(center, scl(*n, 0.5f * d + ra));
    *xb = add(center, scl(*n, 0.5f * d + rb));
}
static int passes_gap(v3 center, v3 n_in, float d, float ra, float rb, float ma, float mb, float gap_sum) {
    v3 xa, xb, n;
    contact_points(center, n_in, d, ra, rb, &xa, &xb, &n);
    float total = ra + rb + ma + mb;
    float dd = dot(sub(xb, xa), n) - total;
    return dd <= gap_sum;
}
/* the writer's own rejection test: d > gap (a NaN distance is not rejected) */
static int passes_gap_not(v3 center, v3 n_in, float d, float ra, float rb, float ma, float mb, float gap_sum) {
    v3 xa, xb, n;
    contact_points(center, n_in, d, ra, rb, &xa, &xb, &n);
    float total = ra + rb + ma + mb;
    float dd = dot(sub(xb, xa), n) - total;
    return dd > gap_sum;
}

/*
 * Pair-level narrowphase.  Returns the number of candidate contacts (0..4)
 * and fills dist/pos/normal/radii/shape order.  Unsupported pair types set
 * *unsupported = 1.
 */
static int narrow_pair(int sa, int sb, const int* shape_type, const float* shape_scale, const tf* X,
                       const float* shape_margin, const float* aabb, float dist[5], v3 pos[5], v3* normal, float* ra_out, float* rb_out, int* shape_a_out, int* shape_b_out,
                       float gap_sum, int* unsupported, int* convex) {
    int ta = shape_type[sa], tb = shape_type[sb];
    if (ta > tb) { int t = sa; sa = sb; sb = t; t = ta; ta = tb; tb = t; }
    *shape_a_out = sa; *shape_b_out = sb;
    v3 scale_a = mk(shape_scale[3 * sa], shape_scale[3 * sa + 1], shape_scale[3 * sa + 2]);
    v3 scale_b = mk(shape_scale[3 * sb], shape_scale[3 * sb + 1], shape_scale[3 * sb + 2]);
    tf Xa = X[sa], Xb = X[sb];
    float ra = 0.0f, rb = 0.0f;
    if (ta == GEO_SPHERE || ta == GEO_CAPSULE) ra = scale_a.x;
    if (tb == GEO_SPHERE || tb == GEO_CAPSULE) rb = scale_b.x;
    for (int i = 0; i < 5; ++i) { dist[i] = MAXVAL; pos[i] = mk(0.0f, 0.0f, 0.0f); }
    *normal = mk(0.0f, 0.0f, 0.0f);
    *unsupported = 0;
    *convex = 0;
    v3 ez = mk(0.0f, 0.0f, 1.0f);
    if (ta == GEO_PLANE && tb == GEO_SPHERE) {
        v3 n = qrot(Xa.q, ez);
        plane_sphere(n, Xa.p, Xb.p, scale_b.x, &dist[0], &pos[0]);
        *normal = n;
    } else if (ta == GEO_PLANE && tb == GEO_BOX) {
        v3 n = qrot(Xa.q, ez);
        plane_box(n, Xa.p, Xb.p, Xb.q, scale_b, gap_sum, dist, pos);
        *normal = n;
    } else if (ta == GEO_SPHERE && tb == GEO_SPHERE) {
        sphere_sphere(Xa.p, scale_a.x, Xb.p, scale_b.x, &dist[0], &pos[0], normal);
    } else if (ta == GEO_PLANE && tb == GEO_CAPSULE) {
        v3 n = qrot(Xa.q, ez);
        v3 ax = qrot(Xb.q, ez);
        plane_capsule(n, Xa.p, Xb.p, ax, scale_b.x, scale_b.y, dist, pos);
        *normal = n;
    } else if (ta == GEO_SPHERE && tb == GEO_CAPSULE) {
        v3 ax = qrot(Xb.q, ez);
        sphere_capsule(Xa.p, scale_a.x, Xb.p, ax, scale_b.x, scale_b.y, &dist[0], &pos[0], normal);
    } else if (ta == GEO_CAPSULE && tb == GEO_CAPSULE) {
        v3 a1 = qrot(Xa.q, ez), a2 = qrot(Xb.q, ez);
        capsule_capsule(Xa.p, a1, scale_a.x, scale_a.y, Xb.p, a2, scale_b.x, scale_b.y, dist, pos, normal);
    } else if (ta == GEO_SPHERE && tb == GEO_BOX) {
        sphere_box(Xa.p, scale_a.x, Xb.p, Xb.q, scale_b, &dist[0], &pos[0], normal);
    } else if (ta == GEO_PLANE && tb == GEO_ELLIPSOID) {
        v3 n = qrot(Xa.q, ez);
        plane_ellipsoid(n, Xa.p, Xb.p, Xb.q, scale_b, &dist[0], &pos[0]);
        *normal = n;
    } else if (ta == GEO_PLANE && tb == GEO_CYLINDER) {
        v3 n = qrot(Xa.q, ez);
        v3 ax = qrot(Xb.q, ez);
        plane_cylinder(n, Xa.p, Xb.p, ax, scale_b.x, scale_b.y, dist, pos);
        *normal = n;
    } else if (ta == GEO_SPHERE && tb == GEO_CYLINDER) {
        v3 ax = qrot(Xb.q, ez);
        sphere_cylinder(Xa.p, scale_a.x, Xb.p, ax, scale_b.x, scale_b.y, &dist[0], &pos[0], normal);
    } else if (cvx_routed(ta, tb)) {
        /* every contact of a convex pair is kept, whatever its distance (the
           writer's own gap check decides) */
        *convex = 1;
        int tta = ta;
        v3 ssa = scale_a, ppa = Xa.p;
        if (ta == GEO_PLANE) {
            /* only an infinite plane (zero half extents) becomes a box proxy (core.py:680-697) */
            if (scale_a.x != 0.0f || scale_a.y != 0.0f) {
                *unsupported = 1;
                return 0;
            }
            /* the other shape's bounding sphere (centre and half diagonal of its collision AABB,
               narrow_phase.py:661-720) must reach the plane's half-space; the plane becomes a box
               of half extent 10 (R + gap) whose top face lies on the plane under the other shape
               (core.py:507-566) */
            const float* B = aabb + 6 * sb;
            const v3 lo = mk(B[0], B[1], B[2]), hi = mk(B[3], B[4], B[5]);
            const v3 c = scl(add(lo, hi), 0.5f);
            const float R = length(scl(sub(hi, lo), 0.5f));
            const v3 n = qrot(Xa.q, ez);
            if (dot(sub(c, Xa.p), n) > R) return 0;
            const float ext = (R + gap_sum) * 10.0f;
            const v3 surf = sub(Xb.p, scl(n, dot(sub(Xb.p, Xa.p), n)));
            ppa = sub(surf, scl(n, ext));
            tta = GEO_BOX;
            ssa = mk(ext, ext, ext);
        }
        int n = cvx_pair(tta, ssa, ppa, Xa.q, tb, scale_b, Xb.p, Xb.q, gap_sum, shape_margin[sa],
                         shape_margin[sb], ra_out, rb_out, pos, dist, normal);
        return n;
    } else {
        *unsupported = 1;
        return 0;
    }
    *ra_out = ra; *rb_out = rb;
    int n = 0;
    for (int i = 0; i < 4; ++i) n += dist[i] < MAXVAL;
    return n ? 4 : 0;
}

/*
 * Full collide for all worlds.  pairs: world-major candidate pair list,
 * pair_world[p] = world of pair p, pairs grouped by world in ascending order.
 * Outputs are written compact (flat) in world-major order; world_start has
 * world_count+1 entries.  Returns the number of unsupported pair encounters.
 */
/* Contacts of one world's candidate pairs [p_begin, p_end), appended at out[base ...] (at most
   max_contacts_per_world; writes stop at out_capacity).  Returns the world's contact count. */
static int collide_world(int p_begin, int p_end, const float* body_q, const int* shape_body, const int* shape_type,
                         const float* shape_scale, const float* shape_margin, const float* shape_gap, const int* pairs,
                         int max_contacts_per_world, int base, int out_capacity, int* out_shape0, int* out_shape1,
                         float* out_point0, float* out_point1, float* out_normal, float* out_margin0,
                         float* out_margin1, int* out_pair, const tf* scratch_tf, const float* scratch_aabb,
                         int* unsupported_total) {
    int wcount = 0, claimed = 0;
    for (int p = p_begin; p < p_end; ++p) {
        int s1 = pairs[2 * p], s2 = pairs[2 * p + 1];
        if (s1 == s2 || s1 < 0 || s2 < 0) continue;
        const float* A = scratch_aabb + 6 * s1;
        const float* B = scratch_aabb + 6 * s2;
        if (!aabb_overlap(mk(A[0], A[1], A[2]), mk(A[3], A[4], A[5]), mk(B[0], B[1], B[2]), mk(B[3], B[4], B[5])))
            continue;
        float dist[5];
        v3 pos[5], normal;
        float ra, rb;
        int sa, sb, unsup, convex;
        float gap_sum = shape_gap[s1] + shape_gap[s2];
        int nc = narrow_pair(s1, s2, shape_type, shape_scale, scratch_tf, shape_margin, scratch_aabb, dist, pos,
                             &normal, &ra, &rb, &sa, &sb, gap_sum, &unsup, &convex);
        *unsupported_total += unsup;
        if (nc == 0) continue;
        float ma = shape_margin[sa], mb = shape_margin[sb];
        int valid[5], nv = 0;
        for (int i = 0; i < nc; ++i) {
            if (convex)  /* the per-contact writer drops d > gap (narrow_phase.py:580-747, pipeline.py:80-88) */
                valid[i] = !passes_gap_not(pos[i], normal, dist[i], ra, rb, ma, mb, gap_sum);
            else
                valid[i] = dist[i] < MAXVAL && passes_gap(pos[i], normal, dist[i], ra, rb, ma, mb, gap_sum);
            nv += valid[i];
        }
        if (nv == 0) continue;
        /* primitive pairs: batch append with the reference's overflow rule
           (the claim always advances, a batch is written only if it fits
           entirely); convex pairs: every contact claims a slot and is
           written if that slot is inside the capacity */
        int first = claimed;
        int fits = claimed + nv <= max_contacts_per_world;
        claimed += nv;
        if (!fits && !convex) continue;
        int ba = shape_body[sa], bb = shape_body[sb];
        for (int i = 0, k = 0; i < nc; ++i) {
            if (!valid[i]) continue;
            if (first + k++ >= max_contacts_per_world) break;
            int idx = base + wcount;
            if (idx < out_capacity) {
                v3 xa, xb, n;
                contact_points(pos[i], normal, dist[i], ra, rb, &xa, &xb, &n);
                v3 la = xa, lb = xb;
                if (ba >= 0) la = tf_inv_point(load_tf(body_q + 7 * ba), xa);
                if (bb >= 0) lb = tf_inv_point(load_tf(body_q + 7 * bb), xb);
                out_shape0[idx] = sa;
                out_shape1[idx] = sb;
                out_point0[3 * idx] = la.x; out_point0[3 * idx + 1] = la.y; out_point0[3 * idx + 2] = la.z;
                out_point1[3 * idx] = lb.x; out_point1[3 * idx + 1] = lb.y; out_point1[3 * idx + 2] = lb.z;
                out_normal[3 * idx] = n.x; out_normal[3 * idx + 1] = n.y; out_normal[3 * idx + 2] = n.z;
                out_margin0[idx] = ra + ma;
                out_margin1[idx] = rb + mb;
                out_pair[idx] = p;
            }
            wcount++;
        }
    }
    return wcount;
}

static void shape_poses(int s_begin, int s_end, const float* body_q, const float* shape_transform,
                        const int* shape_body, const int* shape_type, const float* shape_scale,
                        const float* shape_margin, const float* shape_gap, const float* shape_radius, tf* scratch_tf,
                        float* scratch_aabb) {
    for (int s = s_begin; s < s_end; ++s) {
        tf xs = load_tf(shape_transform + 7 * s);
        int b = shape_body[s];
        if (b >= 0) xs = tf_mul(load_tf(body_q + 7 * b), xs);
        scratch_tf[s] = xs;
        v3 lo, hi;
        v3 sc = mk(shape_scale[3 * s], shape_scale[3 * s + 1], shape_scale[3 * s + 2]);
        shape_aabb(shape_type[s], xs.p, xs.q, sc, shape_radius[s], shape_margin[s] + shape_gap[s], &lo, &hi);
        scratch_aabb[6 * s + 0] = lo.x; scratch_aabb[6 * s + 1] = lo.y; scratch_aabb[6 * s + 2] = lo.z;
        scratch_aabb[6 * s + 3] = hi.x; scratch_aabb[6 * s + 4] = hi.y; scratch_aabb[6 * s + 5] = hi.z;
    }
}

int sapx_oracle_collide(int world_count, int shape_count, const float* body_q, const float* shape_transform,
                        const int* shape_body, const int* shape_type, const float* shape_scale,
                        const float* shape_margin, const float* shape_gap, const float* shape_radius,
                        const int* pairs, const int* pair_world, int pair_count, int max_contacts_per_world,
                        int out_capacity, int* out_count, int* world_start, int* out_shape0, int* out_shape1,
                        float* out_point0, float* out_point1, float* out_normal, float* out_margin0,
                        float* out_margin1, int* out_pair, tf* scratch_tf, float* scratch_aabb) {
    int unsupported_total = 0;
    shape_poses(0, shape_count, body_q, shape_transform, shape_body, shape_type, shape_scale, shape_margin, shape_gap,
                shape_radius, scratch_tf, scratch_aabb);
    int total = 0;
    int p = 0;
    for (int w = 0; w < world_count; ++w) {
        world_start[w] = total;
        int pe = p;
        while (pe < pair_count && pair_world[pe] == w) ++pe;
        total += collide_world(p, pe, body_q, shape_body, shape_type, shape_scale, shape_margin, shape_gap, pairs,
                               max_contacts_per_world, total, out_capacity, out_shape0, out_shape1, out_point0,
                               out_point1, out_normal, out_margin0, out_margin1, out_pair, scratch_tf, scratch_aabb,
                               &unsupported_total);
        p = pe;
    }
    world_start[world_count] = total;
    *out_count = total;
    (void)tf_inv_vector;
    return unsupported_total;
}

/*
 * The same contacts with worlds on `threads` OpenMP threads: each world is generated into its own
 * block of max_contacts_per_world staging slots (stage_* arrays, world_count x C), then the blocks
 * are compacted in world order.  Identical output to sapx_oracle_collide (CPU baseline only).
 */
int sapx_oracle_collide_mt(int world_count, int shape_count, const float* body_q, const float* shape_transform,
                           const int* shape_body, const int* shape_type, const float* shape_scale,
                           const float* shape_margin, const float* shape_gap, const float* shape_radius,
                           const int* pairs, const int* pair_world, int pair_count, int max_contacts_per_world,
                           int out_capacity, int* out_count, int* world_start, int* out_shape0, int* out_shape1,
                           float* out_point0, float* out_point1, float* out_normal, float* out_margin0,
                           float* out_margin1, int* out_pair, tf* scratch_tf, float* scratch_aabb,
                           const int* world_pair_start, int* stage_i, float* stage_f, int* world_count_out,
                           int threads) {
    const int C = max_contacts_per_world;
    int unsupported_total = 0;
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int chunk = 0; chunk < threads; ++chunk) {
        int s0 = (int)((long long)shape_count * chunk / threads), s1 = (int)((long long)shape_count * (chunk + 1) / threads);
        shape_poses(s0, s1, body_q, shape_transform, shape_body, shape_type, shape_scale, shape_margin, shape_gap,
                    shape_radius, scratch_tf, scratch_aabb);
    }
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads) reduction(+ : unsupported_total)
    for (int w = 0; w < world_count; ++w) {
        int* si = stage_i + (size_t)w * C * 3;
        float* sf = stage_f + (size_t)w * C * 11;
        int unsup = 0;
        world_count_out[w] = collide_world(world_pair_start[w], world_pair_start[w + 1], body_q, shape_body,
                                           shape_type, shape_scale, shape_margin, shape_gap, pairs, C, 0, C, si,
                                           si + C, sf, sf + 3 * C, sf + 6 * C, sf + 9 * C, sf + 10 * C, si + 2 * C,
                                           scratch_tf, scratch_aabb, &unsup);
        unsupported_total += unsup;
    }
    int total = 0;
    for (int w = 0; w < world_count; ++w) {
        world_start[w] = total;
        const int* si = stage_i + (size_t)w * C * 3;
        const float* sf = stage_f + (size_t)w * C * 11;
        for (int k = 0; k < world_count_out[w]; ++k) {
            int idx = total + k;
            if (idx >= out_capacity) break;
            out_shape0[idx] = si[k];
            out_shape1[idx] = si[C + k];
            out_pair[idx] = si[2 * C + k];
            for (int e = 0; e < 3; ++e) {
                out_point0[3 * idx + e] = sf[3 * k + e];
                out_point1[3 * idx + e] = sf[3 * C + 3 * k + e];
                out_normal[3 * idx + e] = sf[6 * C + 3 * k + e];
            }
            out_margin0[idx] = sf[9 * C + k];
            out_margin1[idx] = sf[10 * C + k];
        }
        total += world_count_out[w];
    }
    world_start[world_count] = total;
    *out_count = total;
    return unsupported_total;
}

/*
 * One convex pair in isolation (box-box or capsule-box, type_a <= type_b):
 * contact centres (5 x 3), signed distances (5), the world normal (3) and the
 * effective radii (2), as the per-contact writer receives them.  Returns the
 * number of contacts.
 */
int sapx_oracle_convex_pair(int type_a, const float* scale_a, const float* xf_a, int type_b, const float* scale_b,
                            const float* xf_b, float gap_sum, float margin_a, float margin_b, float* out_pos,
                            float* out_dist, float* out_normal, float* out_radii) {
    tf a = load_tf(xf_a), b = load_tf(xf_b);
    v3 pos[5], n;
    float dist[5], ra, rb;
    int c = cvx_pair(type_a, mk(scale_a[0], scale_a[1], scale_a[2]), a.p, a.q, type_b,
                     mk(scale_b[0], scale_b[1], scale_b[2]), b.p, b.q, gap_sum, margin_a, margin_b, &ra, &rb, pos,
                     dist, &n);
    for (int i = 0; i < c; ++i) {
        out_pos[3 * i] = pos[i].x; out_pos[3 * i + 1] = pos[i].y; out_pos[3 * i + 2] = pos[i].z;
        out_dist[i] = dist[i];
    }
    out_normal[0] = n.x; out_normal[1] = n.y; out_normal[2] = n.z;
    out_radii[0] = ra; out_radii[1] = rb;
    return c;
}
