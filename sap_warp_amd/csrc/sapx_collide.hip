// Rigid contact generation on gfx950: shape AABBs, explicit-pair AABB test and
// primitive narrowphase, one workgroup per world, deterministic output order.
//
// Compiled with -ffp-contract=off and IEEE sqrt/div so contact geometry is
// bit-identical to the float32 CPU restatement (same operation order).
//
// Reference: sim/collision/pipeline.py:51-179 (writer, AABBs),
// broad_phase_common.py:18-35 / broad_phase_nxn.py:21-59 (pair test),
// narrow_phase.py:141-577 (primitive kernel), primitive.py:37-657, 1150-1205,
// contact_data.py:36-56 (gap check).
//
// Output: contacts of world w are appended in candidate-pair order; a pair's
// contacts are written only if the whole batch fits into the world's capacity
// (the reference's batch-overflow rule, narrow_phase.py:534-539).  The world's
// contacts are staged in LDS and copied out coalesced into the world's block of
// a W x C staging area with its count; a second kernel places every world's block
// at its offset in the flat SapContacts arrays (sum of the earlier counts) and
// writes the per-world offsets.
#include "sapx_device.h"
#include "sapx_params.h"

#pragma clang fp contract(off)

#include "sapx_convex.h"

#include <mutex>

namespace sapx {

constexpr float kMaxVal = 1e10f;
constexpr float kMinVal = 1e-15f;



__device__ __forceinline__ V3 nsafe(V3 a) {
    float l = length(a);
    if (l > 0.0f) return v3(a.x / l, a.y / l, a.z / l);
    return v3(0.0f, 0.0f, 0.0f);
}
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }
__device__ __forceinline__ void rot_cols(Q4 q, V3 c[3]) {
    c[0] = qrot(q, v3(1.0f, 0.0f, 0.0f));
    c[1] = qrot(q, v3(0.0f, 1.0f, 0.0f));
    c[2] = qrot(q, v3(0.0f, 0.0f, 1.0f));
}
__device__ __forceinline__ V3 rmul(const V3 c[3], V3 v) { return (c[0] * v.x + c[1] * v.y) + c[2] * v.z; }
__device__ __forceinline__ V3 rtmul(const V3 c[3], V3 v) { return v3(dot(c[0], v), dot(c[1], v), dot(c[2], v)); }

__device__ void shape_aabb(int type, V3 pos, Q4 q, V3 scale, float radius, float eff, V3* lo, V3* hi) {
    V3 h;
    if (type == kGeoPlane && scale.x == 0.0f && scale.y == 0.0f) {
        h = v3(radius, radius, radius);
    } else if (type == kGeoSphere) {
        h = v3(scale.x, scale.x, scale.x);
    } else if (type == kGeoCapsule) {
        V3 a = qrot(q, v3(0.0f, 0.0f, 1.0f));
        h = v3(fabsf(a.x) * scale.y + scale.x, fabsf(a.y) * scale.y + scale.x, fabsf(a.z) * scale.y + scale.x);
    } else if (type == kGeoBox) {
        V3 c[3];
        rot_cols(q, c);
        h = v3(fabsf(c[0].x) * scale.x + fabsf(c[1].x) * scale.y + fabsf(c[2].x) * scale.z,
               fabsf(c[0].y) * scale.x + fabsf(c[1].y) * scale.y + fabsf(c[2].y) * scale.z,
               fabsf(c[0].z) * scale.x + fabsf(c[1].z) * scale.y + fabsf(c[2].z) * scale.z);
    } else if (type == kGeoEllipsoid) {
        // extent along world axis i: the support value sqrt(sum_k (R_ik s_k)^2)
        V3 c[3];
        rot_cols(q, c);
        const V3 a = c[0] * scale.x, b = c[1] * scale.y, d = c[2] * scale.z;
        h = v3(sqrtf(a.x * a.x + b.x * b.x + d.x * d.x), sqrtf(a.y * a.y + b.y * b.y + d.y * d.y),
               sqrtf(a.z * a.z + b.z * b.z + d.z * d.z));
    } else if (type == kGeoCylinder) {
        // extent along world axis i: r sqrt(1 - a_i^2) + h |a_i| for the unit axis a
        const V3 a = qrot(q, v3(0.0f, 0.0f, 1.0f));
        h = v3(scale.x * sqrtf(fmaxf(1.0f - a.x * a.x, 0.0f)) + scale.y * fabsf(a.x),
               scale.x * sqrtf(fmaxf(1.0f - a.y * a.y, 0.0f)) + scale.y * fabsf(a.y),
               scale.x * sqrtf(fmaxf(1.0f - a.z * a.z, 0.0f)) + scale.y * fabsf(a.z));
    } else if (type == kGeoCone) {
        // apex at +h along the axis, base rim of radius r at -h: along world axis i the extremes
        // are +-h a_i (apex / base centre) and -h a_i +- r sqrt(1 - a_i^2) (rim)
        const V3 a = qrot(q, v3(0.0f, 0.0f, 1.0f));
        const float av[3] = {a.x, a.y, a.z};
        float lo3[3], hi3[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float rim = scale.x * sqrtf(fmaxf(1.0f - av[i] * av[i], 0.0f));
            const float ap = scale.y * av[i], bc = -scale.y * av[i];
            hi3[i] = fmaxf(ap, bc + rim);
            lo3[i] = fminf(ap, bc - rim);
        }
        *lo = (pos + v3(lo3[0], lo3[1], lo3[2])) - v3(eff, eff, eff);
        *hi = (pos + v3(hi3[0], hi3[1], hi3[2])) + v3(eff, eff, eff);
        return;
    } else {
        h = v3(radius, radius, radius);
    }
    *lo = (pos - h) - v3(eff, eff, eff);
    *hi = (pos + h) + v3(eff, eff, eff);
}

__device__ __forceinline__ V3 closest_segment_point(V3 a, V3 b, V3 pt) {
    V3 ab = b - a;
    float t = dot(pt - a, ab) / (dot(ab, ab) + 1e-6f);
    return a + ab * clampf(t, 0.0f, 1.0f);
}
__device__ __forceinline__ void plane_sphere(V3 n, V3 pp, V3 sp, float r, float* dist, V3* pos) {
    float d = dot(sp - pp, n) - r;
    *dist = d;
    *pos = sp - n * (r + 0.5f * d);
}
__device__ __forceinline__ void sphere_sphere(V3 p1, float r1, V3 p2, float r2, float* dist, V3* pos, V3* nout) {
    V3 dir = p2 - p1;
    float d = length(dir);
    V3 nn;
    if (d == 0.0f) nn = v3(1.0f, 0.0f, 0.0f);
    else nn = v3(dir.x / d, dir.y / d, dir.z / d);
    d = d - (r1 + r2);
    *pos = p1 + nn * (r1 + 0.5f * d);
    *dist = d;
    *nout = nn;
}
__device__ void capsule_capsule(V3 p1, V3 a1, float r1, float h1, V3 p2, V3 a2, float r2, float h2, float dist[4],
                                V3 pos[4], V3* normal) {
    V3 axis1 = a1 * h1;
    V3 axis2 = a2 * h2;
    V3 dif = p1 - p2;
    float ma = dot(axis1, axis1);
    float mb = -dot(axis1, axis2);
    float mc = dot(axis2, axis2);
    float u = -dot(axis1, dif);
    float v = dot(axis2, dif);
    float det = ma * mc - mb * mb;
    if (fabsf(det) >= kMinVal) {
        float inv = 1.0f / det;
        float x1 = (mc * u - mb * v) * inv;
        float x2 = (ma * v - mb * u) * inv;
        if (x1 > 1.0f) { x1 = 1.0f; x2 = (v - mb) / mc; }
        else if (x1 < -1.0f) { x1 = -1.0f; x2 = (v + mb) / mc; }
        if (x2 > 1.0f) { x2 = 1.0f; x1 = clampf((u - mb) / ma, -1.0f, 1.0f); }
        else if (x2 < -1.0f) { x2 = -1.0f; x1 = clampf((u + mb) / ma, -1.0f, 1.0f); }
        V3 v1 = p1 + axis1 * x1;
        V3 v2 = p2 + axis2 * x2;
        sphere_sphere(v1, r1, v2, r2, &dist[0], &pos[0], normal);
    } else {
        V3 v1 = p1 + axis1;
        float x2 = clampf((v - mb) / mc, -1.0f, 1.0f);
        V3 v2 = p2 + axis2 * x2;
        sphere_sphere(v1, r1, v2, r2, &dist[0], &pos[0], normal);
        V3 n2;
        v1 = p1 - axis1;
        x2 = clampf((v + mb) / mc, -1.0f, 1.0f);
        v2 = p2 + axis2 * x2;
        sphere_sphere(v1, r1, v2, r2, &dist[1], &pos[1], &n2);
    }
}
// The contact slots of plane-box / plane-cylinder are separate scalars selected by compares: an
// array (or a struct of adjacent floats) read or written at a runtime slot index becomes an
// indexed access to scratch memory.
__device__ __forceinline__ float sel4(int k, float a, float b, float c, float d) {
    return k == 0 ? a : (k == 1 ? b : (k == 2 ? c : d));
}
__device__ __forceinline__ void put4(int k, float v, V3 p, float& d0, float& d1, float& d2, float& d3, V3& p0, V3& p1,
                                     V3& p2, V3& p3) {
    d0 = k == 0 ? v : d0; d1 = k == 1 ? v : d1; d2 = k == 2 ? v : d2; d3 = k == 3 ? v : d3;
    p0 = vsel(k == 0, p, p0); p1 = vsel(k == 1, p, p1); p2 = vsel(k == 2, p, p2); p3 = vsel(k == 3, p, p3);
}
template <int K>
__device__ void plane_box(V3 n, V3 pp, V3 bp, Q4 bq, V3 size, float margin, float (&dist)[K], V3 (&pos)[K]) {
    V3 c[3];
    rot_cols(bq, c);
    float center_dist = dot(bp - pp, n);
    float d0 = dist[0], d1 = dist[1], d2 = dist[2], d3 = dist[3];
    V3 p0 = pos[0], p1 = pos[1], p2 = pos[2], p3 = pos[3];
    int ncontact = 0, worst = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        V3 corner = v3((i & 1) ? size.x : -size.x, (i & 2) ? size.y : -size.y, (i & 4) ? size.z : -size.z);
        corner = rmul(c, corner);
        float ldist = dot(n, corner);
        float cdist = center_dist + ldist;
        if (cdist > margin) continue;
        V3 cpos = (corner + bp) - n * (0.5f * cdist);
        if (ncontact < 4) {
            put4(ncontact, cdist, cpos, d0, d1, d2, d3, p0, p1, p2, p3);
            if (ncontact == 0 || cdist > sel4(worst, d0, d1, d2, d3)) worst = ncontact;
            ncontact++;
        } else if (cdist < sel4(worst, d0, d1, d2, d3)) {
            put4(worst, cdist, cpos, d0, d1, d2, d3, p0, p1, p2, p3);
            worst = 0;
            if (d1 > sel4(worst, d0, d1, d2, d3)) worst = 1;
            if (d2 > sel4(worst, d0, d1, d2, d3)) worst = 2;
            if (d3 > sel4(worst, d0, d1, d2, d3)) worst = 3;
        }
    }
    dist[0] = d0; dist[1] = d1; dist[2] = d2; dist[3] = d3;
    pos[0] = p0; pos[1] = p1; pos[2] = p2; pos[3] = p3;
}
// plane-ellipsoid (primitive.py:322-351): the ellipsoid's support point against the plane normal
__device__ void plane_ellipsoid(V3 n, V3 pp, V3 ep, Q4 eq, V3 size, float* dist, V3* pos) {
    V3 c[3];
    rot_cols(eq, c);
    const V3 ln = rtmul(c, n);
    V3 sup = nsafe(v3(ln.x * size.x, ln.y * size.y, ln.z * size.z));
    sup = v3(-sup.x, -sup.y, -sup.z);
    const V3 p = ep + rmul(c, v3(sup.x * size.x, sup.y * size.y, sup.z * size.z));
    const float d = dot(n, p - pp);
    *dist = d;
    *pos = p - n * (d * 0.5f);
}

// sphere-cylinder (primitive.py:431-501): side, cap or rim, whichever the sphere centre is nearest
__device__ void sphere_cylinder(V3 sp, float sr, V3 cp, V3 axis, float cr, float hh, float* dist, V3* pos,
                                V3* normal) {
    const V3 vec = sp - cp;
    const float x = dot(vec, axis);
    const V3 a_proj = axis * x;
    V3 p_proj = vec - a_proj;
    const float p_sqr = dot(p_proj, p_proj);
    bool side = fabsf(x) < hh;
    bool cap = p_sqr < cr * cr;
    if (side && cap) {
        const float d_cap = hh - fabsf(x);
        const float d_rad = cr - sqrtf(p_sqr);
        if (d_cap < d_rad) side = false;
        else cap = false;
    }
    if (side) {
        sphere_sphere(sp, sr, cp + a_proj, cr, dist, pos, normal);
    } else if (cap) {
        V3 cpos, pn;
        if (x > 0.0f) { cpos = cp + axis * hh; pn = axis; }
        else { cpos = cp - axis * hh; pn = v3(-axis.x, -axis.y, -axis.z); }
        plane_sphere(pn, cpos, sp, sr, dist, pos);
        *normal = v3(-pn.x, -pn.y, -pn.z);
    } else {
        const float l = sqrtf(p_sqr);
        const float inv = 1.0f / (l != 0.0f ? l : 1e-15f);
        p_proj = p_proj * (cr * inv);
        const V3 cap_off = axis * ((x < 0.0f ? -1.0f : 1.0f) * hh);
        sphere_sphere(sp, sr, (cp + cap_off) + p_proj, 0.0f, dist, pos, normal);
    }
}

// plane-cylinder (primitive.py:504-657): the deepest rim point first, then a fixed tripod on the
// near cap (axis within 22.5 degrees of the normal) or the two side generators and the cap-rim
// point facing the plane, each kept when farther than 1 % of the larger dimension from the first
template <int K>
__device__ void plane_cylinder(V3 n, V3 pp, V3 cc, V3 axis, float r, float hh, float (&dist)[K], V3 (&pos)[K]) {
    constexpr float kFlatCos = 0.9238795325112867f;  // cos(22.5 deg)
    float dna = dot(n, axis);
    if (dna > 0.0f) { axis = v3(-axis.x, -axis.y, -axis.z); dna = -dna; }
    const V3 cap = cc + axis * hh;
    V3 pa = v3(-n.x, -n.y, -n.z) + axis * dna;
    const float pa_sq = dot(pa, pa);
    const bool has_align = pa_sq > 1e-10f;
    if (has_align) pa = pa * (1.0f / sqrtf(pa_sq));
    V3 ref = v3(1.0f, 0.0f, 0.0f);
    if (fabsf(dot(axis, ref)) > 0.9f) ref = v3(0.0f, 1.0f, 0.0f);
    const V3 pf = nsafe(ref - axis * dot(axis, ref));
    const bool flat = -dna >= kFlatCos;
    const V3 dperp = vsel(has_align, pa, pf);
    const V3 dpt = cap + dperp * r;
    const float dd = dot(dpt - pp, n);
    const V3 dpos = dpt - n * (dd * 0.5f);
    dist[0] = dd;
    pos[0] = dpos;
    int nc = 1;
    const float thr = 0.01f * fmaxf(r, hh);
    const float thr_sq = thr * thr;
    V3 pts[3];
    int np;
    if (flat) {
        const V3 u = pf * r;
        const V3 v = cross(axis, pf) * r;
        const float c120 = -0.5f, s120 = 0.8660254f;
        pts[0] = cap + u;
        pts[1] = (cap + u * c120) + v * s120;
        pts[2] = (cap + u * c120) - v * s120;
        np = 3;
    } else {
        const V3 pr = vsel(has_align, pa, pf);
        const V3 u = pr * r;
        const V3 v = cross(axis, pr) * r;
        pts[0] = (cc + axis * hh) + u;
        pts[1] = (cc - axis * hh) + u;
        const V3 ppv = cap + v, pnv = cap - v;
        const float dpv = dot(ppv - pp, n), dnv = dot(pnv - pp, n);
        pts[2] = vsel(dpv <= dnv, ppv, pnv);
        np = 3;
    }
    float d0 = dist[0], d1 = dist[1], d2 = dist[2], d3 = dist[3];
    V3 p0 = pos[0], p1 = pos[1], p2 = pos[2], p3 = pos[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= np) break;
        const float d = dot(pts[k] - pp, n);
        const V3 p = pts[k] - n * (d * 0.5f);
        const V3 e = p - dpos;
        if (nc < 4 && dot(e, e) > thr_sq) {
            put4(nc, d, p, d0, d1, d2, d3, p0, p1, p2, p3);
            ++nc;
        }
    }
    dist[1] = d1; dist[2] = d2; dist[3] = d3;
    pos[1] = p1; pos[2] = p2; pos[3] = p3;
}

__device__ __forceinline__ float vget(V3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
__device__ void sphere_box(V3 sp, float r, V3 bp, Q4 bq, V3 size, float* dist, V3* pos, V3* normal) {
    V3 c[3];
    rot_cols(bq, c);
    V3 center = rtmul(c, sp - bp);
    V3 clamped = v3(fmaxf(-size.x, fminf(size.x, center.x)), fmaxf(-size.y, fminf(size.y, center.y)),
                    fmaxf(-size.z, fminf(size.z, center.z)));
    V3 diff = clamped - center;
    float d = length(diff);
    V3 cdir = diff;
    if (d != 0.0f) cdir = v3(diff.x / d, diff.y / d, diff.z / d);
    V3 lp;
    if (d <= 1e-6f) {
        float closest = 2.0f * (size.x + size.y + size.z);
        int k = 0;
        for (int i = 0; i < 6; ++i) {
            float sgn = (i % 2) ? 1.0f : -1.0f;
            float fd = fabsf(sgn * vget(size, i / 2) - vget(center, i / 2));
            if (closest > fd) { closest = fd; k = i; }
        }
        V3 nearest = v3(0.0f, 0.0f, 0.0f);
        float sv = (k % 2) ? -1.0f : 1.0f;
        if (k / 2 == 0) nearest.x = sv; else if (k / 2 == 1) nearest.y = sv; else nearest.z = sv;
        lp = center + nearest * ((r - closest) / 2.0f);
        *normal = rmul(c, nearest);
        *dist = -closest - r;
    } else {
        V3 deepest = center + cdir * r;
        lp = (clamped + deepest) * 0.5f;
        *normal = rmul(c, cdir);
        *dist = d - r;
    }
    *pos = bp + rmul(c, lp);
}
__device__ __forceinline__ void contact_points(V3 center, V3 n_in, float d, float ra, float rb, V3* xa, V3* xb, V3* n) {
    *n = nsafe(n_in);
    *xa = center - (*n) * (0.5f * d + ra);
    *xb = center + (*n) * (0.5f * d + rb);
}

// LDS layout of the collide kernel: transforms and AABBs (14 floats per shape), NT/64 + 4 ints,
// then the contact staging (13 words per slot), then the convex polygon buffer
__host__ __device__ inline size_t collide_stage_offset(int ns, int nt) {
    return ((size_t)ns * 14 * 4 + (size_t)(nt / 64 + 4) * 4 + 15) / 16 * 16;
}
__host__ __device__ inline size_t collide_poly_offset(int ns, int nt, int C) {
    return (collide_stage_offset(ns, nt) + (size_t)C * 13 * 4 + 15) / 16 * 16;
}

// CVX: the model has candidate pairs on the MPR / GJK path (up to five contacts per pair, a
// 16-slot polygon buffer per lane in LDS)
// Narrowphase of one candidate pair whose AABBs overlap (g1 != g2): orders the pair by shape
// type, runs the primitive routine or the convex path, and marks which of its contacts pass the
// gap test (primitive contacts: d <= gap; a convex pair's: not d > gap).  X / AB hold the
// shapes' world transforms (8 floats) and AABBs (6 floats) at the local indices l1 / l2; S holds
// the per-shape tables at the global indices g1 / g2.
struct ShapeTable {
    const int* type;
    const float* scale;
    const float* margin;
    const float* gap;
};
template <bool CVX, int NC>
__device__ __forceinline__ void narrow_pair(const ShapeTable& S, const float* X, const float* AB, int l1, int l2, int g1,
                                            int g2, const cvx::Poly& buf, int* unsupported, int& la, int& lb, int& ga,
                                            int& gb, float& ra, float& rb, float (&dist)[NC], V3 (&pos)[NC], V3& normal,
                                            int& ncvx, bool (&valid)[NC], int& nv) {
    int ta = S.type[g1], tb = S.type[g2];
    la = l1; lb = l2; ga = g1; gb = g2;
    if (ta > tb) {
        int t = la; la = lb; lb = t;
        t = ga; ga = gb; gb = t;
        t = ta; ta = tb; tb = t;
    }
    const float* sca = S.scale + (size_t)ga * 3;
    const float* scb = S.scale + (size_t)gb * 3;
    V3 scale_a = v3(sca[0], sca[1], sca[2]);
    V3 scale_b = v3(scb[0], scb[1], scb[2]);
    TF Xa = load_tf(X + la * 8), Xb = load_tf(X + lb * 8);
    if (ta == kGeoSphere || ta == kGeoCapsule) ra = scale_a.x;
    if (tb == kGeoSphere || tb == kGeoCapsule) rb = scale_b.x;
    const float gap_sum = S.gap[g1] + S.gap[g2];
    const V3 ez = v3(0.0f, 0.0f, 1.0f);
    if (ta == kGeoPlane && tb == kGeoSphere) {
        V3 nn = qrot(Xa.q, ez);
        plane_sphere(nn, Xa.p, Xb.p, scale_b.x, &dist[0], &pos[0]);
        normal = nn;
    } else if (ta == kGeoPlane && tb == kGeoBox) {
        V3 nn = qrot(Xa.q, ez);
        plane_box(nn, Xa.p, Xb.p, Xb.q, scale_b, gap_sum, dist, pos);
        normal = nn;
    } else if (ta == kGeoSphere && tb == kGeoSphere) {
        sphere_sphere(Xa.p, scale_a.x, Xb.p, scale_b.x, &dist[0], &pos[0], &normal);
    } else if (ta == kGeoPlane && tb == kGeoCapsule) {
        V3 nn = qrot(Xa.q, ez);
        V3 ax = qrot(Xb.q, ez);
        V3 seg = ax * scale_b.y;
        plane_sphere(nn, Xa.p, Xb.p + seg, scale_b.x, &dist[0], &pos[0]);
        plane_sphere(nn, Xa.p, Xb.p - seg, scale_b.x, &dist[1], &pos[1]);
        normal = nn;
    } else if (ta == kGeoSphere && tb == kGeoCapsule) {
        V3 ax = qrot(Xb.q, ez);
        V3 seg = ax * scale_b.y;
        V3 pt = closest_segment_point(Xb.p - seg, Xb.p + seg, Xa.p);
        sphere_sphere(Xa.p, scale_a.x, pt, scale_b.x, &dist[0], &pos[0], &normal);
    } else if (ta == kGeoCapsule && tb == kGeoCapsule) {
        V3 a1 = qrot(Xa.q, ez), a2 = qrot(Xb.q, ez);
        capsule_capsule(Xa.p, a1, scale_a.x, scale_a.y, Xb.p, a2, scale_b.x, scale_b.y, dist, pos, &normal);
    } else if (ta == kGeoSphere && tb == kGeoBox) {
        sphere_box(Xa.p, scale_a.x, Xb.p, Xb.q, scale_b, &dist[0], &pos[0], &normal);
    } else if (ta == kGeoPlane && tb == kGeoEllipsoid) {
        V3 nn = qrot(Xa.q, ez);
        plane_ellipsoid(nn, Xa.p, Xb.p, Xb.q, scale_b, &dist[0], &pos[0]);
        normal = nn;
    } else if (ta == kGeoPlane && tb == kGeoCylinder) {
        V3 nn = qrot(Xa.q, ez);
        plane_cylinder(nn, Xa.p, Xb.p, qrot(Xb.q, ez), scale_b.x, scale_b.y, dist, pos);
        normal = nn;
    } else if (ta == kGeoSphere && tb == kGeoCylinder) {
        sphere_cylinder(Xa.p, scale_a.x, Xb.p, qrot(Xb.q, ez), scale_b.x, scale_b.y, &dist[0], &pos[0],
                        &normal);
    } else if (CVX && pair_convex(ta, tb)) {
        int tta = ta;
        V3 ssa = scale_a, ppa = Xa.p;
        bool reach = true;
        if (ta == kGeoPlane) {
            // the other shape's bounding sphere (centre and half diagonal of its collision
            // AABB) must reach the infinite plane's half-space; the plane becomes a box of
            // half extent 10 (R + gap) whose top face lies on the plane under the shape
            const float* Bb = AB + lb * 6;
            const V3 blo = v3(Bb[0], Bb[1], Bb[2]), bhi = v3(Bb[3], Bb[4], Bb[5]);
            const V3 c = (blo + bhi) * 0.5f;
            const float R = length((bhi - blo) * 0.5f);
            const V3 nn = qrot(Xa.q, ez);
            if (scale_a.x != 0.0f || scale_a.y != 0.0f) {
                atomicAdd(unsupported, 1);
                reach = false;
            } else if (dot(c - Xa.p, nn) > R) {
                reach = false;
            } else {
                const float ext = (R + gap_sum) * 10.0f;
                const V3 surf = Xb.p - nn * dot(Xb.p - Xa.p, nn);
                ppa = surf - nn * ext;
                tta = kGeoBox;
                ssa = v3(ext, ext, ext);
            }
        }
        ncvx = 0;
        if (reach) {
            ncvx = cvx::pair(tta, ssa, ppa, Xa.q, tb, scale_b, Xb.p, Xb.q, gap_sum, S.margin[ga],
                             S.margin[gb], buf, &ra, &rb, pos, dist, &normal);
        }
    } else {
        atomicAdd(unsupported, 1);
    }
    const float ma = S.margin[ga], mbm = S.margin[gb];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        // primitive contacts pass the gap check (d <= gap); a convex pair's contacts go
        // through the per-contact writer, which drops d > gap
        if (ncvx >= 0 ? i < ncvx : dist[i] < kMaxVal) {
            V3 xa, xb, nn;
            contact_points(pos[i], normal, dist[i], ra, rb, &xa, &xb, &nn);
            float total = ((ra + rb) + ma) + mbm;
            float dd = dot(xb - xa, nn) - total;
            valid[i] = ncvx >= 0 ? !(dd > gap_sum) : dd <= gap_sum;
            nv += valid[i] ? 1 : 0;
        }
    }
}

template <int NT, bool CVX>
__global__ __launch_bounds__(NT) void sapx_collide_kernel(CollideParams P) {
    extern __shared__ __attribute__((aligned(16))) char smem_raw[];
    constexpr int NC = CVX ? 5 : 4;
    const DevModel& M = P.m;
    const int tid = threadIdx.x;
    const int nglob = M.total_shapes - M.global_shape_base;
    const int ns = M.S + nglob;
    float* X = reinterpret_cast<float*>(smem_raw);  // ns x 8 world transforms
    float* AB = X + ns * 8;                          // ns x 6 aabbs
    int* scan = reinterpret_cast<int*>(AB + ns * 6); // NT/64 wave totals, staged count
    // the world's contacts, staged: shape0, shape1, margin0, margin1 (C each), point0, point1, normal (3C each)
    int* s_sh0 = reinterpret_cast<int*>(smem_raw + collide_stage_offset(ns, NT));
    int* s_sh1 = s_sh0 + P.C;
    float* s_m0 = reinterpret_cast<float*>(s_sh1 + P.C);
    float* s_m1 = s_m0 + P.C;
    float* s_p0 = s_m1 + P.C;
    float* s_p1 = s_p0 + 3 * P.C;
    float* s_n = s_p1 + 3 * P.C;
    float2* poly = reinterpret_cast<float2*>(smem_raw + collide_poly_offset(ns, NT, P.C));
    const int w = blockIdx.x;
    for (int s = tid; s < ns; s += NT) {
        const int g = (s < M.S) ? w * M.S + s : M.global_shape_base + (s - M.S);
        TF xs = load_tf(M.shape_transform + (size_t)g * 7);
        const int b = M.shape_body[g];
        if (b >= 0) xs = tf_mul(load_tf(P.body_q + (size_t)b * 7), xs);
        float* o = X + s * 8;
        o[0] = xs.p.x; o[1] = xs.p.y; o[2] = xs.p.z; o[3] = xs.q.x; o[4] = xs.q.y; o[5] = xs.q.z; o[6] = xs.q.w;
        V3 lo, hi;
        const float* sc = M.shape_scale + (size_t)g * 3;
        shape_aabb(M.shape_type[g], xs.p, xs.q, v3(sc[0], sc[1], sc[2]), M.shape_radius[g],
                   M.shape_margin[g] + M.shape_gap[g], &lo, &hi);
        float* a = AB + s * 6;
        a[0] = lo.x; a[1] = lo.y; a[2] = lo.z; a[3] = hi.x; a[4] = hi.y; a[5] = hi.z;
    }
    if (tid == 0) scan[NT / 64] = 0;  // running count of staged contacts
    __syncthreads();
    int claim_base = 0;  // uniform running claim of earlier chunks
    const int lane = tid & 63, wave = tid >> 6;
    for (int p0 = 0; p0 < M.P; p0 += NT) {
        const int p = p0 + tid;
        int nv = 0;
        float dist[NC];
        V3 pos[NC];
        bool valid[NC];
#pragma unroll
        for (int i = 0; i < NC; ++i) { dist[i] = kMaxVal; pos[i] = v3(0, 0, 0); valid[i] = false; }
        V3 normal = v3(0, 0, 0);
        float ra = 0.0f, rb = 0.0f;
        int la = -1, lb = -1, ga = -1, gb = -1;
        int ncvx = -1;  // contact count of a convex pair; -1 for primitive pairs
        if (p < M.P) {
            int e0 = M.t_pairs[p * 2], e1 = M.t_pairs[p * 2 + 1];
            int l1 = e0 >= 0 ? e0 : M.S + (-e0 - 1 - M.global_shape_base);
            int l2 = e1 >= 0 ? e1 : M.S + (-e1 - 1 - M.global_shape_base);
            int g1 = e0 >= 0 ? w * M.S + e0 : -e0 - 1;
            int g2 = e1 >= 0 ? w * M.S + e1 : -e1 - 1;
            const float* A = AB + l1 * 6;
            const float* B = AB + l2 * 6;
            bool ov = A[0] <= B[3] && A[3] >= B[0] && A[1] <= B[4] && A[4] >= B[1] && A[2] <= B[5] && A[5] >= B[2];
            if (ov && g1 != g2) {
                narrow_pair<CVX, NC>(ShapeTable{M.shape_type, M.shape_scale, M.shape_margin, M.shape_gap}, X, AB, l1, l2,
                                     g1, g2, cvx::Poly{poly + tid, NT}, P.unsupported, la, lb, ga, gb, ra, rb, dist, pos,
                                     normal, ncvx, valid, nv);
            }
        }
        // block exclusive scan of nv in pair order
        int incl = nv;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            int y = __shfl_up(incl, off, 64);
            if (lane >= off) incl += y;
        }
        if (lane == 63) scan[wave] = incl;
        __syncthreads();
        int wave_off = 0, chunk_total = 0;
        for (int wv = 0; wv < NT / 64; ++wv) {
            if (wv < wave) wave_off += scan[wv];
            chunk_total += scan[wv];
        }
        const int prefix = claim_base + wave_off + incl - nv;
        // a primitive pair's batch is written only if it fits entirely; a convex pair's
        // contacts each claim a slot and are written while the slot is inside the capacity
        if (nv > 0 && (prefix + nv <= P.C || (ncvx >= 0 && prefix < P.C))) {
            int idx = prefix;
            const int bA = M.shape_body[ga], bB = M.shape_body[gb];
            const float ma = M.shape_margin[ga], mbm = M.shape_margin[gb];
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                if (!valid[i]) continue;
                if (idx >= P.C) break;
                V3 xa, xb, nn;
                contact_points(pos[i], normal, dist[i], ra, rb, &xa, &xb, &nn);
                V3 pa = xa, pb = xb;
                if (bA >= 0) {
                    TF t = load_tf(P.body_q + (size_t)bA * 7);
                    pa = qrot(qconj(t.q), xa - t.p);
                }
                if (bB >= 0) {
                    TF t = load_tf(P.body_q + (size_t)bB * 7);
                    pb = qrot(qconj(t.q), xb - t.p);
                }
                const int o = idx;
                s_sh0[o] = ga;
                s_sh1[o] = gb;
                s_p0[o * 3] = pa.x; s_p0[o * 3 + 1] = pa.y; s_p0[o * 3 + 2] = pa.z;
                s_p1[o * 3] = pb.x; s_p1[o * 3 + 1] = pb.y; s_p1[o * 3 + 2] = pb.z;
                s_n[o * 3] = nn.x; s_n[o * 3 + 1] = nn.y; s_n[o * 3 + 2] = nn.z;
                s_m0[o] = ra + ma;
                s_m1[o] = rb + mbm;
                ++idx;
            }
            atomicMax(&scan[NT / 64], idx);
        }
        claim_base += chunk_total;
        __syncthreads();
    }
    // ---- the world's block of the staging area (shape ids, margins, xyz triples), coalesced
    const int nw = scan[NT / 64];  // contacts staged (at most C)
    const size_t C = (size_t)P.C;
    int* gi = P.st_i + (size_t)w * C * 2;
    float* gf = P.st_f + (size_t)w * C * 11;
    for (int k = tid; k < nw; k += NT) {
        gi[k] = s_sh0[k];
        gi[C + k] = s_sh1[k];
        gf[k] = s_m0[k];
        gf[C + k] = s_m1[k];
    }
    for (int e = tid; e < 3 * nw; e += NT) {
        gf[2 * C + e] = s_p0[e];
        gf[5 * C + e] = s_p1[e];
        gf[8 * C + e] = s_n[e];
    }
    if (tid == 0) P.wcount[w] = nw;
}

// Places every world's staged contacts at its offset in the flat, world-grouped contact arrays
// (the sum of the earlier worlds' counts, cut at the buffer's capacity) and writes world_start /
// count.  One workgroup per kPlaceWorlds consecutive worlds (enough workgroups to spread the copy
// over the CUs); each sums the counts before its range itself, so workgroups never wait on each
// other.
constexpr int kPlaceWorlds = 32;
constexpr int kPlaceThreads = 256;
__global__ __launch_bounds__(kPlaceThreads) void sapx_place_kernel(PlaceParams P) {
    __shared__ int part[kPlaceThreads / 64];
    __shared__ int incl_s[kPlaceWorlds];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int w0 = blockIdx.x * kPlaceWorlds;
    // counts of all earlier worlds
    int acc = 0;
    for (int i = tid; i < w0; i += kPlaceThreads) acc += P.wcount[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) part[wave] = acc;
    __syncthreads();
    int base = 0;
#pragma unroll
    for (int k = 0; k < kPlaceThreads / 64; ++k) base += part[k];
    const int cap = P.capacity;
    // inclusive scan of this range's counts (wave 0)
    if (wave == 0) {
        const int w = w0 + lane;
        const int x = (lane < kPlaceWorlds && w < P.W) ? P.wcount[w] : 0;
        int incl = x;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(incl, off, 64);
            if (lane >= off) incl += y;
        }
        if (lane < kPlaceWorlds) incl_s[lane] = incl;
        if (lane < kPlaceWorlds && w < P.W) P.world_start[w] = min(base + incl - x, cap);
        // only a lane of this range closes the list: lanes kPlaceWorlds..63 carry the next range's
        // world ids with a zero count, and one of them may be the last world
        if (lane < kPlaceWorlds && w == P.W - 1) {
            P.world_start[P.W] = min(base + incl, cap);
            P.count[0] = min(base + incl, cap);
        }
    }
    __syncthreads();
    // copy: contact t of the range belongs to the first world whose inclusive count exceeds t
    const int nrange = min(kPlaceWorlds, P.W - w0);
    const int total = incl_s[nrange - 1];
    const size_t C = (size_t)P.C;
    for (int t = tid; t < total; t += kPlaceThreads) {
        int lo = 0, hi = nrange - 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (incl_s[mid] > t) hi = mid;
            else lo = mid + 1;
        }
        const int k = t - (lo > 0 ? incl_s[lo - 1] : 0);
        const int dst = base + t;
        if (dst >= cap) break;
        const size_t ww = (size_t)(w0 + lo);
        const int* gi = P.st_i + ww * C * 2;
        const float* gf = P.st_f + ww * C * 11;
        P.shape0[dst] = gi[k];
        P.shape1[dst] = gi[C + k];
        P.m0[dst] = gf[k];
        P.m1[dst] = gf[C + k];
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            P.p0[(size_t)dst * 3 + e] = gf[2 * C + 3 * k + e];
            P.p1[(size_t)dst * 3 + e] = gf[5 * C + 3 * k + e];
            P.nrm[(size_t)dst * 3 + e] = gf[8 * C + 3 * k + e];
        }
    }
}

// ---------------------------------------------------------------------------
// Static scratch of sapx_collide_noalloc: the W x C contact staging area and the per-world counts
// in one file-scope device array (each device that loads the library holds its own copy), for
// callers that cannot allocate.  Layout in 4-byte words: [0] unsupported-pair counter,
// [64, 64 + kNoAllocWorlds) counts, then kNoAllocSlots x 2 shape ids, then kNoAllocSlots x 11 floats.
// ---------------------------------------------------------------------------
static __device__ __align__(256) int scratch[kNoAllocWords];

namespace {
std::mutex g_scratch_mu;
int* g_scratch_ptr[kNoAllocMaxDevices] = {};
}  // namespace

hipError_t collide_static_scratch(int** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<int*>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

size_t collide_lds_bytes(const DevModel& m, int C) {
    const int ns = m.S + (m.total_shapes - m.global_shape_base);
    const size_t base = collide_poly_offset(ns, kCollideThreads, C);
    return m.has_convex ? base + (size_t)kCollideThreads * cvx::kBuf * sizeof(float2) : base;
}

hipError_t launch_collide(const CollideParams& P, const PlaceParams& Q, hipStream_t stream) {
    constexpr int NT = kCollideThreads;
    const int nglob = P.m.total_shapes - P.m.global_shape_base;
    const int ns = P.m.S + nglob;
    const size_t base = collide_poly_offset(ns, NT, P.C);
    if (P.m.has_convex) {
        const size_t smem = base + (size_t)NT * cvx::kBuf * sizeof(float2);
        hipLaunchKernelGGL((sapx_collide_kernel<NT, true>), dim3(P.m.W), dim3(NT), smem, stream, P);
    } else {
        hipLaunchKernelGGL((sapx_collide_kernel<NT, false>), dim3(P.m.W), dim3(NT), base, stream, P);
    }
    const unsigned blocks = (unsigned)((Q.W + kPlaceWorlds - 1) / kPlaceWorlds);
    hipLaunchKernelGGL(sapx_place_kernel, dim3(blocks), dim3(kPlaceThreads), 0, stream, Q);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// One large world (a bin of a thousand bodies): no per-world LDS staging.  Shape transforms
// and AABBs go to global scratch (one thread per shape); each candidate pair is one thread
// running the same narrowphase and contact writer as the per-world kernel, staging its
// contacts in its own slots; after an exclusive scan of the per-pair counts (pair order), the
// write kernel applies the per-world capacity rule (a primitive pair's batch only if it fits;
// a convex pair's contacts while their slots are inside) and places the contacts.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void sapx_big_shapes_kernel(BigCollideParams B) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= B.nshapes) {
        if (s == B.nshapes) B.count[0] = 0;
        return;
    }
    TF xs = load_tf(B.shape_transform + (size_t)s * 7);
    const int b = B.shape_body[s];
    if (b >= 0) xs = tf_mul(load_tf(B.body_q + (size_t)b * 7), xs);
    float* o = B.X + (size_t)s * 8;
    o[0] = xs.p.x; o[1] = xs.p.y; o[2] = xs.p.z; o[3] = xs.q.x; o[4] = xs.q.y; o[5] = xs.q.z; o[6] = xs.q.w;
    V3 lo, hi;
    const float* sc = B.shape_scale + (size_t)s * 3;
    shape_aabb(B.shape_type[s], xs.p, xs.q, v3(sc[0], sc[1], sc[2]), B.shape_radius[s], B.shape_margin[s] + B.shape_gap[s],
               &lo, &hi);
    float* a = B.AB + (size_t)s * 6;
    a[0] = lo.x; a[1] = lo.y; a[2] = lo.z; a[3] = hi.x; a[4] = hi.y; a[5] = hi.z;
}

constexpr int kBigPairThreads = 128;

__global__ __launch_bounds__(kBigPairThreads) void sapx_big_pairs_kernel(BigCollideParams B) {
    __shared__ float2 poly[kBigPairThreads * cvx::kBuf];
    constexpr int NC = kBigSlots;
    const int p = blockIdx.x * kBigPairThreads + threadIdx.x;
    if (p >= B.P) return;
    const int g1 = B.pairs[2 * p], g2 = B.pairs[2 * p + 1];
    if (g1 < 0 || g2 < 0 || g1 == g2) {
        B.cnt[p] = 0;
        B.cvx[p] = 0;
        return;
    }
    int nv = 0, ncvx = -1;
    float dist[NC];
    V3 pos[NC];
    bool valid[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) { dist[i] = kMaxVal; pos[i] = v3(0, 0, 0); valid[i] = false; }
    V3 normal = v3(0, 0, 0);
    float ra = 0.0f, rb = 0.0f;
    int la = -1, lb = -1, ga = -1, gb = -1;
    const float* A = B.AB + (size_t)g1 * 6;
    const float* Bb = B.AB + (size_t)g2 * 6;
    const bool ov = A[0] <= Bb[3] && A[3] >= Bb[0] && A[1] <= Bb[4] && A[4] >= Bb[1] && A[2] <= Bb[5] && A[5] >= Bb[2];
    if (ov) {
        narrow_pair<true, NC>(ShapeTable{B.shape_type, B.shape_scale, B.shape_margin, B.shape_gap}, B.X, B.AB, g1, g2, g1,
                              g2, cvx::Poly{poly + threadIdx.x, kBigPairThreads}, B.unsupported, la, lb, ga, gb, ra, rb,
                              dist, pos, normal, ncvx, valid, nv);
    }
    B.cnt[p] = nv;
    B.cvx[p] = ncvx >= 0 ? 1 : 0;
    if (nv == 0) return;
    const int bA = B.shape_body[ga], bB = B.shape_body[gb];
    const float ma = B.shape_margin[ga], mbm = B.shape_margin[gb];
    int k = 0;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        if (!valid[i]) continue;
        V3 xa, xb, nn;
        contact_points(pos[i], normal, dist[i], ra, rb, &xa, &xb, &nn);
        V3 pa = xa, pb = xb;
        if (bA >= 0) {
            TF t = load_tf(B.body_q + (size_t)bA * 7);
            pa = qrot(qconj(t.q), xa - t.p);
        }
        if (bB >= 0) {
            TF t = load_tf(B.body_q + (size_t)bB * 7);
            pb = qrot(qconj(t.q), xb - t.p);
        }
        const size_t slot = (size_t)p * NC + k;
        B.st_sh[2 * slot] = ga;
        B.st_sh[2 * slot + 1] = gb;
        float* f = B.st_f + slot * 11;
        f[0] = pa.x; f[1] = pa.y; f[2] = pa.z;
        f[3] = pb.x; f[4] = pb.y; f[5] = pb.z;
        f[6] = nn.x; f[7] = nn.y; f[8] = nn.z;
        f[9] = ra + ma;
        f[10] = rb + mbm;
        ++k;
    }
}

__global__ __launch_bounds__(256) void sapx_big_write_kernel(BigCollideParams B) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= B.P) return;
    const int nv = B.cnt[p];
    const int prefix = B.off[p];
    if (nv <= 0 || !(prefix + nv <= B.C || (B.cvx[p] && prefix < B.C))) return;
    int idx = prefix;
    for (int k = 0; k < nv && idx < B.C; ++k, ++idx) {
        const size_t slot = (size_t)p * kBigSlots + k;
        const float* f = B.st_f + slot * 11;
        B.out_shape0[idx] = B.st_sh[2 * slot];
        B.out_shape1[idx] = B.st_sh[2 * slot + 1];
        B.out_p0[3 * idx] = f[0]; B.out_p0[3 * idx + 1] = f[1]; B.out_p0[3 * idx + 2] = f[2];
        B.out_p1[3 * idx] = f[3]; B.out_p1[3 * idx + 1] = f[4]; B.out_p1[3 * idx + 2] = f[5];
        B.out_n[3 * idx] = f[6]; B.out_n[3 * idx + 1] = f[7]; B.out_n[3 * idx + 2] = f[8];
        B.out_m0[idx] = f[9];
        B.out_m1[idx] = f[10];
        if (B.out_pair) B.out_pair[idx] = p;
    }
    atomicMax(B.count, idx);
}

__global__ void sapx_big_finish_kernel(BigCollideParams B) {
    if (threadIdx.x == 0) {
        B.world_start[0] = 0;
        B.world_start[1] = B.count[0];
    }
}

hipError_t launch_big_pairs(const BigCollideParams& B, hipStream_t stream) {
    hipLaunchKernelGGL(sapx_big_shapes_kernel, dim3((B.nshapes + 1 + 255) / 256), dim3(256), 0, stream, B);
    if (B.P > 0)
        hipLaunchKernelGGL(sapx_big_pairs_kernel, dim3((B.P + kBigPairThreads - 1) / kBigPairThreads),
                           dim3(kBigPairThreads), 0, stream, B);
    return hipGetLastError();
}

hipError_t launch_big_write(const BigCollideParams& B, hipStream_t stream) {
    if (B.P > 0) hipLaunchKernelGGL(sapx_big_write_kernel, dim3((B.P + 255) / 256), dim3(256), 0, stream, B);
    hipLaunchKernelGGL(sapx_big_finish_kernel, dim3(1), dim3(64), 0, stream, B);
    return hipGetLastError();
}

}  // namespace sapx
