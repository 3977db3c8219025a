/*
 * Contacts of the pairs the primitive routines leave to MPR / GJK (two
 * support-mapped shapes among sphere, capsule, ellipsoid, cylinder, box, cone;
 * an infinite plane against a cone as a box proxy) for the float32 oracle —
 * TEST INFRASTRUCTURE ONLY; included by narrowphase.c after its vector helpers.
 *
 * Restates the reference's GJK/MPR path for pairs the primitive kernel routes on:
 *   routing + per-pair driver   sim/collision/narrow_phase.py:580-747,
 *                               core.py:625-724, core.py:278-392
 *   fused MPR-then-GJK solver   sim/collision/convex.py:26-118
 *   support maps                sim/collision/support.py:94-276 (box, sphere,
 *                               capsule, ellipsoid, cylinder, cone; capsule /
 *                               sphere radius replaced by 1e-4 and restored
 *                               afterwards, core.py:340-351)
 *   Minkowski portal refinement sim/collision/mpr.py:27-237 (30 iterations,
 *                               epsilon 1e-5, inflate 1e-4)
 *   GJK distance                sim/collision/simplex.py:39-436 (30
 *                               iterations, epsilon 1e-4, Johnson sub-simplex
 *                               selection with the 1e-8 degeneracy guard)
 *   contact manifold            sim/collision/multicontact.py:20-955 (five
 *                               support probes tilted 2 degrees about the
 *                               normal, incremental plane trackers, body
 *                               projectors, polygon trim, rotating calipers to
 *                               at most four points, plus the deepest contact
 *                               when the face normals deviate by more than 0.1
 *                               degrees)
 *   tangent basis               sim/collision/math.py:20-44
 *   post-process                sim/collision/core.py:172-275 (Minkowski shift;
 *                               a cylinder / cone rolling on a box or plane:
 *                               contact points projected onto the plane
 *                               through its axis and the normal)
 *   single contact              core.py:360-367 (a sphere or ellipsoid pair)
 *   infinite plane as a box     core.py:507-566 (half extents 10 x the other
 *                               shape's bounding-sphere radius + gap, top face
 *                               on the plane under it), narrow_phase.py:640-720
 *                               (bounding-sphere half-space test)
 *
 * The 2D polygon buffers keep the reference's layout: the shape-B polygon
 * starts at slot 0 and the shape-A polygon at slot 5 of one buffer, so the
 * B polygon growing during trimming overwrites A vertices that the trim has
 * already consumed (multicontact.py:797-799, 464-468).  The buffer has 16
 * slots and every index is clamped into it.
 */
#ifndef SAPX_ORACLE_CONVEX_H
#define SAPX_ORACLE_CONVEX_H

typedef struct { float x, y; } v2;
typedef struct { v3 B, BtoA; } mvert;
typedef struct { int type; v3 scale; } cvx_geom;
typedef struct { v3 ref, prev, n; float area; } cvx_tracker;
typedef struct { float d; v3 n; } cvx_projector;

#define CVX_SIMPLEX_EPS 1e-8f
#define CVX_MC_EPS 1e-5f
#define CVX_BUF 16
#define CVX_SIN_TILT 0.034899496f
#define CVX_COS_TILT 0.99939084f
#define CVX_COS_DEEPEST 0.99999845f

static v3 v3neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static float len_sq(v3 a) { return dot(a, a); }
static q4 qconj4(q4 q) { q.x = -q.x; q.y = -q.y; q.z = -q.z; return q; }
static v3 qrot_inv(q4 q, v3 v) { return qrot(qconj4(q), v); }
static v2 mk2(float x, float y) { v2 r = {x, y}; return r; }
static float len_sq2(v2 a) { return a.x * a.x + a.y * a.y; }
static v2 sub2(v2 a, v2 b) { return mk2(a.x - b.x, a.y - b.y); }

/* ---------------------------------------------------------------- support */
static v3 cvx_support(const cvx_geom* g, v3 d) {
    if (g->type == GEO_BOX) {
        return mk(d.x >= 0.0f ? g->scale.x : -g->scale.x, d.y >= 0.0f ? g->scale.y : -g->scale.y,
                  d.z >= 0.0f ? g->scale.z : -g->scale.z);
    }
    if (g->type == GEO_SPHERE || g->type == GEO_CAPSULE) {
        v3 n = len_sq(d) > 1e-12f ? normalize(d) : mk(1.0f, 0.0f, 0.0f);
        v3 r = scl(n, g->scale.x);
        if (g->type == GEO_CAPSULE) r = add(r, mk(0.0f, 0.0f, d.z >= 0.0f ? g->scale.y : -g->scale.y));
        return r;
    }
    if (g->type == GEO_ELLIPSOID) {
        const float a = g->scale.x, b = g->scale.y, c = g->scale.z;
        if (len_sq(d) > 1e-12f) {
            const float adx = a * d.x, bdy = b * d.y, cdz = c * d.z;
            const float den2 = adx * adx + bdy * bdy + cdz * cdz;
            if (den2 > 1e-12f) {
                const float den = sqrtf(den2);
                return mk((a * a) * d.x / den, (b * b) * d.y / den, (c * c) * d.z / den);
            }
        }
        return mk(a, 0.0f, 0.0f);
    }
    if (g->type == GEO_CYLINDER) {
        const float r = g->scale.x, hh = g->scale.y;
        v3 lat = mk(r, 0.0f, 0.0f);
        v3 dxy = mk(d.x, d.y, 0.0f);
        if (len_sq(dxy) > 1e-12f) {
            v3 nxy = normalize(dxy);
            lat = mk(nxy.x * r, nxy.y * r, 0.0f);
        }
        if (d.z > 0.0f) return mk(lat.x, lat.y, hh);
        if (d.z < 0.0f) return mk(lat.x, lat.y, -hh);
        return lat;
    }
    if (g->type == GEO_CONE) {
        /* apex at +hh, base disk at -hh: the apex when dz >= r / (2 hh) |d_xy| */
        const float r = g->scale.x, hh = g->scale.y;
        const float lxy = length(mk(d.x, d.y, 0.0f));
        const float k = hh > 1e-12f ? r / (2.0f * hh) : 0.0f;
        if (lxy <= 1e-12f) return d.z >= 0.0f ? mk(0.0f, 0.0f, hh) : mk(r, 0.0f, -hh);
        if (d.z >= k * lxy) return mk(0.0f, 0.0f, hh);
        return mk(d.x / lxy * r, d.y / lxy * r, -hh);
    }
    return mk(0.0f, 0.0f, 0.0f);
}
/* shape B lives in A's frame at (qb, pb) */
static v3 cvx_support_b(const cvx_geom* gb, v3 d, q4 qb, v3 pb) {
    return add(qrot(qb, cvx_support(gb, qrot_inv(qb, d))), pb);
}
static mvert cvx_msupport(const cvx_geom* ga, const cvx_geom* gb, v3 d, q4 qb, v3 pb, float extend) {
    mvert v;
    v3 pa = cvx_support(ga, d);
    v.B = cvx_support_b(gb, v3neg(d), qb, pb);
    if (extend != 0.0f) {
        v3 e = scl(scl(normalize(d), extend), 0.5f);
        pa = add(pa, e);
        v.B = sub(v.B, e);
    }
    v.BtoA = sub(pa, v.B);
    return v;
}
static v3 mvert_a(mvert v) { return add(v.B, v.BtoA); }

/* ------------------------------------------------------------------- MPR */
static int cvx_mpr(const cvx_geom* ga, const cvx_geom* gb, q4 qb, v3 pb, float extend, v3* pa_out, v3* pb_out,
                   v3* n_out, float* pen_out) {
    const float ne = 1e-16f;
    const float ce = 1e-5f;
    const int max_iter = 30;
    float pen = 0.0f;
    v3 pa = mk(0.0f, 0.0f, 0.0f), pbb = mk(0.0f, 0.0f, 0.0f), normal;
    mvert v0, v1, v2, v3_, v4;
    int hit = 0, ret = 0;
    v0.B = pb;
    v0.BtoA = v3neg(pb);
    normal = v0.BtoA;
    if (fabsf(normal.x) < ne && fabsf(normal.y) < ne && fabsf(normal.z) < ne) v0.BtoA = mk(1e-5f, 0.0f, 0.0f);
    normal = v3neg(v0.BtoA);
    v1 = cvx_msupport(ga, gb, normal, qb, pb, extend);
    pa = mvert_a(v1);
    pbb = v1.B;
    if (dot(v1.BtoA, normal) <= 0.0f) goto done;
    normal = cross(v1.BtoA, v0.BtoA);
    if (len_sq(normal) < ne * ne) {
        normal = normalize(sub(v1.BtoA, v0.BtoA));
        pen = dot(v1.BtoA, normal);
        ret = 1;
        goto done;
    }
    v2 = cvx_msupport(ga, gb, normal, qb, pb, extend);
    if (dot(v2.BtoA, normal) <= 0.0f) goto done;
    normal = cross(sub(v1.BtoA, v0.BtoA), sub(v2.BtoA, v0.BtoA));
    if (dot(normal, v0.BtoA) > 0.0f) {
        mvert t = v1; v1 = v2; v2 = t;
        normal = v3neg(normal);
    }
    for (int phase1 = 0;;) {
        if (phase1 > max_iter) goto done;
        phase1++;
        v3_ = cvx_msupport(ga, gb, normal, qb, pb, extend);
        if (dot(v3_.BtoA, normal) <= 0.0f) goto done;
        if (dot(cross(v1.BtoA, v3_.BtoA), v0.BtoA) < 0.0f) {
            v2 = v3_;
            normal = cross(sub(v1.BtoA, v0.BtoA), sub(v3_.BtoA, v0.BtoA));
            continue;
        }
        if (dot(cross(v3_.BtoA, v2.BtoA), v0.BtoA) < 0.0f) {
            v1 = v3_;
            normal = cross(sub(v3_.BtoA, v0.BtoA), sub(v2.BtoA, v0.BtoA));
            continue;
        }
        break;
    }
    for (int phase2 = 0;;) {
        phase2++;
        v3 t1 = sub(v2.BtoA, v1.BtoA), t2 = sub(v3_.BtoA, v1.BtoA);
        normal = cross(t1, t2);
        float nsq = len_sq(normal);
        if (nsq < ne * ne) goto done;
        if (!hit) hit = dot(normal, v1.BtoA) >= 0.0f;
        v4 = cvx_msupport(ga, gb, normal, qb, pb, extend);
        float delta = dot(sub(v4.BtoA, v3_.BtoA), normal);
        pen = dot(v4.BtoA, normal);
        if (delta * delta <= ce * ce * nsq || pen <= 0.0f || phase2 > max_iter) {
            if (hit) {
                float inv = 1.0f / sqrtf(nsq);
                pen *= inv;
                normal = scl(normal, inv);
                float gamma = dot(cross(v1.BtoA, t1), normal) * inv;
                float beta = dot(cross(t2, v1.BtoA), normal) * inv;
                float alpha = 1.0f - gamma - beta;
                pa = add(add(scl(mvert_a(v1), alpha), scl(mvert_a(v2), beta)), scl(mvert_a(v3_), gamma));
                pbb = add(add(scl(v1.B, alpha), scl(v2.B, beta)), scl(v3_.B, gamma));
            }
            ret = hit;
            goto done;
        }
        v3 t4 = cross(v4.BtoA, v0.BtoA);
        if (dot(t4, v1.BtoA) >= 0.0f) {
            if (dot(t4, v2.BtoA) >= 0.0f) v1 = v4; else v3_ = v4;
        } else {
            if (dot(t4, v3_.BtoA) >= 0.0f) v2 = v4; else v1 = v4;
        }
    }
done:
    *pa_out = pa; *pb_out = pbb; *n_out = normal; *pen_out = pen;
    return ret;
}

/* ------------------------------------------------------------------- GJK */
/* simplex slot i holds B in sv[2i] and the Minkowski point BtoA in sv[2i+1] */
static v3 cvx_segment(const v3 sv[8], int i0, int i1, float bc[4], unsigned* mask) {
    v3 a = sv[2 * i0 + 1], b = sv[2 * i1 + 1];
    v3 e = sub(b, a);
    float vsq = len_sq(e);
    int deg = vsq < CVX_SIMPLEX_EPS;
    float t = -dot(a, e) / (deg ? CVX_SIMPLEX_EPS : vsq);
    float l0 = 1.0f - t, l1 = t;
    unsigned m = (1u << i0) | (1u << i1);
    if (l0 < 0.0f || deg) { m = 1u << i1; l0 = 0.0f; l1 = 1.0f; }
    else if (l1 < 0.0f) { m = 1u << i0; l0 = 1.0f; l1 = 0.0f; }
    bc[0] = bc[1] = bc[2] = bc[3] = 0.0f;
    bc[i0] = l0;
    bc[i1] = l1;
    *mask = m;
    return add(scl(a, l0), scl(b, l1));
}
static v3 cvx_triangle(const v3 sv[8], int i0, int i1, int i2, float bc[4], unsigned* mask) {
    v3 a = sv[2 * i0 + 1], b = sv[2 * i1 + 1], c = sv[2 * i2 + 1];
    v3 u = sub(a, b), w = sub(a, c);
    v3 n = cross(u, w);
    float t = len_sq(n);
    int deg = t < CVX_SIMPLEX_EPS;
    float it = 1.0f / (deg ? CVX_SIMPLEX_EPS : t);
    float l2 = dot(cross(u, a), n) * it;
    float l1 = dot(cross(a, w), n) * it;
    float l0 = 1.0f - l2 - l1;
    float best = 1e30f;
    v3 closest = mk(0.0f, 0.0f, 0.0f);
    float bct[4];
    unsigned m = 0, mt;
    bc[0] = bc[1] = bc[2] = bc[3] = 0.0f;
    if (l0 < 0.0f || deg) {
        v3 cl = cvx_segment(sv, i1, i2, bct, &mt);
        float d = len_sq(cl);
        if (d < best) { memcpy(bc, bct, sizeof bct); m = mt; best = d; closest = cl; }
    }
    if (l1 < 0.0f || deg) {
        v3 cl = cvx_segment(sv, i0, i2, bct, &mt);
        float d = len_sq(cl);
        if (d < best) { memcpy(bc, bct, sizeof bct); m = mt; best = d; closest = cl; }
    }
    if (l2 < 0.0f || deg) {
        v3 cl = cvx_segment(sv, i0, i1, bct, &mt);
        float d = len_sq(cl);
        if (d < best) { memcpy(bc, bct, sizeof bct); m = mt; closest = cl; }
    }
    if (m != 0) { *mask = m; return closest; }
    bc[i0] = l0; bc[i1] = l1; bc[i2] = l2;
    *mask = (1u << i0) | (1u << i1) | (1u << i2);
    return add(add(scl(a, l0), scl(b, l1)), scl(c, l2));
}
static float cvx_det(v3 a, v3 b, v3 c, v3 d) { return dot(sub(b, a), cross(sub(c, a), sub(d, a))); }
static v3 cvx_tetra(const v3 sv[8], float bc[4], unsigned* mask) {
    v3 p0 = sv[1], p1 = sv[3], p2 = sv[5], p3 = sv[7];
    v3 z = mk(0.0f, 0.0f, 0.0f);
    float dt = cvx_det(p0, p1, p2, p3);
    int deg = fabsf(dt) < CVX_SIMPLEX_EPS;
    float inv = 1.0f / (deg ? CVX_SIMPLEX_EPS : dt);
    float l0 = cvx_det(z, p1, p2, p3) * inv;
    float l1 = cvx_det(p0, z, p2, p3) * inv;
    float l2 = cvx_det(p0, p1, z, p3) * inv;
    float l3 = 1.0f - l0 - l1 - l2;
    float best = 1e30f;
    v3 closest = z;
    float bct[4];
    unsigned m = 0, mt;
    bc[0] = bc[1] = bc[2] = bc[3] = 0.0f;
    if (l0 < 0.0f || deg) {
        v3 cl = cvx_triangle(sv, 1, 2, 3, bct, &mt);
        float d = len_sq(cl);
        if (d < best) { memcpy(bc, bct, sizeof bct); m = mt; best = d; closest = cl; }
    }
    if (l1 < 0.0f || deg) {
        v3 cl = cvx_triangle(sv, 0, 2, 3, bct, &mt);
        float d = len_sq(cl);
        if (d < best) { memcpy(bc, bct, sizeof bct); m = mt; best = d; closest = cl; }
    }
    if (l2 < 0.0f || deg) {
        v3 cl = cvx_triangle(sv, 0, 1, 3, bct, &mt);
        float d = len_sq(cl);
        if (d < best) { memcpy(bc, bct, sizeof bct); m = mt; best = d; closest = cl; }
    }
    if (l3 < 0.0f || deg) {
        v3 cl = cvx_triangle(sv, 0, 1, 2, bct, &mt);
        float d = len_sq(cl);
        if (d < best) { memcpy(bc, bct, sizeof bct); m = mt; closest = cl; }
    }
    if (m != 0) { *mask = m; return closest; }
    bc[0] = l0; bc[1] = l1; bc[2] = l2; bc[3] = l3;
    *mask = 15u;
    return z;
}
static void cvx_simplex_closest(const v3 sv[8], const float bc[4], unsigned mask, v3* pa, v3* pb) {
    v3 a = mk(0.0f, 0.0f, 0.0f), b = mk(0.0f, 0.0f, 0.0f);
    for (int i = 0; i < 4; ++i) {
        if (!(mask & (1u << i))) continue;
        a = add(a, scl(add(sv[2 * i], sv[2 * i + 1]), bc[i]));
        b = add(b, scl(sv[2 * i], bc[i]));
    }
    *pa = a; *pb = b;
}
static void cvx_gjk(const cvx_geom* ga, const cvx_geom* gb, q4 qb, v3 pb, float extend, v3* pa_out, v3* pb_out,
                    v3* n_out, float* dist_out) {
    const float ce = 1e-4f;
    v3 sv[8];
    float bc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    unsigned mask = 0;
    for (int i = 0; i < 8; ++i) sv[i] = mk(0.0f, 0.0f, 0.0f);
    v3 v = v3neg(pb);
    float dist_sq = len_sq(v);
    v3 last = mk(1.0f, 0.0f, 0.0f);
    for (int iter = 30; iter > 0;) {
        iter--;
        if (dist_sq < ce * ce) goto overlap;
        int fallback = 0;
        v3 sd = v3neg(v);
        if (dist_sq < 1.0e-12f) { sd = mk(1.0f, 0.0f, 0.0f); fallback = 1; }
        last = sd;
        mvert w = cvx_msupport(ga, gb, sd, qb, pb, extend);
        if (!fallback && dot(v, sub(v, w.BtoA)) < ce * sqrtf(dist_sq)) break;
        int dup = 0;
        for (int i = 0; i < 4; ++i)
            if ((mask & (1u << i)) && len_sq(sub(sv[2 * i + 1], w.BtoA)) < ce * ce) { dup = 1; break; }
        if (dup) break;
        int idx[4] = {0, 0, 0, 0}, use = 0, free_slot = 0;
        for (int i = 0; i < 4; ++i) {
            if (mask & (1u << i)) idx[use++] = i;
            else free_slot = i;
        }
        idx[use++] = free_slot;
        sv[2 * free_slot] = w.B;
        sv[2 * free_slot + 1] = w.BtoA;
        v3 closest;
        if (use == 1) {
            closest = sv[2 * idx[0] + 1];
            mask = 1u << idx[0];
            bc[idx[0]] = 1.0f;
        } else if (use == 2) {
            closest = cvx_segment(sv, idx[0], idx[1], bc, &mask);
        } else if (use == 3) {
            closest = cvx_triangle(sv, idx[0], idx[1], idx[2], bc, &mask);
        } else {
            closest = cvx_tetra(sv, bc, &mask);
            if (mask == 15u) goto overlap;
        }
        v = closest;
        dist_sq = len_sq(v);
    }
    {
        float distance = sqrtf(dist_sq);
        v3 pa, pbb, normal;
        cvx_simplex_closest(sv, bc, mask, &pa, &pbb);
        v3 delta = sub(pbb, pa);
        float dl = len_sq(delta);
        if (dl > CVX_SIMPLEX_EPS * CVX_SIMPLEX_EPS) {
            normal = scl(delta, 1.0f / sqrtf(dl));
        } else if (distance > ce) {
            normal = scl(v, -1.0f / distance);
        } else {
            float nsq = len_sq(last);
            normal = nsq > 0.0f ? scl(last, 1.0f / sqrtf(nsq)) : mk(1.0f, 0.0f, 0.0f);
        }
        *pa_out = pa; *pb_out = pbb; *n_out = normal; *dist_out = distance;
        return;
    }
overlap:
    cvx_simplex_closest(sv, bc, mask, pa_out, pb_out);
    *n_out = mk(0.0f, 0.0f, 0.0f);
    *dist_out = 0.0f;
}

/* ------------------------------------------------------------- manifold */
static void cvx_basis(v3 n, v3* b1, v3* b2) {
    if (n.z < 0.0f) {
        float a = 1.0f / (1.0f - n.z);
        float b = n.x * n.y * a;
        *b1 = mk(1.0f - n.x * n.x * a, -b, n.x);
        *b2 = mk(b, n.y * n.y * a - 1.0f, -n.y);
    } else {
        float a = 1.0f / (1.0f + n.z);
        float b = -n.x * n.y * a;
        *b1 = mk(1.0f - n.x * n.x * a, b, -n.x);
        *b2 = mk(b, 1.0f - n.y * n.y * a, -n.y);
    }
}
static void cvx_tracker_update(cvx_tracker* t, v3 p, int id) {
    if (id == 0) {
        t->ref = p;
        t->area = 0.0f;
    } else if (id == 1) {
        t->prev = p;
    } else {
        v3 c = cross(sub(t->prev, t->ref), sub(p, t->ref));
        float a = dot(c, c);
        if (a > t->area) { t->area = a; t->n = c; }
        t->prev = p;
    }
}
static v3 cvx_segment_normal(v3 dir, v3 refn) {
    v3 n = cross(cross(dir, refn), dir);
    float l = length(n);
    return l > 1.0e-12f ? mk(n.x / l, n.y / l, n.z / l) : refn;
}
static void cvx_projectors(const cvx_tracker* ta, v3 anchor_a, const cvx_tracker* tb, v3 anchor_b, v3 cn,
                           cvx_projector* pa, cvx_projector* pb) {
    pa->d = 0.0f; pa->n = mk(0.0f, 0.0f, 0.0f);
    pb->d = 0.0f; pb->n = mk(0.0f, 0.0f, 0.0f);
    if (ta->area == 0.0f && tb->area == 0.0f) {
        pa->n = cvx_segment_normal(sub(ta->prev, ta->ref), cn);
        pa->d = -dot(scl(add(ta->ref, ta->prev), 0.5f), pa->n);
        pb->n = cvx_segment_normal(sub(tb->prev, tb->ref), cn);
        pb->d = -dot(scl(add(tb->ref, tb->prev), 0.5f), pb->n);
        return;
    }
    if (ta->area > 0.0f) {
        float l = sqrtf(fmaxf(1.0e-12f, ta->area));
        pa->n = mk(ta->n.x / l, ta->n.y / l, ta->n.z / l);
        pa->d = -dot(anchor_a, pa->n);
    }
    if (tb->area > 0.0f) {
        float l = sqrtf(fmaxf(1.0e-12f, tb->area));
        pb->n = mk(tb->n.x / l, tb->n.y / l, tb->n.z / l);
        pb->d = -dot(anchor_b, pb->n);
    }
    if (ta->area == 0.0f) {
        pa->n = cvx_segment_normal(sub(ta->prev, ta->ref), pb->n);
        pa->d = -dot(scl(add(ta->ref, ta->prev), 0.5f), pa->n);
    }
    if (tb->area == 0.0f) {
        pb->n = cvx_segment_normal(sub(tb->prev, tb->ref), pa->n);
        pb->d = -dot(scl(add(tb->ref, tb->prev), 0.5f), pb->n);
    }
}
static v3 cvx_ray_plane(v3 o, v3 dir, float pd, v3 pn) {
    float den = dot(dir, pn);
    if (fabsf(den) < 1.0e-12f) return o;
    float t = -(dot(o, pn) + pd) / den;
    return add(o, scl(dir, t));
}

/* 2D polygon buffer; slot indices are clamped into the buffer */
static int cb(int k) { return k < 0 ? 0 : (k > CVX_BUF - 1 ? CVX_BUF - 1 : k); }
static float sarea(v2 a, v2 b, v2 q) { return (b.x - a.x) * (q.y - a.y) - (b.y - a.y) * (q.x - a.x); }
static v2 cvx_isect(v2 s, v2 e, v2 a, v2 b) {
    float sa = sarea(s, e, a), sb = sarea(s, e, b);
    float t = fabsf(sa) / fabsf(sa - sb);
    return mk2((1.0f - t) * a.x + t * b.x, (1.0f - t) * a.y + t * b.y);
}
static int cvx_trim(v2 s, v2 e, v2* buf, int base, int count) {
    if (count < 3) return count;
    v2 ia = mk2(0.0f, 0.0f), ib = mk2(0.0f, 0.0f);
    int ca = -1, cbn = -1, keep = 0;
    int prev_out = sarea(s, e, buf[cb(base)]) <= 0.0f;
    for (int i = 0; i < count; ++i) {
        int nx = (i + 1) % count;
        int out = sarea(s, e, buf[cb(base + nx)]) <= 0.0f;
        if (out != prev_out) {
            v2 pt = cvx_isect(s, e, buf[cb(base + i)], buf[cb(base + nx)]);
            if (ca < 0) { ca = i; keep = !prev_out; ia = pt; }
            else { cbn = i; ib = pt; }
        }
        prev_out = out;
    }
    if (ca >= 0 && cbn >= 0) {
        int li = -1, new_count = count;
        for (int i = 0; i < count && i < CVX_BUF; ++i) {
            if (keep) { li++; buf[cb(base + li)] = buf[cb(base + i)]; }
            if (i == ca || i == cbn) {
                v2 pt = i == ca ? ia : ib;
                if (li == i && !keep) {
                    li++;
                    for (int k = new_count < CVX_BUF ? new_count : CVX_BUF - 1; k > li; --k)
                        buf[cb(base + k)] = buf[cb(base + k - 1)];
                    buf[cb(base + li)] = pt;
                    new_count++;
                    i++;
                    cbn++;
                    count++;
                } else {
                    li++;
                    buf[cb(base + li)] = pt;
                }
                keep = !keep;
            }
        }
        return li + 1;
    }
    return prev_out ? 0 : count;
}
static int cvx_thin_rect(v2* buf, int base) {
    v2 p0 = buf[cb(base)], p1 = buf[cb(base + 1)];
    float dx = p1.x - p0.x, dy = p1.y - p0.y;
    float dl = sqrtf(dx * dx + dy * dy);
    if (!(dl > 1e-10f)) return 0;
    float ox = (-dy / dl) * 1e-5f, oy = (dx / dl) * 1e-5f;
    buf[cb(base)] = mk2(p0.x - ox, p0.y - oy);
    buf[cb(base + 1)] = mk2(p1.x - ox, p1.y - oy);
    buf[cb(base + 2)] = mk2(p1.x + ox, p1.y + oy);
    buf[cb(base + 3)] = mk2(p0.x + ox, p0.y + oy);
    return 1;
}
/* trim the loop at slot 0 (count lc) by the convex polygon at slot 5 (count tc) */
static int cvx_trim_all(v2* buf, int tc, int lc) {
    if (tc <= 1) return lc < 1 ? lc : 1;
    if (tc == 2) {
        if (!cvx_thin_rect(buf, 5)) return lc < 1 ? lc : 1;
        tc = 4;
    }
    if (lc == 2) {
        if (!cvx_thin_rect(buf, 0)) return lc < 1 ? lc : 1;
        lc = 4;
    }
    v2 t0 = buf[cb(5)];
    for (int i = 0; i < tc; ++i) {
        v2 s = buf[cb(5 + i)];
        v2 e = i == tc - 1 ? t0 : buf[cb(5 + i + 1)];
        lc = cvx_trim(s, e, buf, 0, lc);
    }
    return lc;
}
static int cvx_remove_short_edges(v2* buf, int count, float eps) {
    if (count < 2) return 0;
    int wi = 0;
    for (int ri = 1; ri < count && ri < CVX_BUF; ++ri) {
        if (len_sq2(sub2(buf[cb(ri)], buf[cb(wi)])) > eps) {
            wi++;
            buf[cb(wi)] = buf[cb(ri)];
        }
    }
    int nc;
    if (wi > 0) nc = len_sq2(sub2(buf[cb(wi)], buf[cb(0)])) < eps ? wi : wi + 1;
    else nc = wi + 1;
    return nc < 2 ? 0 : nc;
}
static void cvx_calipers(const v2* buf, int n, int out[4]) {
    const float tie = 1.0f + 1.0e-3f;
    int p1 = 0, p3 = 1;
    float best = len_sq2(sub2(buf[cb(p1)], buf[cb(p3)]));
    int j = 1;
    for (int i = 0; i < n; ++i) {
        v2 hi = buf[cb(i)], hi1 = buf[cb((i + 1) % n)];
        for (int guard = 0; guard < n; ++guard) {
            float aj1 = sarea(hi, hi1, buf[cb((j + 1) % n)]);
            float aj = sarea(hi, hi1, buf[cb(j)]);
            if (aj1 > aj) j = (j + 1) % n;
            else break;
        }
        v2 hj = buf[cb(j)];
        float d1 = len_sq2(sub2(hi, hj));
        if (d1 > best * tie) { best = d1; p1 = i; p3 = j; }
        float d2 = len_sq2(sub2(hi1, hj));
        if (d2 > best * tie) { best = d2; p1 = (i + 1) % n; p3 = j; }
    }
    int p2 = 0, p4 = 0;
    float m1 = 0.0f, m2 = 0.0f;
    v2 a = buf[cb(p1)], b = buf[cb(p3)];
    for (int i = 0; i < n; ++i) {
        float ar = sarea(a, b, buf[cb(i)]);
        if (ar > m1 * tie) { m1 = ar; p2 = i; }
        else if (-ar > m2 * tie) { m2 = -ar; p4 = i; }
    }
    out[0] = p1; out[1] = p2; out[2] = p3; out[3] = p4;
}
static int cvx_add_unique(v2* buf, int base, int count, v2 p) {
    if (count > 0 && len_sq2(sub2(buf[cb(base)], p)) < CVX_MC_EPS) return 0;
    if (count > 1 && len_sq2(sub2(buf[cb(base + count - 1)], p)) < CVX_MC_EPS) return 0;
    buf[cb(base + count)] = p;
    return 1;
}

static const float cvx_pent_c[5] = {1.0f, 0.309017f, -0.809017f, -0.809017f, 0.309017f};
static const float cvx_pent_s[5] = {0.0f, 0.95105654f, 0.58778524f, -0.58778524f, -0.95105654f};

/*
 * Manifold in A's frame (A at the origin, B at (qr, pr)); contacts go to
 * world space through (qa, pa_w).  Returns the number of contacts (0..5):
 * the trimmed polygon's points first, then the deepest point if included.
 */
static int cvx_manifold(const cvx_geom* ga, const cvx_geom* gb, q4 qa, v3 pa_w, q4 qr, v3 pr, v3 p_a, v3 p_b,
                        v3 normal, v3 pos[5], float dist[5]) {
    v2 buf[CVX_BUF];
    for (int i = 0; i < CVX_BUF; ++i) buf[i] = mk2(0.0f, 0.0f);
    cvx_tracker tra, trb;
    memset(&tra, 0, sizeof tra);
    memset(&trb, 0, sizeof trb);
    v3 ta, tb;
    cvx_basis(normal, &ta, &tb);
    v3 center = scl(add(p_a, p_b), 0.5f);
    v3 ln_b = qrot_inv(qr, v3neg(normal));
    v3 lta_b = qrot_inv(qr, v3neg(ta));
    v3 ltb_b = qrot_inv(qr, v3neg(tb));
    int na = 0, nb = 0;
    for (int e = 0; e < 5; ++e) {
        float cs = cvx_pent_c[e] * CVX_SIN_TILT, ss = cvx_pent_s[e] * CVX_SIN_TILT;
        v3 da = add(add(scl(normal, CVX_COS_TILT), scl(ta, cs)), scl(tb, ss));
        v3 qa3 = cvx_support(ga, da);
        v3 pj = sub(qa3, center);
        if (cvx_add_unique(buf, 5, na, mk2(dot(ta, pj), dot(tb, pj)))) {
            na++;
            cvx_tracker_update(&tra, qa3, na - 1);
        }
        v3 db = add(add(scl(ln_b, CVX_COS_TILT), scl(lta_b, cs)), scl(ltb_b, ss));
        v3 qb3 = add(qrot(qr, cvx_support(gb, db)), pr);
        pj = sub(qb3, center);
        if (cvx_add_unique(buf, 0, nb, mk2(dot(ta, pj), dot(tb, pj)))) {
            nb++;
            cvx_tracker_update(&trb, qb3, nb - 1);
        }
    }
    int count = 0;
    float ndot = 0.0f;
    if (na >= 2 && nb >= 2) {
        cvx_projector pja, pjb;
        cvx_projectors(&tra, p_a, &trb, p_b, normal, &pja, &pjb);
        if (!(fabsf(dot(normal, pja.n)) < CVX_COS_TILT) && !(fabsf(dot(normal, pjb.n)) < CVX_COS_TILT)) {
            ndot = fabsf(dot(pja.n, pjb.n));
            int lc = cvx_trim_all(buf, na, nb);
            lc = cvx_remove_short_edges(buf, lc, CVX_MC_EPS);
            if (lc > 1) {
                int sel[4] = {0, 1, 2, 3};
                if (lc > 4) {
                    cvx_calipers(buf, lc, sel);
                    lc = 4;
                }
                for (int i = 0; i < lc; ++i) {
                    v2 q = buf[cb(sel[i])];
                    v3 pl = add(add(scl(ta, q.x), scl(tb, q.y)), center);
                    v3 a = cvx_ray_plane(pl, normal, pja.d, pja.n);
                    v3 b = cvx_ray_plane(pl, normal, pjb.d, pjb.n);
                    pos[count] = add(qrot(qa, scl(add(a, b), 0.5f)), pa_w);
                    dist[count] = dot(sub(b, a), normal);
                    count++;
                }
            } else {
                ndot = 0.0f;
            }
        }
    }
    if (ndot < CVX_COS_DEEPEST || count == 0) {
        pos[count] = add(qrot(qa, scl(add(p_a, p_b), 0.5f)), pa_w);
        dist[count] = dot(sub(p_b, p_a), normal);
        count++;
    }
    return count;
}

/*
 * Convex pair driver (type_a <= type_b; a capsule is always shape A against a
 * box).  Fills up to five contact centres and signed distances (after the
 * Minkowski radius shift), the shared world normal and the effective radii.
 */
/* A cylinder or cone against a box (or a plane's box proxy) in rolling contact: the contact points
   move onto the plane through the shape's centre spanned by its axis and the normal (core.py:215-273).
   A cylinder rolls when its axis is within 2 degrees of perpendicular to the normal, a cone when the
   angle between axis and normal is within 2 degrees of its half angle atan2(r, 2 hh); sin(theta +- 2
   deg) is formed from sin / cos of the half angle (r and 2 hh over their hypotenuse). */
#define CVX_SIN2 0.034899496702500969f
#define CVX_COS2 0.99939082701909576f
static void cvx_axial_rolling(int ta_, v3 sa_, v3 pa_w, q4 qa, int tb_, v3 sb_, v3 pb_w, q4 qb, v3 nw, v3 pos[5],
                              int count) {
    const int disc_a = ta_ == GEO_BOX, disc_b = tb_ == GEO_BOX;
    const int ax_a = ta_ == GEO_CYLINDER || ta_ == GEO_CONE, ax_b = tb_ == GEO_CYLINDER || tb_ == GEO_CONE;
    if (!((disc_a && ax_b) || (disc_b && ax_a))) return;
    v3 axis, spos, an;
    float r, hh;
    int cone;
    if (disc_a && ax_b) {
        axis = qrot(qb, mk(0.0f, 0.0f, 1.0f)); r = sb_.x; hh = sb_.y; cone = tb_ == GEO_CONE; spos = pb_w; an = nw;
    } else {
        axis = qrot(qa, mk(0.0f, 0.0f, 1.0f)); r = sa_.x; hh = sa_.y; cone = ta_ == GEO_CONE; spos = pa_w; an = v3neg(nw);
    }
    const float adot = fabsf(dot(axis, an));
    int rolling;
    if (cone) {
        const float hyp = sqrtf(r * r + 4.0f * hh * hh);
        const float sh = r / hyp, ch = (2.0f * hh) / hyp;
        rolling = adot >= sh * CVX_COS2 - ch * CVX_SIN2 && adot <= sh * CVX_COS2 + ch * CVX_SIN2;
    } else {
        rolling = adot <= CVX_SIN2;
    }
    if (!rolling) return;
    const v3 pn = normalize(cross(axis, an));
    for (int i = 0; i < count; ++i) pos[i] = sub(pos[i], scl(pn, dot(sub(pos[i], spos), pn)));
}

static int cvx_pair(int ta_, v3 sa_, v3 pa_w, q4 qa, int tb_, v3 sb_, v3 pb_w, q4 qb, float gap_sum, float ma,
                    float mb, float* ra_out, float* rb_out, v3 pos[5], float dist[5], v3* normal_out) {
    cvx_geom ga = {ta_, sa_}, gb = {tb_, sb_};
    float ra = 0.0f, rb = 0.0f;
    if (ta_ == GEO_SPHERE || ta_ == GEO_CAPSULE) { ra = sa_.x; ga.scale.x = 0.0001f; }
    if (tb_ == GEO_SPHERE || tb_ == GEO_CAPSULE) { rb = sb_.x; gb.scale.x = 0.0001f; }
    const float threshold = gap_sum + ra + rb + ma + mb;
    const int skip_multi = ta_ == GEO_SPHERE || tb_ == GEO_SPHERE || ta_ == GEO_ELLIPSOID || tb_ == GEO_ELLIPSOID;
    q4 qr = qmul(qconj4(qa), qb);
    v3 pr = qrot_inv(qa, sub(pb_w, pa_w));
    const float enlarge = 1e-4f;
    v3 p_a, p_b, n;
    float pen, sd;
    if (cvx_mpr(&ga, &gb, qr, pr, 0.0f + enlarge, &p_a, &p_b, &n, &pen)) {
        sd = -pen + enlarge;
    } else {
        cvx_gjk(&ga, &gb, qr, pr, 0.0f, &p_a, &p_b, &n, &sd);
    }
    v3 nw = qrot(qa, n);
    int count;
    if (skip_multi || sd > threshold) {
        pos[0] = add(qrot(qa, scl(add(p_a, p_b), 0.5f)), pa_w);
        dist[0] = sd;
        count = 1;
    } else {
        count = cvx_manifold(&ga, &gb, qa, pa_w, qr, pr, p_a, p_b, n, pos, dist);
    }
    for (int i = 0; i < count; ++i) {
        if (ta_ == GEO_SPHERE || ta_ == GEO_CAPSULE) {
            pos[i] = add(pos[i], scl(nw, ra * 0.5f));
            dist[i] = dist[i] - ra;
        }
        if (tb_ == GEO_SPHERE || tb_ == GEO_CAPSULE) {
            pos[i] = sub(pos[i], scl(nw, rb * 0.5f));
            dist[i] = dist[i] - rb;
        }
    }
    cvx_axial_rolling(ta_, sa_, pa_w, qa, tb_, sb_, pb_w, qb, nw, pos, count);
    *ra_out = ra; *rb_out = rb; *normal_out = nw;
    return count;
}

#endif
