// Contacts of the pairs the analytic routines leave (two support-mapped shapes among sphere,
// capsule, ellipsoid, cylinder, box, cone; an infinite plane against a cone enters as a box) on
// gfx950: fused MPR + GJK and the tilted-probe contact manifold, one lane per candidate pair.
//
// Included by sapx_collide.hip (compiled with -ffp-contract=off and IEEE
// float32 division / square root) so the results are bit-identical to the CPU
// restatement of the same steps.
//
// Reference: sim/collision/narrow_phase.py:580-747 (GJK/MPR kernel),
// core.py:278-392 and 625-724 (per-pair driver: capsule radius replaced by
// 1e-4, MPR inflated by 1e-4, GJK when MPR finds no overlap, multi-contact
// threshold gap + radii + margins), convex.py:26-118, mpr.py:27-237 (portal
// refinement, 30 iterations, epsilon 1e-5), simplex.py:39-436 (GJK distance,
// 30 iterations, epsilon 1e-4, Johnson sub-simplex selection with the 1e-8
// degeneracy guard), multicontact.py:20-955 (manifold), support.py:94-276
// (support maps), math.py:20-44 (tangent basis), core.py:172-275 (Minkowski
// post-process and the rolling projection of cylinders and cones), core.py:360-378
// (one contact for sphere and ellipsoid pairs).
//
// Device layout: the simplex (4 Minkowski vertices, barycentric weights) lives
// in registers, indexed through unrolled selects; the manifold's 2D polygon
// buffer (16 float2 per lane) lives in LDS with a stride of one workgroup
// between a lane's slots, so a wave's accesses to one slot are contiguous.
// Shape B's polygon starts at slot 0 and shape A's at slot 5 of the same
// buffer, as in the reference, so the B polygon growing while it is trimmed
// overwrites A vertices the trim has already consumed.  Every slot index is
// clamped into the buffer.
#pragma once

namespace sapx {
namespace cvx {

constexpr float kSimplexEps = 1e-8f;
constexpr float kMcEps = 1e-5f;
constexpr int kBuf = 16;
constexpr float kSinTilt = 0.034899496f;     // sin(2 deg)
constexpr float kCosTilt = 0.99939084f;      // cos(2 deg)
constexpr float kCosDeepest = 0.99999845f;   // cos(0.1 deg)

struct Geom {
    int type;
    V3 scale;
};
struct MVert {
    V3 B, BtoA;
};

__device__ __forceinline__ float len_sq(V3 a) { return dot(a, a); }
__device__ __forceinline__ V3 nrm(V3 a) {
    float l = length(a);
    if (l > 0.0f) return v3(a.x / l, a.y / l, a.z / l);
    return v3(0.0f, 0.0f, 0.0f);
}
__device__ __forceinline__ V3 qrot_inv(Q4 q, V3 v) { return qrot(qconj(q), v); }

// ---------------------------------------------------------------- support
__device__ __forceinline__ V3 support(const Geom& g, V3 d) {
    if (g.type == kGeoBox)
        return v3(d.x >= 0.0f ? g.scale.x : -g.scale.x, d.y >= 0.0f ? g.scale.y : -g.scale.y,
                  d.z >= 0.0f ? g.scale.z : -g.scale.z);
    if (g.type == kGeoSphere || g.type == kGeoCapsule) {
        V3 n = len_sq(d) > 1e-12f ? nrm(d) : v3(1.0f, 0.0f, 0.0f);
        V3 r = n * g.scale.x;
        if (g.type == kGeoCapsule) r = r + v3(0.0f, 0.0f, d.z >= 0.0f ? g.scale.y : -g.scale.y);
        return r;
    }
    if (g.type == kGeoEllipsoid) {
        const float a = g.scale.x, b = g.scale.y, c = g.scale.z;
        if (len_sq(d) > 1e-12f) {
            const float adx = a * d.x, bdy = b * d.y, cdz = c * d.z;
            const float den2 = adx * adx + bdy * bdy + cdz * cdz;
            if (den2 > 1e-12f) {
                const float den = sqrtf(den2);
                return v3((a * a) * d.x / den, (b * b) * d.y / den, (c * c) * d.z / den);
            }
        }
        return v3(a, 0.0f, 0.0f);
    }
    if (g.type == kGeoCylinder) {
        const float r = g.scale.x, hh = g.scale.y;
        V3 lat = v3(r, 0.0f, 0.0f);
        const V3 dxy = v3(d.x, d.y, 0.0f);
        if (len_sq(dxy) > 1e-12f) {
            const V3 nxy = nrm(dxy);
            lat = v3(nxy.x * r, nxy.y * r, 0.0f);
        }
        if (d.z > 0.0f) return v3(lat.x, lat.y, hh);
        if (d.z < 0.0f) return v3(lat.x, lat.y, -hh);
        return lat;
    }
    if (g.type == kGeoCone) {
        // apex at +hh, base disk at -hh: the apex when dz >= r / (2 hh) |d_xy|
        const float r = g.scale.x, hh = g.scale.y;
        const float lxy = length(v3(d.x, d.y, 0.0f));
        const float k = hh > 1e-12f ? r / (2.0f * hh) : 0.0f;
        if (lxy <= 1e-12f) return d.z >= 0.0f ? v3(0.0f, 0.0f, hh) : v3(r, 0.0f, -hh);
        if (d.z >= k * lxy) return v3(0.0f, 0.0f, hh);
        return v3(d.x / lxy * r, d.y / lxy * r, -hh);
    }
    return v3(0.0f, 0.0f, 0.0f);
}
__device__ __forceinline__ MVert msupport(const Geom& ga, const Geom& gb, V3 d, Q4 qb, V3 pb, float extend) {
    MVert v;
    V3 pa = support(ga, d);
    v.B = qrot(qb, support(gb, qrot_inv(qb, -d))) + pb;
    if (extend != 0.0f) {
        V3 e = (nrm(d) * extend) * 0.5f;
        pa = pa + e;
        v.B = v.B - e;
    }
    v.BtoA = pa - v.B;
    return v;
}
__device__ __forceinline__ V3 vert_a(const MVert& v) { return v.B + v.BtoA; }

// -------------------------------------------------------------------- MPR
__device__ bool mpr(const Geom& ga, const Geom& gb, Q4 qb, V3 pb, float extend, V3* pa_out, V3* pb_out, V3* n_out,
                    float* pen_out) {
    const float ne = 1e-16f;
    const float ce = 1e-5f;
    const int max_iter = 30;
    float pen = 0.0f;
    V3 normal;
    MVert v0, v1, v2, v3_, v4;
    bool hit = false, ret = false;
    v0.B = pb;
    v0.BtoA = -pb;
    normal = v0.BtoA;
    if (fabsf(normal.x) < ne && fabsf(normal.y) < ne && fabsf(normal.z) < ne) v0.BtoA = v3(1e-5f, 0.0f, 0.0f);
    normal = -v0.BtoA;
    v1 = msupport(ga, gb, normal, qb, pb, extend);
    V3 pa = vert_a(v1), pbb = v1.B;
    do {
        if (dot(v1.BtoA, normal) <= 0.0f) break;
        normal = cross(v1.BtoA, v0.BtoA);
        if (len_sq(normal) < ne * ne) {
            normal = nrm(v1.BtoA - v0.BtoA);
            pen = dot(v1.BtoA, normal);
            ret = true;
            break;
        }
        v2 = msupport(ga, gb, normal, qb, pb, extend);
        if (dot(v2.BtoA, normal) <= 0.0f) break;
        normal = cross(v1.BtoA - v0.BtoA, v2.BtoA - v0.BtoA);
        if (dot(normal, v0.BtoA) > 0.0f) {
            MVert t = v1;
            v1 = v2;
            v2 = t;
            normal = -normal;
        }
        bool found = false;
        for (int phase1 = 0; phase1 <= max_iter;) {
            phase1++;
            v3_ = msupport(ga, gb, normal, qb, pb, extend);
            if (dot(v3_.BtoA, normal) <= 0.0f) break;
            if (dot(cross(v1.BtoA, v3_.BtoA), v0.BtoA) < 0.0f) {
                v2 = v3_;
                normal = cross(v1.BtoA - v0.BtoA, v3_.BtoA - v0.BtoA);
                continue;
            }
            if (dot(cross(v3_.BtoA, v2.BtoA), v0.BtoA) < 0.0f) {
                v1 = v3_;
                normal = cross(v3_.BtoA - v0.BtoA, v2.BtoA - v0.BtoA);
                continue;
            }
            found = true;
            break;
        }
        if (!found) break;
        // each pass either exits or replaces a portal vertex; phase 2 ends after at most
        // max_iter + 1 passes
        for (int phase2 = 1;; ++phase2) {
            V3 t1 = v2.BtoA - v1.BtoA, t2 = v3_.BtoA - v1.BtoA;
            normal = cross(t1, t2);
            float nsq = len_sq(normal);
            if (nsq < ne * ne) break;
            if (!hit) hit = dot(normal, v1.BtoA) >= 0.0f;
            v4 = msupport(ga, gb, normal, qb, pb, extend);
            float delta = dot(v4.BtoA - v3_.BtoA, normal);
            pen = dot(v4.BtoA, normal);
            if (delta * delta <= ce * ce * nsq || pen <= 0.0f || phase2 > max_iter) {
                if (hit) {
                    float inv = 1.0f / sqrtf(nsq);
                    pen *= inv;
                    normal = normal * inv;
                    float gamma = dot(cross(v1.BtoA, t1), normal) * inv;
                    float beta = dot(cross(t2, v1.BtoA), normal) * inv;
                    float alpha = 1.0f - gamma - beta;
                    pa = vert_a(v1) * alpha + vert_a(v2) * beta + vert_a(v3_) * gamma;
                    pbb = v1.B * alpha + v2.B * beta + v3_.B * gamma;
                }
                ret = hit;
                break;
            }
            V3 t4 = cross(v4.BtoA, v0.BtoA);
            if (dot(t4, v1.BtoA) >= 0.0f) {
                if (dot(t4, v2.BtoA) >= 0.0f) v1 = v4; else v3_ = v4;
            } else {
                if (dot(t4, v3_.BtoA) >= 0.0f) v2 = v4; else v1 = v4;
            }
        }
    } while (false);
    *pa_out = pa;
    *pb_out = pbb;
    *n_out = normal;
    *pen_out = pen;
    return ret;
}

// -------------------------------------------------------------------- GJK
struct Simplex {
    V3 b[4], m[4];  // point on B, Minkowski point BtoA
    float bc[4];
    unsigned mask;
};
__device__ __forceinline__ V3 pick(const V3 (&a)[4], int k) {
    V3 r = a[0];
    if (k == 1) r = a[1];
    if (k == 2) r = a[2];
    if (k == 3) r = a[3];
    return r;
}
__device__ __forceinline__ void put(V3 (&a)[4], int k, V3 v) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (j == k) a[j] = v;
}
__device__ __forceinline__ void putf(float (&a)[4], int k, float v) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (j == k) a[j] = v;
}
__device__ __forceinline__ V3 segment(const V3 (&m)[4], int i0, int i1, float (&bc)[4], unsigned* mask) {
    V3 a = pick(m, i0), b = pick(m, i1);
    V3 e = b - a;
    float vsq = len_sq(e);
    bool deg = vsq < kSimplexEps;
    float t = -dot(a, e) / (deg ? kSimplexEps : vsq);
    float l0 = 1.0f - t, l1 = t;
    unsigned mk = (1u << i0) | (1u << i1);
    if (l0 < 0.0f || deg) { mk = 1u << i1; l0 = 0.0f; l1 = 1.0f; }
    else if (l1 < 0.0f) { mk = 1u << i0; l0 = 1.0f; l1 = 0.0f; }
    bc[0] = bc[1] = bc[2] = bc[3] = 0.0f;
    putf(bc, i0, l0);
    putf(bc, i1, l1);
    *mask = mk;
    return a * l0 + b * l1;
}
__device__ V3 triangle(const V3 (&m)[4], int i0, int i1, int i2, float (&bc)[4], unsigned* mask) {
    V3 a = pick(m, i0), b = pick(m, i1), c = pick(m, i2);
    V3 u = a - b, w = a - c;
    V3 n = cross(u, w);
    float t = len_sq(n);
    bool deg = t < kSimplexEps;
    float it = 1.0f / (deg ? kSimplexEps : t);
    float l2 = dot(cross(u, a), n) * it;
    float l1 = dot(cross(a, w), n) * it;
    float l0 = 1.0f - l2 - l1;
    float best = 1e30f;
    V3 closest = v3(0.0f, 0.0f, 0.0f);
    float bct[4];
    unsigned mk = 0, mt;
    bc[0] = bc[1] = bc[2] = bc[3] = 0.0f;
    if (l0 < 0.0f || deg) {
        V3 cl = segment(m, i1, i2, bct, &mt);
        float d = len_sq(cl);
        if (d < best) { bc[0] = bct[0]; bc[1] = bct[1]; bc[2] = bct[2]; bc[3] = bct[3]; mk = mt; best = d; closest = cl; }
    }
    if (l1 < 0.0f || deg) {
        V3 cl = segment(m, i0, i2, bct, &mt);
        float d = len_sq(cl);
        if (d < best) { bc[0] = bct[0]; bc[1] = bct[1]; bc[2] = bct[2]; bc[3] = bct[3]; mk = mt; best = d; closest = cl; }
    }
    if (l2 < 0.0f || deg) {
        V3 cl = segment(m, i0, i1, bct, &mt);
        float d = len_sq(cl);
        if (d < best) { bc[0] = bct[0]; bc[1] = bct[1]; bc[2] = bct[2]; bc[3] = bct[3]; mk = mt; closest = cl; }
    }
    if (mk != 0) { *mask = mk; return closest; }
    putf(bc, i0, l0);
    putf(bc, i1, l1);
    putf(bc, i2, l2);
    *mask = (1u << i0) | (1u << i1) | (1u << i2);
    return a * l0 + b * l1 + c * l2;
}
__device__ __forceinline__ float det4(V3 a, V3 b, V3 c, V3 d) { return dot(b - a, cross(c - a, d - a)); }
__device__ V3 tetra(const V3 (&m)[4], float (&bc)[4], unsigned* mask) {
    const V3 z = v3(0.0f, 0.0f, 0.0f);
    float dt = det4(m[0], m[1], m[2], m[3]);
    bool deg = fabsf(dt) < kSimplexEps;
    float inv = 1.0f / (deg ? kSimplexEps : dt);
    float l0 = det4(z, m[1], m[2], m[3]) * inv;
    float l1 = det4(m[0], z, m[2], m[3]) * inv;
    float l2 = det4(m[0], m[1], z, m[3]) * inv;
    float l3 = 1.0f - l0 - l1 - l2;
    float best = 1e30f;
    V3 closest = z;
    float bct[4];
    unsigned mk = 0, mt;
    bc[0] = bc[1] = bc[2] = bc[3] = 0.0f;
    if (l0 < 0.0f || deg) {
        V3 cl = triangle(m, 1, 2, 3, bct, &mt);
        float d = len_sq(cl);
        if (d < best) { bc[0] = bct[0]; bc[1] = bct[1]; bc[2] = bct[2]; bc[3] = bct[3]; mk = mt; best = d; closest = cl; }
    }
    if (l1 < 0.0f || deg) {
        V3 cl = triangle(m, 0, 2, 3, bct, &mt);
        float d = len_sq(cl);
        if (d < best) { bc[0] = bct[0]; bc[1] = bct[1]; bc[2] = bct[2]; bc[3] = bct[3]; mk = mt; best = d; closest = cl; }
    }
    if (l2 < 0.0f || deg) {
        V3 cl = triangle(m, 0, 1, 3, bct, &mt);
        float d = len_sq(cl);
        if (d < best) { bc[0] = bct[0]; bc[1] = bct[1]; bc[2] = bct[2]; bc[3] = bct[3]; mk = mt; best = d; closest = cl; }
    }
    if (l3 < 0.0f || deg) {
        V3 cl = triangle(m, 0, 1, 2, bct, &mt);
        float d = len_sq(cl);
        if (d < best) { bc[0] = bct[0]; bc[1] = bct[1]; bc[2] = bct[2]; bc[3] = bct[3]; mk = mt; closest = cl; }
    }
    if (mk != 0) { *mask = mk; return closest; }
    bc[0] = l0; bc[1] = l1; bc[2] = l2; bc[3] = l3;
    *mask = 15u;
    return z;
}
__device__ __forceinline__ void simplex_closest(const Simplex& s, V3* pa, V3* pb) {
    V3 a = v3(0.0f, 0.0f, 0.0f), b = v3(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (!(s.mask & (1u << i))) continue;
        a = a + (s.b[i] + s.m[i]) * s.bc[i];
        b = b + s.b[i] * s.bc[i];
    }
    *pa = a;
    *pb = b;
}
__device__ void gjk(const Geom& ga, const Geom& gb, Q4 qb, V3 pb, float extend, V3* pa_out, V3* pb_out, V3* n_out,
                    float* dist_out) {
    const float ce = 1e-4f;
    Simplex s;
#pragma unroll
    for (int i = 0; i < 4; ++i) { s.b[i] = v3(0.0f, 0.0f, 0.0f); s.m[i] = s.b[i]; s.bc[i] = 0.0f; }
    s.mask = 0;
    V3 v = -pb;
    float dist_sq = len_sq(v);
    V3 last = v3(1.0f, 0.0f, 0.0f);
    bool overlap = false;
    for (int iter = 30; iter > 0;) {
        iter--;
        if (dist_sq < ce * ce) { overlap = true; break; }
        bool fallback = false;
        V3 sd = -v;
        if (dist_sq < 1.0e-12f) { sd = v3(1.0f, 0.0f, 0.0f); fallback = true; }
        last = sd;
        MVert w = msupport(ga, gb, sd, qb, pb, extend);
        if (!fallback && dot(v, v - w.BtoA) < ce * sqrtf(dist_sq)) break;
        bool dup = false;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (!dup && (s.mask & (1u << i)) && len_sq(s.m[i] - w.BtoA) < ce * ce) dup = true;
        if (dup) break;
        int i0 = 0, i1 = 0, i2 = 0, use = 0, free_slot = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (s.mask & (1u << i)) {
                if (use == 0) i0 = i; else if (use == 1) i1 = i; else i2 = i;
                use++;
            } else {
                free_slot = i;
            }
        }
        if (use == 0) i0 = free_slot; else if (use == 1) i1 = free_slot; else if (use == 2) i2 = free_slot;
        use++;
        put(s.b, free_slot, w.B);
        put(s.m, free_slot, w.BtoA);
        V3 closest;
        if (use == 1) {
            closest = pick(s.m, i0);
            s.mask = 1u << i0;
            putf(s.bc, i0, 1.0f);
        } else if (use == 2) {
            closest = segment(s.m, i0, i1, s.bc, &s.mask);
        } else if (use == 3) {
            closest = triangle(s.m, i0, i1, i2, s.bc, &s.mask);
        } else {
            closest = tetra(s.m, s.bc, &s.mask);
            if (s.mask == 15u) { overlap = true; break; }
        }
        v = closest;
        dist_sq = len_sq(v);
    }
    if (overlap) {
        simplex_closest(s, pa_out, pb_out);
        *n_out = v3(0.0f, 0.0f, 0.0f);
        *dist_out = 0.0f;
        return;
    }
    float distance = sqrtf(dist_sq);
    V3 pa, pbb, normal;
    simplex_closest(s, &pa, &pbb);
    V3 delta = pbb - pa;
    float dl = len_sq(delta);
    if (dl > kSimplexEps * kSimplexEps) {
        normal = delta * (1.0f / sqrtf(dl));
    } else if (distance > ce) {
        normal = v * (-1.0f / distance);
    } else {
        float nsq = len_sq(last);
        normal = nsq > 0.0f ? last * (1.0f / sqrtf(nsq)) : v3(1.0f, 0.0f, 0.0f);
    }
    *pa_out = pa;
    *pb_out = pbb;
    *n_out = normal;
    *dist_out = distance;
}

// --------------------------------------------------------------- manifold
struct Tracker {
    V3 ref, prev, n;
    float area;
};
struct Projector {
    float d;
    V3 n;
};
__device__ __forceinline__ void basis(V3 n, V3* b1, V3* b2) {
    if (n.z < 0.0f) {
        float a = 1.0f / (1.0f - n.z);
        float b = n.x * n.y * a;
        *b1 = v3(1.0f - n.x * n.x * a, -b, n.x);
        *b2 = v3(b, n.y * n.y * a - 1.0f, -n.y);
    } else {
        float a = 1.0f / (1.0f + n.z);
        float b = -n.x * n.y * a;
        *b1 = v3(1.0f - n.x * n.x * a, b, -n.x);
        *b2 = v3(b, 1.0f - n.y * n.y * a, -n.y);
    }
}
__device__ __forceinline__ void tracker_update(Tracker& t, V3 p, int id) {
    if (id == 0) {
        t.ref = p;
        t.area = 0.0f;
    } else if (id == 1) {
        t.prev = p;
    } else {
        V3 c = cross(t.prev - t.ref, p - t.ref);
        float a = dot(c, c);
        if (a > t.area) { t.area = a; t.n = c; }
        t.prev = p;
    }
}
__device__ __forceinline__ V3 segment_normal(V3 dir, V3 refn) {
    V3 n = cross(cross(dir, refn), dir);
    float l = length(n);
    return l > 1.0e-12f ? v3(n.x / l, n.y / l, n.z / l) : refn;
}
__device__ void projectors(const Tracker& ta, V3 anchor_a, const Tracker& tb, V3 anchor_b, V3 cn, Projector* pa,
                           Projector* pb) {
    pa->d = 0.0f; pa->n = v3(0.0f, 0.0f, 0.0f);
    pb->d = 0.0f; pb->n = v3(0.0f, 0.0f, 0.0f);
    if (ta.area == 0.0f && tb.area == 0.0f) {
        pa->n = segment_normal(ta.prev - ta.ref, cn);
        pa->d = -dot((ta.ref + ta.prev) * 0.5f, pa->n);
        pb->n = segment_normal(tb.prev - tb.ref, cn);
        pb->d = -dot((tb.ref + tb.prev) * 0.5f, pb->n);
        return;
    }
    if (ta.area > 0.0f) {
        float l = sqrtf(fmaxf(1.0e-12f, ta.area));
        pa->n = v3(ta.n.x / l, ta.n.y / l, ta.n.z / l);
        pa->d = -dot(anchor_a, pa->n);
    }
    if (tb.area > 0.0f) {
        float l = sqrtf(fmaxf(1.0e-12f, tb.area));
        pb->n = v3(tb.n.x / l, tb.n.y / l, tb.n.z / l);
        pb->d = -dot(anchor_b, pb->n);
    }
    if (ta.area == 0.0f) {
        pa->n = segment_normal(ta.prev - ta.ref, pb->n);
        pa->d = -dot((ta.ref + ta.prev) * 0.5f, pa->n);
    }
    if (tb.area == 0.0f) {
        pb->n = segment_normal(tb.prev - tb.ref, pa->n);
        pb->d = -dot((tb.ref + tb.prev) * 0.5f, pb->n);
    }
}
__device__ __forceinline__ V3 ray_plane(V3 o, V3 dir, float pd, V3 pn) {
    float den = dot(dir, pn);
    if (fabsf(den) < 1.0e-12f) return o;
    float t = -(dot(o, pn) + pd) / den;
    return o + dir * t;
}

// a lane's 16-slot polygon buffer in LDS (slot k at p[k * stride])
struct Poly {
    float2* p;
    int stride;
    __device__ __forceinline__ float2 ld(int k) const { return p[(k < 0 ? 0 : (k > kBuf - 1 ? kBuf - 1 : k)) * stride]; }
    __device__ __forceinline__ void st(int k, float2 v) const {
        p[(k < 0 ? 0 : (k > kBuf - 1 ? kBuf - 1 : k)) * stride] = v;
    }
};
__device__ __forceinline__ float2 f2(float x, float y) { return make_float2(x, y); }
__device__ __forceinline__ float len_sq2(float2 a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ float2 sub2(float2 a, float2 b) { return f2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float sarea(float2 a, float2 b, float2 q) {
    return (b.x - a.x) * (q.y - a.y) - (b.y - a.y) * (q.x - a.x);
}
__device__ __forceinline__ float2 isect(float2 s, float2 e, float2 a, float2 b) {
    float sa = sarea(s, e, a), sb = sarea(s, e, b);
    float t = fabsf(sa) / fabsf(sa - sb);
    return f2((1.0f - t) * a.x + t * b.x, (1.0f - t) * a.y + t * b.y);
}
// Sutherland-Hodgman clip of the loop at slot 0 against the line s->e, in place
__device__ int trim(float2 s, float2 e, const Poly& L, int count) {
    if (count < 3) return count;
    float2 ia = f2(0.0f, 0.0f), ib = f2(0.0f, 0.0f);
    int ca = -1, cbn = -1;
    bool keep = false;
    bool prev_out = sarea(s, e, L.ld(0)) <= 0.0f;
    for (int i = 0; i < count; ++i) {
        int nx = (i + 1) % count;
        bool out = sarea(s, e, L.ld(nx)) <= 0.0f;
        if (out != prev_out) {
            float2 pt = isect(s, e, L.ld(i), L.ld(nx));
            if (ca < 0) { ca = i; keep = !prev_out; ia = pt; }
            else { cbn = i; ib = pt; }
        }
        prev_out = out;
    }
    if (ca >= 0 && cbn >= 0) {
        int li = -1, new_count = count;
        // the loop below visits each vertex once plus at most two insertions
        for (int i = 0; i < count && i < kBuf; ++i) {
            if (keep) { li++; L.st(li, L.ld(i)); }
            if (i == ca || i == cbn) {
                float2 pt = i == ca ? ia : ib;
                if (li == i && !keep) {
                    li++;
                    for (int k = new_count < kBuf ? new_count : kBuf - 1; k > li; --k) L.st(k, L.ld(k - 1));
                    L.st(li, pt);
                    new_count++;
                    i++;
                    cbn++;
                    count++;
                } else {
                    li++;
                    L.st(li, pt);
                }
                keep = !keep;
            }
        }
        return li + 1;
    }
    return prev_out ? 0 : count;
}
__device__ bool thin_rect(const Poly& B, int base) {
    float2 p0 = B.ld(base), p1 = B.ld(base + 1);
    float dx = p1.x - p0.x, dy = p1.y - p0.y;
    float dl = sqrtf(dx * dx + dy * dy);
    if (!(dl > 1e-10f)) return false;
    float ox = (-dy / dl) * 1e-5f, oy = (dx / dl) * 1e-5f;
    B.st(base, f2(p0.x - ox, p0.y - oy));
    B.st(base + 1, f2(p1.x - ox, p1.y - oy));
    B.st(base + 2, f2(p1.x + ox, p1.y + oy));
    B.st(base + 3, f2(p0.x + ox, p0.y + oy));
    return true;
}
__device__ int trim_all(const Poly& B, int tc, int lc) {
    if (tc <= 1) return lc < 1 ? lc : 1;
    if (tc == 2) {
        if (!thin_rect(B, 5)) return lc < 1 ? lc : 1;
        tc = 4;
    }
    if (lc == 2) {
        if (!thin_rect(B, 0)) return lc < 1 ? lc : 1;
        lc = 4;
    }
    float2 t0 = B.ld(5);
    for (int i = 0; i < tc; ++i) {
        float2 s = B.ld(5 + i);
        float2 e = i == tc - 1 ? t0 : B.ld(5 + i + 1);
        lc = trim(s, e, B, lc);
    }
    return lc;
}
__device__ int remove_short_edges(const Poly& B, int count, float eps) {
    if (count < 2) return 0;
    int wi = 0;
    for (int ri = 1; ri < count && ri < kBuf; ++ri) {
        if (len_sq2(sub2(B.ld(ri), B.ld(wi))) > eps) {
            wi++;
            B.st(wi, B.ld(ri));
        }
    }
    int nc;
    if (wi > 0) nc = len_sq2(sub2(B.ld(wi), B.ld(0))) < eps ? wi : wi + 1;
    else nc = wi + 1;
    return nc < 2 ? 0 : nc;
}
__device__ void calipers(const Poly& B, int n, int out[4]) {
    const float tie = 1.0f + 1.0e-3f;
    int p1 = 0, p3 = 1;
    float best = len_sq2(sub2(B.ld(p1), B.ld(p3)));
    int j = 1;
    for (int i = 0; i < n; ++i) {
        float2 hi = B.ld(i), hi1 = B.ld((i + 1) % n);
        for (int guard = 0; guard < n; ++guard) {
            float aj1 = sarea(hi, hi1, B.ld((j + 1) % n));
            float aj = sarea(hi, hi1, B.ld(j));
            if (aj1 > aj) j = (j + 1) % n;
            else break;
        }
        float2 hj = B.ld(j);
        float d1 = len_sq2(sub2(hi, hj));
        if (d1 > best * tie) { best = d1; p1 = i; p3 = j; }
        float d2 = len_sq2(sub2(hi1, hj));
        if (d2 > best * tie) { best = d2; p1 = (i + 1) % n; p3 = j; }
    }
    int p2 = 0, p4 = 0;
    float m1 = 0.0f, m2 = 0.0f;
    float2 a = B.ld(p1), b = B.ld(p3);
    for (int i = 0; i < n; ++i) {
        float ar = sarea(a, b, B.ld(i));
        if (ar > m1 * tie) { m1 = ar; p2 = i; }
        else if (-ar > m2 * tie) { m2 = -ar; p4 = i; }
    }
    out[0] = p1; out[1] = p2; out[2] = p3; out[3] = p4;
}
__device__ __forceinline__ bool add_unique(const Poly& B, int base, int count, float2 p) {
    if (count > 0 && len_sq2(sub2(B.ld(base), p)) < kMcEps) return false;
    if (count > 1 && len_sq2(sub2(B.ld(base + count - 1), p)) < kMcEps) return false;
    B.st(base + count, p);
    return true;
}
__device__ __forceinline__ void pent(int e, float* c, float* s) {
    // cos / sin of 0, 72, 144, 216, 288 degrees
    *c = e == 0 ? 1.0f : (e == 1 || e == 4 ? 0.309017f : -0.809017f);
    *s = e == 0 ? 0.0f : (e == 1 ? 0.95105654f : (e == 2 ? 0.58778524f : (e == 3 ? -0.58778524f : -0.95105654f)));
}

// manifold in A's frame (A at the origin, B at (qr, pr)); contacts to world through (qa, pa_w)
__device__ int manifold(const Geom& ga, const Geom& gb, Q4 qa, V3 pa_w, Q4 qr, V3 pr, V3 p_a, V3 p_b, V3 normal,
                        const Poly& B, V3 pos[5], float dist[5]) {
    for (int i = 0; i < kBuf; ++i) B.st(i, f2(0.0f, 0.0f));
    Tracker tra, trb;
    tra.ref = tra.prev = tra.n = v3(0.0f, 0.0f, 0.0f); tra.area = 0.0f;
    trb = tra;
    V3 ta, tb;
    basis(normal, &ta, &tb);
    V3 center = (p_a + p_b) * 0.5f;
    V3 ln_b = qrot_inv(qr, -normal);
    V3 lta_b = qrot_inv(qr, -ta);
    V3 ltb_b = qrot_inv(qr, -tb);
    int na = 0, nb = 0;
    for (int e = 0; e < 5; ++e) {
        float c, s;
        pent(e, &c, &s);
        float cs = c * kSinTilt, ss = s * kSinTilt;
        V3 da = normal * kCosTilt + ta * cs + tb * ss;
        V3 qa3 = support(ga, da);
        V3 pj = qa3 - center;
        if (add_unique(B, 5, na, f2(dot(ta, pj), dot(tb, pj)))) {
            na++;
            tracker_update(tra, qa3, na - 1);
        }
        V3 db = ln_b * kCosTilt + lta_b * cs + ltb_b * ss;
        V3 qb3 = qrot(qr, support(gb, db)) + pr;
        pj = qb3 - center;
        if (add_unique(B, 0, nb, f2(dot(ta, pj), dot(tb, pj)))) {
            nb++;
            tracker_update(trb, qb3, nb - 1);
        }
    }
    int count = 0;
    float ndot = 0.0f;
    if (na >= 2 && nb >= 2) {
        Projector pja, pjb;
        projectors(tra, p_a, trb, p_b, normal, &pja, &pjb);
        if (!(fabsf(dot(normal, pja.n)) < kCosTilt) && !(fabsf(dot(normal, pjb.n)) < kCosTilt)) {
            ndot = fabsf(dot(pja.n, pjb.n));
            int lc = trim_all(B, na, nb);
            lc = remove_short_edges(B, lc, kMcEps);
            if (lc > 1) {
                int sel[4] = {0, 1, 2, 3};
                if (lc > 4) {
                    calipers(B, lc, sel);
                    lc = 4;
                }
                for (int i = 0; i < lc; ++i) {
                    float2 q = B.ld(sel[i]);
                    V3 pl = ta * q.x + tb * q.y + center;
                    V3 a = ray_plane(pl, normal, pja.d, pja.n);
                    V3 b = ray_plane(pl, normal, pjb.d, pjb.n);
                    pos[count] = qrot(qa, (a + b) * 0.5f) + pa_w;
                    dist[count] = dot(b - a, normal);
                    count++;
                }
            } else {
                ndot = 0.0f;
            }
        }
    }
    if (ndot < kCosDeepest || count == 0) {
        pos[count] = qrot(qa, (p_a + p_b) * 0.5f) + pa_w;
        dist[count] = dot(p_b - p_a, normal);
        count++;
    }
    return count;
}

// A cylinder or cone against a box (or an infinite plane's box proxy) in rolling contact: the
// contact points move onto the plane through the shape's centre spanned by its axis and the
// normal.  A cylinder rolls when its axis is within 2 degrees of perpendicular to the normal, a
// cone when the angle between axis and normal is within 2 degrees of its half angle atan2(r, 2 hh);
// sin(half angle +- 2 deg) is formed from the half angle's sine and cosine (r and 2 hh over their
// hypotenuse).
constexpr float kSin2 = 0.034899496702500969f;
constexpr float kCos2 = 0.99939082701909576f;
__device__ void axial_rolling(int ta_, V3 sa_, V3 pa_w, Q4 qa, int tb_, V3 sb_, V3 pb_w, Q4 qb, V3 nw, V3 pos[5],
                              int count) {
    const bool disc_a = ta_ == kGeoBox, disc_b = tb_ == kGeoBox;
    const bool ax_a = ta_ == kGeoCylinder || ta_ == kGeoCone, ax_b = tb_ == kGeoCylinder || tb_ == kGeoCone;
    if (!((disc_a && ax_b) || (disc_b && ax_a))) return;
    V3 axis, spos, an;
    float r, hh;
    bool cone;
    if (disc_a && ax_b) {
        axis = qrot(qb, v3(0.0f, 0.0f, 1.0f)); r = sb_.x; hh = sb_.y; cone = tb_ == kGeoCone; spos = pb_w; an = nw;
    } else {
        axis = qrot(qa, v3(0.0f, 0.0f, 1.0f)); r = sa_.x; hh = sa_.y; cone = ta_ == kGeoCone; spos = pa_w; an = -nw;
    }
    const float adot = fabsf(dot(axis, an));
    bool rolling;
    if (cone) {
        const float hyp = sqrtf(r * r + 4.0f * hh * hh);
        const float sh = r / hyp, ch = (2.0f * hh) / hyp;
        rolling = adot >= sh * kCos2 - ch * kSin2 && adot <= sh * kCos2 + ch * kSin2;
    } else {
        rolling = adot <= kSin2;
    }
    if (!rolling) return;
    const V3 pn = nrm(cross(axis, an));
    for (int i = 0; i < 5; ++i) {
        if (i >= count) break;
        pos[i] = pos[i] - pn * dot(pos[i] - spos, pn);
    }
}

// Pair driver (type_a <= type_b): up to five contact centres and signed distances after
// the Minkowski radius shift, the shared world normal and the effective radii.
__device__ int pair(int ta_, V3 sa_, V3 pa_w, Q4 qa, int tb_, V3 sb_, V3 pb_w, Q4 qb, float gap_sum, float ma,
                    float mb, const Poly& B, float* ra_out, float* rb_out, V3 pos[5], float dist[5], V3* normal_out) {
    Geom ga{ta_, sa_}, gb{tb_, sb_};
    float ra = 0.0f, rb = 0.0f;
    if (ta_ == kGeoSphere || ta_ == kGeoCapsule) { ra = sa_.x; ga.scale.x = 0.0001f; }
    if (tb_ == kGeoSphere || tb_ == kGeoCapsule) { rb = sb_.x; gb.scale.x = 0.0001f; }
    const float threshold = gap_sum + ra + rb + ma + mb;
    const bool skip_multi =
        ta_ == kGeoSphere || tb_ == kGeoSphere || ta_ == kGeoEllipsoid || tb_ == kGeoEllipsoid;
    Q4 qr = qmul(qconj(qa), qb);
    V3 pr = qrot_inv(qa, pb_w - pa_w);
    const float enlarge = 1e-4f;
    V3 p_a, p_b, n;
    float pen, sd;
    if (mpr(ga, gb, qr, pr, 0.0f + enlarge, &p_a, &p_b, &n, &pen)) {
        sd = -pen + enlarge;
    } else {
        gjk(ga, gb, qr, pr, 0.0f, &p_a, &p_b, &n, &sd);
    }
    V3 nw = qrot(qa, n);
    int count;
    if (skip_multi || sd > threshold) {
        pos[0] = qrot(qa, (p_a + p_b) * 0.5f) + pa_w;
        dist[0] = sd;
        count = 1;
    } else {
        count = manifold(ga, gb, qa, pa_w, qr, pr, p_a, p_b, n, B, pos, dist);
    }
    for (int i = 0; i < 5; ++i) {
        if (i >= count) break;
        if (ta_ == kGeoSphere || ta_ == kGeoCapsule) {
            pos[i] = pos[i] + nw * (ra * 0.5f);
            dist[i] = dist[i] - ra;
        }
        if (tb_ == kGeoSphere || tb_ == kGeoCapsule) {
            pos[i] = pos[i] - nw * (rb * 0.5f);
            dist[i] = dist[i] - rb;
        }
    }
    axial_rolling(ta_, sa_, pa_w, qa, tb_, sb_, pb_w, qb, nw, pos, count);
    *ra_out = ra;
    *rb_out = rb;
    *normal_out = nw;
    return count;
}

}  // namespace cvx
}  // namespace sapx
