// dem_device.hpp -- FP64 contact law and integrator arithmetic for gfx950.
//
// Device functions used by the kernels in dem_hip.hip.  They follow the
// reference's CUDA law expression by expression (src/cuda/cuda_contact_backend.cu
// and src/cuda/cuda_particle_backend.cu; lines cited per function) so that,
// compiled without FMA contraction, every force, torque and history value
// is the correctly rounded result of the same operation sequence.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

// Law arithmetic compiles for the device and for the host (the host copies
// serve the derivative checks; the kernels use the device copies).
#define DEM_FN __host__ __device__ __forceinline__

namespace dem {

constexpr double kPi = 3.1415926535897932384626433832795;
constexpr double kGravityConstant = 6.67430e-11;

struct V3 {
    double x, y, z;
};

DEM_FN V3 mk(double x, double y, double z) { return V3{x, y, z}; }
DEM_FN V3 add(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
DEM_FN V3 sub(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
DEM_FN V3 scl(V3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
DEM_FN double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DEM_FN V3 cross(V3 a, V3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
DEM_FN double len(V3 a) { return sqrt(dot(a, a)); }
// normalized(): zero below 1e-14 (contact:67-71)
DEM_FN V3 unit(V3 a)
{
    const double l = len(a);
    return l > 1.0e-14 ? scl(a, 1.0 / l) : mk(0.0, 0.0, 0.0);
}
DEM_FN V3 zero3() { return mk(0.0, 0.0, 0.0); }

// world = R(q) body, R in the reference's rotateIv form (contact:105-112)
DEM_FN V3 rot(double qw, double qx, double qy, double qz, V3 v)
{
    return mk((1.0 - 2.0 * (qy * qy + qz * qz)) * v.x + 2.0 * (qx * qy - qw * qz) * v.y +
                  2.0 * (qx * qz + qw * qy) * v.z,
              2.0 * (qx * qy + qw * qz) * v.x + (1.0 - 2.0 * (qx * qx + qz * qz)) * v.y +
                  2.0 * (qy * qz - qw * qx) * v.z,
              2.0 * (qx * qz - qw * qy) * v.x + 2.0 * (qy * qz + qw * qx) * v.y +
                  (1.0 - 2.0 * (qx * qx + qy * qy)) * v.z);
}

// Per-material constants.  log(eps) is evaluated once on the host; the
// damping coefficient below is otherwise formed exactly as at contact:619-626.
struct Material {
    double epsS, mu, epsN, cohesion, beta, muT, muR, kN, dt;
    double logN, logS;  // log(epsN), log(epsS)
};

DEM_FN double damping(double l, double k, double m)
{
    return -2.0 * l * sqrt(k * m / (kPi * kPi + l * l));
}

// History slot values: [slot][roll(0), tangent(1), twist(2)][xyz]
DEM_FN V3 hget(const double* v, int cat)
{
    return mk(v[cat * 3], v[cat * 3 + 1], v[cat * 3 + 2]);
}
DEM_FN void hput(double* v, int cat, V3 a)
{
    v[cat * 3] = a.x;
    v[cat * 3 + 1] = a.y;
    v[cat * 3 + 2] = a.z;
}

// transportTangentHistory / transportAxialHistory (contact:245-265)
DEM_FN V3 transport_tangent(V3 prev, V3 n)
{
    const double pn = len(prev);
    const V3 proj = sub(prev, scl(n, dot(prev, n)));
    const double qn = len(proj);
    if (pn <= 1.0e-30 || qn <= 1.0e-30) return zero3();
    return scl(proj, pn / qn);
}
DEM_FN V3 transport_axial(V3 prev, V3 n)
{
    const double pn = len(prev);
    if (pn <= 1.0e-30) return zero3();
    const double al = dot(prev, n);
    if (fabs(al) <= 1.0e-30) return zero3();
    return scl(n, al < 0.0 ? -pn : pn);
}

// limitedHistoryResponseRates (contact:267-299): elastic trial from the
// history rate, damping from the response rate, Coulomb-type cap on the
// elastic trial, yield fallback rescaling the previous value.
DEM_FN V3 limited(V3 hrate, V3 rrate, double k, double c, double dt, V3 prev,
                                      double limit, double* slot, int cat, bool store,
                                      bool* yielded)
{
    const V3 delta = scl(hrate, dt);
    const V3 inc = scl(delta, k);
    const V3 et = add(inc, prev);
    const V3 trial = add(scl(rrate, c), et);
    const double tn = len(et);
    V3 r;
    const bool y = tn > limit;
    if (yielded) *yielded = y;
    if (y) {
        const double dn = len(delta);
        if (dn > 1.0e-14) {
            r = scl(et, limit / tn);
        } else {
            const double pn = len(prev);
            r = pn > 1.0e-14 ? scl(prev, limit / pn) : zero3();
        }
        if (store) hput(slot, cat, r);
    } else {
        r = trial;
        if (store) hput(slot, cat, add(prev, inc));
    }
    return r;
}

// wallHistoryResponse (contact:951-973)
DEM_FN V3 wall_limited(V3 rate, double k, double c, double dt, V3 prev,
                                           double limit, bool checkDamped, bool zeroAtZero,
                                           double* slot, int cat)
{
    const V3 delta = scl(rate, dt);
    const V3 et = add(prev, scl(delta, k));
    const V3 st = add(et, scl(rate, c));
    const double tn = len(checkDamped ? st : et);
    if (tn > limit) {
        const double dn = len(delta);
        const V3 lim = checkDamped ? st : et;
        const V3 r = dn > 1.0e-14 && tn > 1.0e-14 ? scl(lim, limit / tn)
                                                 : (zeroAtZero ? zero3() : prev);
        hput(slot, cat, r);
        return r;
    }
    hput(slot, cat, et);
    return st;
}

// One component as the contact law sees it.
struct Body {
    V3 pos;      // component centre
    V3 vel;      // component endpoint velocity
    V3 vh;       // component half-step velocity
    V3 off;      // world offset from the owner centre
    V3 w;        // owner angular velocity, world frame
    V3 wh;       // owner half-step angular velocity, world frame
    double mass, r;
    int status;
    uint32_t cid;
};

// particleContact (contact:579-706) after the overlap test has passed and
// the history slot has been found: the contact force on i (cf) and the term
// the owner adds to its torque sum (couple + offset x cf).
DEM_FN void pp_law(const Body& I, const Body& J, V3 delta, double d2,
                                       const Material& m, double* slot, V3& cf_out, V3& tq_out)
{
    const double ri = I.r, rj = J.r, rs = ri + rj;
    const double dist = sqrt(d2);
    const double depth = rs - dist;
    const V3 dir = scl(delta, 1.0 / dist);
    const V3 erv = sub(J.vel, I.vel);
    const bool midI = m.dt > 0.0 && I.status != -1;
    const bool midJ = m.dt > 0.0 && J.status != -1;
    const V3 mvi = midI ? I.vh : I.vel;
    const V3 mvj = midJ ? J.vh : J.vel;
    const V3 mrv = sub(mvj, mvi);
    const V3 mwi = midI ? I.wh : I.w;
    const V3 mwj = midJ ? J.wh : J.w;
    const double reduced = I.mass * J.mass / (I.mass + J.mass);
    const double kS = 2.0 * m.kN / 7.0;
    const double cN = damping(m.logN, m.kN, reduced);
    const double cS = damping(m.logS, kS, reduced);
    const double R = ri * rj / rs;
    const double brs = m.beta * R * m.beta * R;
    const double kT = 2.0 * kS * brs, cT = 2.0 * cS * brs, kR = m.kN * brs, cR = cN * brs;
    V3 fn = add(scl(dir, -m.kN * depth), scl(dir, cN * dot(erv, dir)));
    fn = add(fn, scl(dir, m.cohesion * 4.0 * brs));
    const double li = (ri * ri - rj * rj + d2) / (2.0 * dist);
    const double lj = dist - li;
    V3 etv = sub(erv, scl(dir, dot(erv, dir)));
    etv = add(etv, scl(cross(dir, I.w), li));
    etv = add(etv, scl(cross(dir, J.w), lj));
    V3 mtv = sub(mrv, scl(dir, dot(mrv, dir)));
    mtv = add(mtv, scl(cross(dir, mwi), li));
    mtv = add(mtv, scl(cross(dir, mwj), lj));
    const V3 oldT = transport_tangent(hget(slot, 1), dir);
    const V3 fr = limited(mtv, etv, kS, cS, m.dt, oldT, m.mu * len(fn), slot, 1, true, nullptr);
    const V3 erw = sub(J.w, I.w), mrw = sub(mwj, mwi);
    const V3 etw = scl(dir, dot(erw, dir)), mtw = scl(dir, dot(mrw, dir));
    const V3 erl = sub(erw, etw), mrl = sub(mrw, mtw);
    const double ms = m.beta * R * len(fn);
    const V3 tw = limited(mtw, etw, kT, cT, m.dt, transport_axial(hget(slot, 2), dir),
                          m.mu * m.muT * ms, slot, 2, true, nullptr);
    const V3 rl = limited(mrl, erl, kR, cR, m.dt, transport_tangent(hget(slot, 0), dir),
                          m.muR * ms, slot, 0, true, nullptr);
    const V3 cf = add(fn, fr);
    const V3 couple = add(add(tw, rl), scl(cross(dir, fr), li));
    cf_out = cf;
    tq_out = add(couple, cross(I.off, cf));
}

// wallContact (contact:975-1063) after the depth test and slot lookup.
// Accumulates the wall's force and torque contribution; returns the force.
DEM_FN V3 pw_law(const Body& I, V3 outward, double depth,
                                       const Material& m, double* slot, V3& fsum, V3& tsum)
{
    const V3 normal = scl(outward, -1.0);
    const double radius = I.r;
    const double lever = radius - depth;
    const V3 rv = scl(I.vel, -1.0);
    V3 tv = sub(rv, scl(normal, dot(rv, normal)));
    tv = add(tv, scl(cross(normal, I.w), lever));
    const V3 rw = scl(I.w, -1.0);
    const V3 twr = scl(normal, dot(rw, normal));
    V3 rlr = sub(rw, twr);
    if (fabs(lever) > 1.0e-30) rlr = add(rlr, scl(cross(normal, tv), 0.5 / lever));
    const double kS = 2.0 * m.kN / 7.0;
    const double cN = damping(m.logN, m.kN, I.mass);
    const double cS = damping(m.logS, kS, I.mass);
    const double brs = m.beta * radius * m.beta * radius;
    const double kT = 2.0 * kS * brs, cT = 2.0 * cS * brs, kR = m.kN * brs, cR = cN * brs;
    const V3 en = scl(normal, -m.kN * depth);
    const V3 dn = scl(normal, cN * dot(rv, normal));
    const V3 fn = add(en, dn);
    const double flim = m.mu * len(fn);
    const double ms = m.beta * radius * len(fn);
    const V3 fr = wall_limited(tv, kS, cS, m.dt, transport_tangent(hget(slot, 1), normal), flim,
                               false, false, slot, 1);
    const V3 tw = wall_limited(twr, kT, cT, m.dt, transport_axial(hget(slot, 2), normal),
                               m.mu * m.muT * ms, false, true, slot, 2);
    const V3 rl = wall_limited(rlr, kR, cR, m.dt, transport_tangent(hget(slot, 0), normal),
                               m.muR * ms, true, true, slot, 0);
    const V3 cf = add(fn, fr);
    const V3 couple = add(add(tw, rl), scl(cross(normal, fr), lever));
    fsum = add(fsum, cf);
    tsum = add(tsum, add(couple, cross(I.off, cf)));
    return cf;
}

// ---- arc tangent --------------------------------------------------------
// Built only from correctly rounded +, -, *, /, sqrt so the particle-triangle
// path evaluates identically on any IEEE-754 machine (the yield test of the
// read-only triangle pass compares a stored value of norm exactly mu|Fn|
// against mu|Fn|; a last-bit difference in atan2 would flip that branch).
// One half-angle reduction atan(t) = 2 atan(t / (1 + sqrt(1 + t^2))) takes
// t <= 1 below tan(pi/8); the Taylor series to z^43/43 follows (within 3 ulp).
DEM_FN double series_atan(double t)
{
    t = t / (1.0 + sqrt(1.0 + t * t));
    const double z = t * t;
    double p = -1.0 / 43.0;
    p = 1.0 / 41.0 + z * p;
    p = -1.0 / 39.0 + z * p;
    p = 1.0 / 37.0 + z * p;
    p = -1.0 / 35.0 + z * p;
    p = 1.0 / 33.0 + z * p;
    p = -1.0 / 31.0 + z * p;
    p = 1.0 / 29.0 + z * p;
    p = -1.0 / 27.0 + z * p;
    p = 1.0 / 25.0 + z * p;
    p = -1.0 / 23.0 + z * p;
    p = 1.0 / 21.0 + z * p;
    p = -1.0 / 19.0 + z * p;
    p = 1.0 / 17.0 + z * p;
    p = -1.0 / 15.0 + z * p;
    p = 1.0 / 13.0 + z * p;
    p = -1.0 / 11.0 + z * p;
    p = 1.0 / 9.0 + z * p;
    p = -1.0 / 7.0 + z * p;
    p = 1.0 / 5.0 + z * p;
    p = -1.0 / 3.0 + z * p;
    p = 1.0 + z * p;
    return (t * p) * 2.0;
}

DEM_FN double arc_tan2(double y, double x)
{
    const double kHalfPi = 1.5707963267948966192313216916398;
    if (isnan(x) || isnan(y)) return x + y;
    const double ax = fabs(x), ay = fabs(y);
    double a;
    if (ax == 0.0 && ay == 0.0) a = 0.0;
    else if (isinf(ax) && isinf(ay)) a = 0.5 * kHalfPi;
    else if (isinf(ax)) a = 0.0;
    else if (isinf(ay)) a = kHalfPi;
    else if (ay <= ax) a = series_atan(ay / ax);
    else a = kHalfPi - series_atan(ax / ay);
    if (signbit(x)) a = kPi - a;
    return signbit(y) ? -a : a;
}

// ---- sphere-section / triangle overlap area (contact:313-407) -----------
struct V2 {
    double x, y;
};
DEM_FN double cross2(V2 a, V2 b) { return a.x * b.y - a.y * b.x; }
DEM_FN double dot2(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }

DEM_FN double circle_edge_area(V2 a, V2 b, double radius)
{
    const V2 d{b.x - a.x, b.y - a.y};
    const double aa = dot2(d, d);
    double c0 = 0.0, c1 = 1.0, c2 = 0.0, c3 = 0.0;
    int n = 2;
    if (aa > 1.0e-30) {
        const double bb = 2.0 * dot2(a, d);
        const double cc = dot2(a, a) - radius * radius;
        const double disc = bb * bb - 4.0 * aa * cc;
        if (disc > 0.0) {
            const double root = sqrt(disc);
            const double t0 = (-bb - root) / (2.0 * aa);
            const double t1 = (-bb + root) / (2.0 * aa);
            if (t0 > 0.0 && t0 < 1.0) { if (n == 2) c2 = t0; else c3 = t0; ++n; }
            if (t1 > 0.0 && t1 < 1.0) { if (n == 2) c2 = t1; else c3 = t1; ++n; }
        }
    }
    double cuts[4] = {c0, c1, c2, c3};
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (cuts[j] < cuts[i]) {
                const double t = cuts[i];
                cuts[i] = cuts[j];
                cuts[j] = t;
            }
    double area = 0.0;
    for (int i = 0; i + 1 < n; ++i) {
        const double ta = cuts[i], tb = cuts[i + 1];
        const V2 u{a.x + d.x * ta, a.y + d.y * ta};
        const V2 v{a.x + d.x * tb, a.y + d.y * tb};
        const double h = 0.5 * (ta + tb);
        const V2 mid{a.x + d.x * h, a.y + d.y * h};
        if (dot2(mid, mid) <= radius * radius)
            area += 0.5 * cross2(u, v);
        else
            area += 0.5 * radius * radius * arc_tan2(cross2(u, v), dot2(u, v));
    }
    return area;
}

DEM_FN double circle_triangle_area(V3 center, V3 normal, V3 a, V3 b, V3 c,
                                              double radius)
{
    const V3 axis = fabs(normal.x) < 0.8 ? mk(1.0, 0.0, 0.0) : mk(0.0, 1.0, 0.0);
    const V3 u = unit(cross(normal, axis));
    const V3 v = cross(normal, u);
    const V3 ra = sub(a, center), rb = sub(b, center), rc = sub(c, center);
    const V2 pa{dot(ra, u), dot(ra, v)}, pb{dot(rb, u), dot(rb, v)}, pc{dot(rc, u), dot(rc, v)};
    double area = circle_edge_area(pa, pb, radius) + circle_edge_area(pb, pc, radius) +
                  circle_edge_area(pc, pa, radius);
    area = fabs(area);
    return fmin(area, kPi * radius * radius);
}

// World triangle: vertices v2, v3, v4 (a, b, c) and outward unit normal.
struct Tri {
    V3 a, b, c, n;
    int body;
};

// True unless the ball (c, r) stays clear of the axis-aligned box
// box[0..2]-box[3..5] by more than a relative 1e-6 plus 1e-12 m, margins
// many orders above the rounding of the coordinates involved.
DEM_FN bool ball_meets_box(V3 c, double r, const double* box)
{
    const double m = r * (1.0 + 1.0e-6) + 1.0e-12;
    const double ex = fmax(0.0, fmax(box[0] - c.x, c.x - box[3]));
    const double ey = fmax(0.0, fmax(box[1] - c.y, c.y - box[4]));
    const double ez = fmax(0.0, fmax(box[2] - c.z, c.z - box[5]));
    return ex * ex + ey * ey + ez * ez <= m * m;
}

// bodyTriangleAreaScale (contact:383-407).  With `box` (a bounding box of
// the triangle), a contact circle clear of the box returns 0 without the
// area evaluation: the circle then misses every edge and encloses no
// vertex, so the three signed arc terms of circle_triangle_area cancel to
// rounding (about 1e-16 of the circle's area), and the 1e-12 threshold below
// gives 0 as well.
DEM_FN double tri_area_scale(V3 pos, double r, const Tri& t, const double* box = nullptr)
{
    const double distance = dot(t.n, sub(pos, t.a));
    if (distance <= 0.0) return 0.0;
    const double crs = r * r - distance * distance;
    if (crs <= 1.0e-30) return 0.0;
    const double cr = sqrt(crs);
    const V3 cc = sub(pos, scl(t.n, distance));
    if (box && !ball_meets_box(cc, cr, box)) return 0.0;
    const double area = circle_triangle_area(cc, t.n, t.a, t.b, t.c, cr);
    if (area <= 0.0) return 0.0;
    const double s = fmin(1.0, area / (kPi * crs));
    return s > 1.0e-12 ? s : 0.0;
}

// Rigid body pose as the contact law reads it (13 doubles per body).
struct Pose {
    V3 c, v, w;  // centre, velocity, angular velocity (world)
    double qw, qx, qy, qz;
};

// Output of one particle-triangle contact (bodyContact, contact:708-853).
struct PtOut {
    V3 pf, ptq;               // on the particle
    V3 bnf, bfr, bnt, bft, btw, brl;  // on the body, the six ForceBody groups
    double depth;
    bool hit;
};

// bodyContact (contact:708-853) past its geometry: the triangle's outward
// normal tn, the plane distance, the contact centre cc and the area scale sc
// (> 1e-12) as pt_law forms them.
DEM_FN PtOut pt_contact(const Body& I, V3 tn, double distance, V3 cc, double sc, const Pose& P,
                        const Material& m, const double* slot)
{
    PtOut o;
    const double depth = I.r - distance;
    const V3 normal = scl(tn, -1.0);
    const V3 arm = sub(cc, P.c);
    const V3 bv = add(P.v, cross(P.w, arm));
    const V3 rv = sub(bv, I.vel);
    const double kS = 2.0 * m.kN / 7.0;
    const double cN = damping(m.logN, m.kN, I.mass);
    const double cS = damping(m.logS, kS, I.mass);
    const double radius = I.r;
    const double brs = m.beta * radius * m.beta * radius;
    const double kT = 2.0 * kS * brs, cT = 2.0 * cS * brs, kR = m.kN * brs, cR = cN * brs;
    V3 fn = add(scl(normal, -m.kN * depth), scl(normal, cN * dot(rv, normal)));
    fn = add(fn, scl(normal, m.cohesion * 4.0 * brs));
    V3 tv = sub(rv, scl(normal, dot(rv, normal)));
    tv = add(tv, scl(cross(normal, I.w), distance));
    double* none = nullptr;
    const V3 fr = limited(tv, tv, kS, cS, m.dt, transport_tangent(hget(slot, 1), normal),
                          m.mu * len(fn), none, 1, false, nullptr);
    const V3 rw = sub(P.w, I.w);
    const V3 twr = scl(normal, dot(rw, normal));
    const V3 rlr = add(sub(rw, twr), scl(cross(normal, tv), 0.5 / fmax(distance, 1.0e-14)));
    const double ms = m.beta * radius * len(fn);
    const V3 tw = limited(twr, twr, kT, cT, m.dt, transport_axial(hget(slot, 2), normal),
                          m.mu * m.muT * ms, none, 2, false, nullptr);
    const V3 rl = limited(rlr, rlr, kR, cR, m.dt, transport_tangent(hget(slot, 0), normal),
                          m.muR * ms, none, 0, false, nullptr);
    o.pf = scl(add(fn, fr), sc);
    const V3 moment = add(tw, rl);
    const V3 couple = scl(add(moment, scl(cross(normal, fr), distance)), sc);
    o.ptq = add(couple, cross(I.off, o.pf));
    o.bnf = scl(fn, -sc);
    o.bfr = scl(fr, -sc);
    o.bnt = scl(cross(arm, fn), -sc);
    o.bft = scl(cross(arm, fr), -sc);
    o.btw = scl(tw, -sc);
    o.brl = scl(rl, -sc);
    o.depth = depth;
    o.hit = true;
    return o;
}

// bodyContact with the manifold history read-only (dt = 0 in m).
DEM_FN PtOut pt_law(const Body& I, const Tri& t, const Pose& P, const Material& m,
                               const double* slot)
{
    PtOut o;
    o.hit = false;
    const double distance = dot(t.n, sub(I.pos, t.a));
    if (distance <= 0.0) return o;
    const double depth = I.r - distance;
    if (depth <= 0.0) return o;
    const double crs = I.r * I.r - distance * distance;
    if (crs <= 1.0e-30) return o;
    const double cr = sqrt(crs);
    const V3 cc = sub(I.pos, scl(t.n, distance));
    const double area = circle_triangle_area(cc, t.n, t.a, t.b, t.c, cr);
    if (area <= 0.0) return o;
    const double sc = fmin(1.0, area / (kPi * crs));
    if (sc <= 1.0e-12) return o;
    return pt_contact(I, t.n, distance, cc, sc, P, m, slot);
}

// advanceBodyManifoldHistory (contact:855-949) after the slot is found.
DEM_FN void manifold_law(const Body& I, const Pose& P, const Material& m, V3 normal,
                                    V3 cc, double distance, double* slot, bool* slide,
                                    bool* twisty)
{
    const V3 arm = sub(cc, P.c);
    const bool dyn = m.dt > 0.0 && I.status != -1;
    const V3 mw = dyn ? I.wh : I.w;
    const V3 bv = add(P.v, cross(P.w, arm));
    const V3 mpv = dyn ? I.vh : I.vel;
    const V3 erv = sub(bv, I.vel), mrv = sub(bv, mpv);
    const double kS = 2.0 * m.kN / 7.0;
    const double cN = damping(m.logN, m.kN, I.mass);
    const double cS = damping(m.logS, kS, I.mass);
    const double radius = I.r;
    const double brs = m.beta * radius * m.beta * radius;
    const double kT = 2.0 * kS * brs, cT = 2.0 * cS * brs, kR = m.kN * brs, cR = cN * brs;
    const double depth = radius - distance;
    V3 fn = add(scl(normal, -m.kN * depth), scl(normal, cN * dot(erv, normal)));
    fn = add(fn, scl(normal, m.cohesion * 4.0 * brs));
    V3 tv = sub(mrv, scl(normal, dot(mrv, normal)));
    tv = add(tv, scl(cross(normal, mw), distance));
    limited(tv, tv, kS, cS, m.dt, transport_tangent(hget(slot, 1), normal), m.mu * len(fn), slot,
            1, true, slide);
    const V3 rw = sub(P.w, mw);
    const V3 twr = scl(normal, dot(rw, normal));
    const V3 rlr = add(sub(rw, twr), scl(cross(normal, tv), 0.5 / fmax(distance, 1.0e-14)));
    const double ms = m.beta * radius * len(fn);
    limited(twr, twr, kT, cT, m.dt, transport_axial(hget(slot, 2), normal), m.mu * m.muT * ms,
            slot, 2, true, twisty);
    limited(rlr, rlr, kR, cR, m.dt, transport_tangent(hget(slot, 0), normal), m.muR * ms, slot, 0,
            true, nullptr);
}

// quatDerivative / normalizeQuat (particle:149-183)
DEM_FN void quat_dot(double qw, double qx, double qy, double qz, double wx,
                                         double wy, double wz, double& dw, double& dx, double& dy,
                                         double& dz)
{
    dw = 0.5 * (-qx * wx - qy * wy - qz * wz);
    dx = 0.5 * (qw * wx + qy * wz - qz * wy);
    dy = 0.5 * (qw * wy - qx * wz + qz * wx);
    dz = 0.5 * (qw * wz + qx * wy - qy * wx);
}
DEM_FN void quat_unit(double& qw, double& qx, double& qy, double& qz)
{
    const double n2 = qw * qw + qx * qx + qy * qy + qz * qz;
    if (n2 <= 1.0e-24) {
        qw = 1.0; qx = 0.0; qy = 0.0; qz = 0.0;
        return;
    }
    const double inv = 1.0 / sqrt(n2);
    qw *= inv; qx *= inv; qy *= inv; qz *= inv;
}

}  // namespace dem
