// dem_adjoint.hpp -- reverse-mode derivatives (vector-Jacobian products) of
// the contact laws and the leapfrog arithmetic in dem_device.hpp.
//
// Each *_vjp function recomputes the forward intermediates of its forward
// counterpart from the same inputs (so every branch -- contact on/off is
// fixed by the caller, yield / no yield, the 1e-14 and 1e-30 guards -- is
// taken exactly as in the forward step) and then propagates the output
// adjoints back to the inputs, accumulating (+=) into the input adjoints.
// Inside a branch the law is smooth; the derivative of a branch switch
// itself is zero, as for any piecewise-smooth map.
#pragma once

#include "dem_device.hpp"

namespace dem {

// Adjoint of the material constants a law reads.  epsN and epsS enter
// through their logarithms (logN, logS); kS = 2 kN / 7 is folded into kN.
struct MatGrad {
    double mu, logN, logS, kN;
};

DEM_FN void acc(V3& a, V3 b) { a = add(a, b); }

// y = |a|
DEM_FN void len_vjp(V3 a, double y, double yb, V3& ab)
{
    if (y > 0.0) acc(ab, scl(a, yb / y));
}
// s = a . b
DEM_FN void dot_vjp(V3 a, V3 b, double sb, V3& ab, V3& bb)
{
    acc(ab, scl(b, sb));
    acc(bb, scl(a, sb));
}
// c = a x b
DEM_FN void cross_vjp(V3 a, V3 b, V3 cb, V3& ab, V3& bb)
{
    acc(ab, cross(b, cb));
    acc(bb, cross(cb, a));
}

// damping(l, k, m) = -2 l sqrt(k m / (pi^2 + l^2)), m constant
DEM_FN void damping_vjp(double l, double k, double m, double cb, double& lb, double& kb)
{
    const double D = kPi * kPi + l * l;
    const double s = sqrt(k * m / D);
    if (s > 0.0) {
        kb += cb * (-l * m / (D * s));
        lb += cb * (-2.0 * s + 2.0 * l * l * k * m / (D * D * s));
    } else {
        lb += cb * (-2.0 * s);
    }
}

// out = transport_tangent(prev, n)
DEM_FN void transport_tangent_vjp(V3 prev, V3 n, V3 ob, V3& prevb, V3& nb, bool wantN)
{
    const double pn = len(prev);
    const double d = dot(prev, n);
    const V3 proj = sub(prev, scl(n, d));
    const double qn = len(proj);
    if (pn <= 1.0e-30 || qn <= 1.0e-30) return;
    const double r = pn / qn;
    V3 projb = scl(ob, r);
    const double rb = dot(proj, ob);
    const double pnb = rb / qn;
    const double qnb = -rb * pn / (qn * qn);
    acc(projb, scl(proj, qnb / qn));
    acc(prevb, scl(prev, pnb / pn));
    acc(prevb, projb);
    const double db = -dot(n, projb);
    acc(prevb, scl(n, db));
    if (wantN) {
        acc(nb, scl(projb, -d));
        acc(nb, scl(prev, db));
    }
}

// out = transport_axial(prev, n)
DEM_FN void transport_axial_vjp(V3 prev, V3 n, V3 ob, V3& prevb, V3& nb, bool wantN)
{
    const double pn = len(prev);
    if (pn <= 1.0e-30) return;
    const double al = dot(prev, n);
    if (fabs(al) <= 1.0e-30) return;
    const double sg = al < 0.0 ? -1.0 : 1.0;
    if (wantN) acc(nb, scl(ob, sg * pn));
    const double pnb = sg * dot(n, ob);
    acc(prevb, scl(prev, pnb / pn));
}

// (r, stored) = limited(hrate, rrate, k, c, dt, prev, limit); adjoints rb
// (returned response) and hb (stored history value).
DEM_FN void limited_vjp(V3 hrate, V3 rrate, double k, double c, double dt, V3 prev, double limit,
                        V3 rb, V3 hb, V3& hrateb, V3& rrateb, double& kb, double& cb, V3& prevb,
                        double& limitb)
{
    const V3 delta = scl(hrate, dt);
    const V3 inc = scl(delta, k);
    const V3 et = add(inc, prev);
    const double tn = len(et);
    V3 incb = zero3();
    if (tn > limit) {
        const V3 ub = add(rb, hb);
        const double dn = len(delta);
        if (dn > 1.0e-14) {
            const double s = limit / tn;
            V3 etb = scl(ub, s);
            const double sb = dot(et, ub);
            limitb += sb / tn;
            const double tnb = -sb * limit / (tn * tn);
            acc(etb, scl(et, tnb / tn));
            acc(incb, etb);
            acc(prevb, etb);
        } else {
            const double pn = len(prev);
            if (pn > 1.0e-14) {
                const double s = limit / pn;
                acc(prevb, scl(ub, s));
                const double sb = dot(prev, ub);
                limitb += sb / pn;
                const double pnb = -sb * limit / (pn * pn);
                acc(prevb, scl(prev, pnb / pn));
            }
        }
    } else {
        // r = rrate c + inc + prev, stored = prev + inc
        acc(rrateb, scl(rb, c));
        cb += dot(rrate, rb);
        acc(incb, add(rb, hb));
        acc(prevb, add(rb, hb));
    }
    // inc = hrate dt k
    acc(hrateb, scl(incb, dt * k));
    kb += dot(delta, incb);
}

// (r, stored) = wall_limited(rate, k, c, dt, prev, limit, checkDamped, zeroAtZero)
DEM_FN void wall_limited_vjp(V3 rate, double k, double c, double dt, V3 prev, double limit,
                             bool checkDamped, bool zeroAtZero, V3 rb, V3 hb, V3& rateb, double& kb,
                             double& cb, V3& prevb, double& limitb)
{
    const V3 delta = scl(rate, dt);
    const V3 et = add(prev, scl(delta, k));
    const V3 st = add(et, scl(rate, c));
    const V3 lim = checkDamped ? st : et;
    const double tn = len(lim);
    V3 etb = zero3(), stb = zero3();
    if (tn > limit) {
        const V3 ub = add(rb, hb);
        const double dn = len(delta);
        if (dn > 1.0e-14 && tn > 1.0e-14) {
            const double s = limit / tn;
            V3 limb = scl(ub, s);
            const double sb = dot(lim, ub);
            limitb += sb / tn;
            const double tnb = -sb * limit / (tn * tn);
            acc(limb, scl(lim, tnb / tn));
            if (checkDamped) acc(stb, limb);
            else acc(etb, limb);
        } else if (!zeroAtZero) {
            acc(prevb, ub);
        }
    } else {
        acc(stb, rb);   // returned st
        acc(etb, hb);   // stored et
    }
    // st = et + rate c
    acc(etb, stb);
    acc(rateb, scl(stb, c));
    cb += dot(rate, stb);
    // et = prev + rate dt k
    acc(prevb, etb);
    acc(rateb, scl(etb, dt * k));
    kb += dot(delta, etb);
}

// Adjoint of one side of a particle-particle contact.
struct BodyGrad {
    V3 pos, vel, vh, w, wh;
};

DEM_FN BodyGrad zero_body_grad()
{
    BodyGrad g;
    g.pos = zero3(); g.vel = zero3(); g.vh = zero3(); g.w = zero3(); g.wh = zero3();
    return g;
}

// Reverse of pp_law (dem_device.hpp).  prev[9] holds the slot values the
// forward law read; cfb, tqb are the adjoints of the force on i and of its
// torque term, hb[9] the adjoint of the values the law stored.  Accumulates
// into Ib, Jb, prevb[9] and g.
DEM_FN void pp_law_vjp(const Body& I, const Body& J, const Material& m, const double* prev, V3 cfb,
                       V3 tqb, const double* hb, BodyGrad& Ib, BodyGrad& Jb, double* prevb, MatGrad& g)
{
    // ---- forward recomputation
    const V3 delta = sub(J.pos, I.pos);
    const double d2 = dot(delta, delta);
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
    const double vn = dot(erv, dir);
    const double A = -m.kN * depth + cN * vn + m.cohesion * 4.0 * brs;
    const V3 fn = scl(dir, A);
    const double fnn = len(fn);
    const double li = (ri * ri - rj * rj + d2) / (2.0 * dist);
    const double lj = dist - li;
    const double ev = dot(erv, dir);
    V3 etv = sub(erv, scl(dir, ev));
    etv = add(etv, scl(cross(dir, I.w), li));
    etv = add(etv, scl(cross(dir, J.w), lj));
    const double mv = dot(mrv, dir);
    V3 mtv = sub(mrv, scl(dir, mv));
    mtv = add(mtv, scl(cross(dir, mwi), li));
    mtv = add(mtv, scl(cross(dir, mwj), lj));
    const V3 pR = mk(prev[0], prev[1], prev[2]);
    const V3 pT = mk(prev[3], prev[4], prev[5]);
    const V3 pW = mk(prev[6], prev[7], prev[8]);
    const V3 oldT = transport_tangent(pT, dir);
    const double limT = m.mu * fnn;
    double scratch[9];
    const V3 fr = limited(mtv, etv, kS, cS, m.dt, oldT, limT, scratch, 1, false, nullptr);
    const V3 erw = sub(J.w, I.w), mrw = sub(mwj, mwi);
    const double ew = dot(erw, dir), mw = dot(mrw, dir);
    const V3 etw = scl(dir, ew), mtw = scl(dir, mw);
    const V3 erl = sub(erw, etw), mrl = sub(mrw, mtw);
    const double ms = m.beta * R * fnn;
    const V3 oldW = transport_axial(pW, dir);
    const double limW = m.mu * m.muT * ms;
    const V3 oldR = transport_tangent(pR, dir);
    const double limR = m.muR * ms;
    (void)scratch;

    // ---- reverse
    V3 dirb = zero3(), frb = zero3(), fnb = zero3(), twb, rlb, mtvb = zero3(), etvb = zero3();
    double lib = 0.0, ljb = 0.0, fnnb = 0.0, msb = 0.0;
    double kSb = 0.0, cSb = 0.0, kTb = 0.0, cTb = 0.0, kRb = 0.0, cRb = 0.0, cNb = 0.0;
    // tq = couple + off x cf
    V3 cfb2 = cfb;
    acc(cfb2, cross(tqb, I.off));
    const V3 coupleb = tqb;
    // couple = tw + rl + (dir x fr) li
    twb = coupleb;
    rlb = coupleb;
    {
        const V3 x = cross(dir, fr);
        const V3 xb = scl(coupleb, li);
        lib += dot(x, coupleb);
        cross_vjp(dir, fr, xb, dirb, frb);
    }
    // cf = fn + fr
    acc(fnb, cfb2);
    acc(frb, cfb2);
    // rl = limited(mrl, erl, kR, cR, dt, oldR, limR)
    V3 mrlb = zero3(), erlb = zero3(), oldRb = zero3();
    double limRb = 0.0;
    limited_vjp(mrl, erl, kR, cR, m.dt, oldR, limR, rlb, mk(hb[0], hb[1], hb[2]), mrlb, erlb, kRb, cRb,
                oldRb, limRb);
    msb += m.muR * limRb;
    // tw = limited(mtw, etw, kT, cT, dt, oldW, limW)
    V3 mtwb = zero3(), etwb = zero3(), oldWb = zero3();
    double limWb = 0.0;
    limited_vjp(mtw, etw, kT, cT, m.dt, oldW, limW, twb, mk(hb[6], hb[7], hb[8]), mtwb, etwb, kTb, cTb,
                oldWb, limWb);
    g.mu += m.muT * ms * limWb;
    msb += m.mu * m.muT * limWb;
    // fr = limited(mtv, etv, kS, cS, dt, oldT, limT)
    V3 oldTb = zero3();
    double limTb = 0.0;
    limited_vjp(mtv, etv, kS, cS, m.dt, oldT, limT, frb, mk(hb[3], hb[4], hb[5]), mtvb, etvb, kSb, cSb,
                oldTb, limTb);
    g.mu += fnn * limTb;
    fnnb += m.mu * limTb;
    // transports of the stored values
    V3 pRb = zero3(), pTb = zero3(), pWb = zero3();
    transport_tangent_vjp(pR, dir, oldRb, pRb, dirb, true);
    transport_axial_vjp(pW, dir, oldWb, pWb, dirb, true);
    transport_tangent_vjp(pT, dir, oldTb, pTb, dirb, true);
    prevb[0] += pRb.x; prevb[1] += pRb.y; prevb[2] += pRb.z;
    prevb[3] += pTb.x; prevb[4] += pTb.y; prevb[5] += pTb.z;
    prevb[6] += pWb.x; prevb[7] += pWb.y; prevb[8] += pWb.z;
    // ms = beta R |fn|
    fnnb += m.beta * R * msb;
    // erl = erw - etw, mrl = mrw - mtw
    V3 erwb = erlb, mrwb = mrlb;
    acc(etwb, scl(erlb, -1.0));
    acc(mtwb, scl(mrlb, -1.0));
    // etw = dir (erw . dir), mtw = dir (mrw . dir)
    {
        acc(dirb, scl(etwb, ew));
        const double sb = dot(dir, etwb);
        dot_vjp(erw, dir, sb, erwb, dirb);
        acc(dirb, scl(mtwb, mw));
        const double tb = dot(dir, mtwb);
        dot_vjp(mrw, dir, tb, mrwb, dirb);
    }
    // erw = J.w - I.w, mrw = mwj - mwi
    V3 mwib = zero3(), mwjb = zero3();
    acc(Jb.w, erwb);
    acc(Ib.w, scl(erwb, -1.0));
    acc(mwjb, mrwb);
    acc(mwib, scl(mrwb, -1.0));
    // mtv = mrv - dir (mrv . dir) + (dir x mwi) li + (dir x mwj) lj
    V3 mrvb = mtvb;
    {
        acc(dirb, scl(mtvb, -mv));
        const double sb = -dot(dir, mtvb);
        dot_vjp(mrv, dir, sb, mrvb, dirb);
        const V3 y = cross(dir, mwi);
        lib += dot(y, mtvb);
        cross_vjp(dir, mwi, scl(mtvb, li), dirb, mwib);
        const V3 z = cross(dir, mwj);
        ljb += dot(z, mtvb);
        cross_vjp(dir, mwj, scl(mtvb, lj), dirb, mwjb);
    }
    // etv = erv - dir (erv . dir) + (dir x I.w) li + (dir x J.w) lj
    V3 ervb = etvb;
    {
        acc(dirb, scl(etvb, -ev));
        const double sb = -dot(dir, etvb);
        dot_vjp(erv, dir, sb, ervb, dirb);
        const V3 y = cross(dir, I.w);
        lib += dot(y, etvb);
        cross_vjp(dir, I.w, scl(etvb, li), dirb, Ib.w);
        const V3 z = cross(dir, J.w);
        ljb += dot(z, etvb);
        cross_vjp(dir, J.w, scl(etvb, lj), dirb, Jb.w);
    }
    // lj = dist - li ; li = (ri^2 - rj^2 + d2) / (2 dist)
    double distb = ljb;
    lib -= ljb;
    double d2b = lib / (2.0 * dist);
    distb += -lib * li / dist;
    // fn = dir A ; A = -kN depth + cN vn + coh 4 brs ; |fn|
    len_vjp(fn, fnn, fnnb, fnb);
    acc(dirb, scl(fnb, A));
    const double Ab = dot(dir, fnb);
    g.kN += -depth * Ab;
    const double depthb = -m.kN * Ab;
    cNb += vn * Ab;
    const double vnb = cN * Ab;
    dot_vjp(erv, dir, vnb, ervb, dirb);
    // kT = 2 kS brs, cT = 2 cS brs, kR = kN brs, cR = cN brs
    kSb += 2.0 * brs * kTb;
    cSb += 2.0 * brs * cTb;
    g.kN += brs * kRb;
    cNb += brs * cRb;
    // cS = damping(logS, kS, reduced), cN = damping(logN, kN, reduced)
    damping_vjp(m.logS, kS, reduced, cSb, g.logS, kSb);
    damping_vjp(m.logN, m.kN, reduced, cNb, g.logN, g.kN);
    g.kN += 2.0 / 7.0 * kSb;
    // mid-rate selections
    if (midI) acc(Ib.wh, mwib); else acc(Ib.w, mwib);
    if (midJ) acc(Jb.wh, mwjb); else acc(Jb.w, mwjb);
    if (midJ) acc(Jb.vh, mrvb); else acc(Jb.vel, mrvb);
    if (midI) acc(Ib.vh, scl(mrvb, -1.0)); else acc(Ib.vel, scl(mrvb, -1.0));
    // erv = J.vel - I.vel
    acc(Jb.vel, ervb);
    acc(Ib.vel, scl(ervb, -1.0));
    // dir = delta / dist ; depth = rs - dist ; dist = sqrt(d2) ; d2 = delta . delta
    V3 deltab = scl(dirb, 1.0 / dist);
    distb += -dot(delta, dirb) / (dist * dist);
    distb -= depthb;
    d2b += distb / (2.0 * dist);
    acc(deltab, scl(delta, 2.0 * d2b));
    acc(Jb.pos, deltab);
    acc(Ib.pos, scl(deltab, -1.0));
}

// Reverse of pw_law for one wall: cfb, tqb adjoints of the wall's force and
// torque contributions, hb[9] of the stored values.
DEM_FN void pw_law_vjp(const Body& I, V3 outward, V3 origin, const Material& m, const double* prev,
                       V3 cfb, V3 tqb, const double* hb, BodyGrad& Ib, double* prevb, MatGrad& g)
{
    const V3 normal = scl(outward, -1.0);
    const double radius = I.r;
    const double distance = dot(sub(I.pos, origin), outward);
    const double depth = radius - distance;
    const double lever = radius - depth;
    const V3 rv = scl(I.vel, -1.0);
    const double rvn = dot(rv, normal);
    V3 tv = sub(rv, scl(normal, rvn));
    const V3 cw = cross(normal, I.w);
    tv = add(tv, scl(cw, lever));
    const V3 rw = scl(I.w, -1.0);
    const double rwn = dot(rw, normal);
    const V3 twr = scl(normal, rwn);
    V3 rlr = sub(rw, twr);
    const bool useLever = fabs(lever) > 1.0e-30;
    const V3 ct = cross(normal, tv);
    if (useLever) rlr = add(rlr, scl(ct, 0.5 / lever));
    const double kS = 2.0 * m.kN / 7.0;
    const double cN = damping(m.logN, m.kN, I.mass);
    const double cS = damping(m.logS, kS, I.mass);
    const double brs = m.beta * radius * m.beta * radius;
    const double kT = 2.0 * kS * brs, cT = 2.0 * cS * brs, kR = m.kN * brs, cR = cN * brs;
    const double A = -m.kN * depth + cN * rvn;
    const V3 fn = scl(normal, A);
    const double fnn = len(fn);
    const double flim = m.mu * fnn;
    const double ms = m.beta * radius * fnn;
    const V3 pR = mk(prev[0], prev[1], prev[2]);
    const V3 pT = mk(prev[3], prev[4], prev[5]);
    const V3 pW = mk(prev[6], prev[7], prev[8]);
    const V3 oldT = transport_tangent(pT, normal);
    const V3 oldW = transport_axial(pW, normal);
    const V3 oldR = transport_tangent(pR, normal);
    double scratch[9];
    const V3 fr = wall_limited(tv, kS, cS, m.dt, oldT, flim, false, false, scratch, 1);
    (void)scratch;

    V3 unused = zero3();
    // T += couple + off x cf ; couple = tw + rl + (normal x fr) lever
    V3 cfb2 = cfb;
    acc(cfb2, cross(tqb, I.off));
    const V3 twb = tqb, rlb = tqb;
    V3 frb = zero3();
    cross_vjp(normal, fr, scl(tqb, lever), unused, frb);
    double leverb = dot(cross(normal, fr), tqb);
    // cf = fn + fr
    V3 fnb = cfb2;
    acc(frb, cfb2);
    double kSb = 0.0, cSb = 0.0, kTb = 0.0, cTb = 0.0, kRb = 0.0, cRb = 0.0, cNb = 0.0;
    double fnnb = 0.0, msb = 0.0;
    // rl = wall_limited(rlr, kR, cR, dt, oldR, muR ms, checkDamped, zeroAtZero)
    V3 rlrb = zero3(), oldRb = zero3();
    double limRb = 0.0;
    wall_limited_vjp(rlr, kR, cR, m.dt, oldR, m.muR * ms, true, true, rlb, mk(hb[0], hb[1], hb[2]), rlrb,
                     kRb, cRb, oldRb, limRb);
    msb += m.muR * limRb;
    // tw = wall_limited(twr, kT, cT, dt, oldW, mu muT ms, false, true)
    V3 twrb = zero3(), oldWb = zero3();
    double limWb = 0.0;
    wall_limited_vjp(twr, kT, cT, m.dt, oldW, m.mu * m.muT * ms, false, true, twb, mk(hb[6], hb[7], hb[8]),
                     twrb, kTb, cTb, oldWb, limWb);
    g.mu += m.muT * ms * limWb;
    msb += m.mu * m.muT * limWb;
    // fr = wall_limited(tv, kS, cS, dt, oldT, flim, false, false)
    V3 tvb = zero3(), oldTb = zero3();
    double flimb = 0.0;
    wall_limited_vjp(tv, kS, cS, m.dt, oldT, flim, false, false, frb, mk(hb[3], hb[4], hb[5]), tvb, kSb,
                     cSb, oldTb, flimb);
    g.mu += fnn * flimb;
    fnnb += m.mu * flimb;
    V3 pRb = zero3(), pTb = zero3(), pWb = zero3();
    transport_tangent_vjp(pR, normal, oldRb, pRb, unused, false);
    transport_axial_vjp(pW, normal, oldWb, pWb, unused, false);
    transport_tangent_vjp(pT, normal, oldTb, pTb, unused, false);
    prevb[0] += pRb.x; prevb[1] += pRb.y; prevb[2] += pRb.z;
    prevb[3] += pTb.x; prevb[4] += pTb.y; prevb[5] += pTb.z;
    prevb[6] += pWb.x; prevb[7] += pWb.y; prevb[8] += pWb.z;
    // ms = beta r |fn| ; flim handled ; |fn|
    fnnb += m.beta * radius * msb;
    len_vjp(fn, fnn, fnnb, fnb);
    // fn = normal A ; A = -kN depth + cN rvn
    const double Ab = dot(normal, fnb);
    g.kN += -depth * Ab;
    double depthb = -m.kN * Ab;
    cNb += rvn * Ab;
    double rvnb = cN * Ab;
    // kT, cT, kR, cR
    kSb += 2.0 * brs * kTb;
    cSb += 2.0 * brs * cTb;
    g.kN += brs * kRb;
    cNb += brs * cRb;
    damping_vjp(m.logS, kS, I.mass, cSb, g.logS, kSb);
    damping_vjp(m.logN, m.kN, I.mass, cNb, g.logN, g.kN);
    g.kN += 2.0 / 7.0 * kSb;
    // rlr = rw - twr + (normal x tv) (0.5 / lever)
    V3 rwb = rlrb;
    acc(twrb, scl(rlrb, -1.0));
    if (useLever) {
        cross_vjp(normal, tv, scl(rlrb, 0.5 / lever), unused, tvb);
        leverb += dot(ct, rlrb) * (-0.5 / (lever * lever));
    }
    // twr = normal (rw . normal)
    rwb = add(rwb, scl(normal, dot(normal, twrb)));
    // rw = -I.w
    acc(Ib.w, scl(rwb, -1.0));
    // tv = rv - normal (rv . normal) + (normal x I.w) lever
    V3 rvb = tvb;
    rvnb += -dot(normal, tvb);
    leverb += dot(cw, tvb);
    cross_vjp(normal, I.w, scl(tvb, lever), unused, Ib.w);
    acc(rvb, scl(normal, rvnb));
    acc(Ib.vel, scl(rvb, -1.0));
    // lever = radius - depth ; depth = radius - (pos - origin) . outward
    depthb -= leverb;
    acc(Ib.pos, scl(outward, -depthb));
}

// ---- rotations and the rigid-body update ---------------------------------
// w = rot(q, v): adjoints of q (4) and v
DEM_FN void rot_vjp(double qw, double qx, double qy, double qz, V3 v, V3 wb, double* qb, V3& vb)
{
    // v-adjoint = R^T wb
    vb.x += (1.0 - 2.0 * (qy * qy + qz * qz)) * wb.x + 2.0 * (qx * qy + qw * qz) * wb.y +
            2.0 * (qx * qz - qw * qy) * wb.z;
    vb.y += 2.0 * (qx * qy - qw * qz) * wb.x + (1.0 - 2.0 * (qx * qx + qz * qz)) * wb.y +
            2.0 * (qy * qz + qw * qx) * wb.z;
    vb.z += 2.0 * (qx * qz + qw * qy) * wb.x + 2.0 * (qy * qz - qw * qx) * wb.y +
            (1.0 - 2.0 * (qx * qx + qy * qy)) * wb.z;
    const double vx = v.x, vy = v.y, vz = v.z;
    qb[0] += wb.x * (-2.0 * qz * vy + 2.0 * qy * vz) + wb.y * (2.0 * qz * vx - 2.0 * qx * vz) +
             wb.z * (-2.0 * qy * vx + 2.0 * qx * vy);
    qb[1] += wb.x * (2.0 * qy * vy + 2.0 * qz * vz) + wb.y * (2.0 * qy * vx - 4.0 * qx * vy - 2.0 * qw * vz) +
             wb.z * (2.0 * qz * vx + 2.0 * qw * vy - 4.0 * qx * vz);
    qb[2] += wb.x * (-4.0 * qy * vx + 2.0 * qx * vy + 2.0 * qw * vz) + wb.y * (2.0 * qx * vx + 2.0 * qz * vz) +
             wb.z * (-2.0 * qw * vx + 2.0 * qz * vy - 4.0 * qy * vz);
    qb[3] += wb.x * (-4.0 * qz * vx - 2.0 * qw * vy + 2.0 * qx * vz) +
             wb.y * (2.0 * qw * vx - 4.0 * qz * vy + 2.0 * qy * vz) + wb.z * (2.0 * qx * vx + 2.0 * qy * vy);
}

// (dw, dx, dy, dz) = quat_dot(q, w)
DEM_FN void quat_dot_vjp(const double* q, V3 w, const double* db, double* qb, V3& wb)
{
    const double qw = q[0], qx = q[1], qy = q[2], qz = q[3];
    qb[0] += 0.5 * (w.x * db[1] + w.y * db[2] + w.z * db[3]);
    qb[1] += 0.5 * (-w.x * db[0] - w.z * db[2] + w.y * db[3]);
    qb[2] += 0.5 * (-w.y * db[0] + w.z * db[1] - w.x * db[3]);
    qb[3] += 0.5 * (-w.z * db[0] - w.y * db[1] + w.x * db[2]);
    wb.x += 0.5 * (-qx * db[0] + qw * db[1] + qz * db[2] - qy * db[3]);
    wb.y += 0.5 * (-qy * db[0] - qz * db[1] + qw * db[2] + qx * db[3]);
    wb.z += 0.5 * (-qz * db[0] + qy * db[1] - qx * db[2] + qw * db[3]);
}

// u = quat_unit(p) (identity below |p|^2 = 1e-24: zero derivative)
DEM_FN void quat_unit_vjp(const double* p, const double* ub, double* pb)
{
    const double n2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2] + p[3] * p[3];
    if (n2 <= 1.0e-24) return;
    const double n = sqrt(n2);
    double u[4], s = 0.0;
    for (int k = 0; k < 4; ++k) {
        u[k] = p[k] / n;
        s += u[k] * ub[k];
    }
    for (int k = 0; k < 4; ++k) pb[k] += (ub[k] - u[k] * s) / n;
}

// alpha = Euler(R(q)^T T, w, I) as in the acceleration kernel: adjoints of
// the world torque T, q and the body angular velocity w.
DEM_FN void euler_vjp(const double* q, V3 T, V3 w, V3 inertia, V3 ab, V3& Tb, double* qb, V3& wb)
{
    const double ix = inertia.x, iy = inertia.y, iz = inertia.z;
    // ax = (tx - (iz - iy) wz wy) / ix, ...
    const V3 tb = mk(ab.x / ix, ab.y / iy, ab.z / iz);
    wb.z += -tb.x * (iz - iy) * w.y;
    wb.y += -tb.x * (iz - iy) * w.z;
    wb.x += -tb.y * (ix - iz) * w.z;
    wb.z += -tb.y * (ix - iz) * w.x;
    wb.y += -tb.z * (iy - ix) * w.x;
    wb.x += -tb.z * (iy - ix) * w.y;
    // t = R(q)^T T = rot(conj(q), T)
    double qcb[4] = {0.0, 0.0, 0.0, 0.0};
    rot_vjp(q[0], -q[1], -q[2], -q[3], T, tb, qcb, Tb);
    qb[0] += qcb[0];
    qb[1] -= qcb[1];
    qb[2] -= qcb[2];
    qb[3] -= qcb[3];
}

}  // namespace dem
