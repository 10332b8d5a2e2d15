// phx_adjoint_numerics.hpp -- reverse-mode (adjoint) counterparts of phx_numerics.hpp.
//
// Each function takes the primal inputs, recomputes the primal intermediates it needs, and
// propagates an adjoint seed backwards through the statements of the primal in reverse order
// (hand-derived, statement by statement).  Non-smooth operations follow the branch the primal took:
// |x| -> sign(x), max/min -> the selected operand, the Harten fix -> the corrected or uncorrected
// formula, the limiter -> the active face of the min and the active branch of the slope.
// Adjoint arithmetic does not need to reproduce the primal's rounding; it is checked against
// central finite differences of the primal.
#pragma once
#include "phx_numerics.hpp"

namespace phx {

struct PrimBar {
    double r, u, v, p;
};

__host__ __device__ inline double sgn(double x) { return (x > 0.0) ? 1.0 : ((x < 0.0) ? -1.0 : 0.0); }

// Adjoint of cons_to_prim: given the adjoint of W = (rho, u, v, p), the adjoint of q.
__host__ __device__ inline Q4 cons_to_prim_vjp(const Q4& q, const PrimBar& b, const Phys& ph) {
    const double r = q.v[0];
    const double u = q.v[1] / r, v = q.v[2] / r;
    // p = gm1*(q3 - ek), ek = r*(0.5*(u*u+v*v))
    const double bek = -ph.gm1 * b.p;
    double bq3 = ph.gm1 * b.p;
    double br = b.r + bek * (0.5 * (u * u + v * v));
    double bu = b.u + bek * r * u;
    double bv = b.v + bek * r * v;
    // u = q1/r, v = q2/r
    Q4 out;
    out.v[1] = bu / r;
    out.v[2] = bv / r;
    out.v[0] = br - (bu * u + bv * v) / r;
    out.v[3] = bq3;
    return out;
}

// Adjoint of the x-flux of a primitive state, F = (r u, r u u + p, r u v, u (p/(g-1) + ek + p)),
// accumulated into b.
__host__ __device__ inline void prim_flux_vjp(const Prim& W, const double* bF, const Phys& ph, PrimBar& b) {
    const double ru = W.r * W.u;
    const double ek = (0.5 * W.r) * (W.u * W.u + W.v * W.v);
    const double bru = bF[0] + bF[1] * W.u + bF[2] * W.v;
    const double bek = bF[3] * W.u;
    b.u += bF[1] * ru + bF[3] * ((ph.one_over_gm1 * W.p + ek) + W.p) + bru * W.r + bek * W.r * W.u;
    b.p += bF[1] + bF[3] * W.u * (ph.one_over_gm1 + 1.0);
    b.v += bF[2] * ru + bek * W.r * W.v;
    b.r += bru * W.u + bek * 0.5 * (W.u * W.u + W.v * W.v);
}

// Adjoint of the conservative state of a primitive state (PrimitiveConverter.to_conservative).
__host__ __device__ inline void prim_cons_vjp(const Prim& W, const double* bU, const Phys& ph, PrimBar& b) {
    b.r += bU[0] + bU[1] * W.u + bU[2] * W.v + bU[3] * 0.5 * (W.u * W.u + W.v * W.v);
    b.u += bU[1] * W.r + bU[3] * W.r * W.u;
    b.v += bU[2] * W.r + bU[3] * W.r * W.v;
    b.p += bU[3] / ph.gm1;
}

// Adjoint of roe_flux (flux/Roe.py:262-304 as restated in roe_flux): lamF -> (bL, bR).
__host__ __device__ inline void roe_flux_vjp(const Prim& L, const Prim& R, const Q4& lamF, const Phys& ph,
                                             PrimBar& bL, PrimBar& bR) {
    const double g = ph.gamma, gm = ph.g_over_gm1;
    // ---------------- primal recompute (the dissipation's intermediates; the central part 0.5*(F_L + F_R)
    // has its own adjoint below)
    const double sL = sqrt(L.r), sR = sqrt(R.r);
    const double inv = 1.0 / (sL + sR);
    const double rs = sqrt(L.r * R.r);
    const double nu = L.u * sL + R.u * sR, nv = L.v * sL + R.v * sR, np_ = L.p * sL + R.p * sR;
    const double us = nu * inv, vs = nv * inv, ps = np_ * inv;
    const double aL = sqrt(g * L.p / L.r), aR = sqrt(g * R.p / R.r), as = sqrt(g * ps / rs);
    const double l1_0 = us - as, l4_0 = us + as;
    const double tp_raw = 2.0 * ((R.u - aR) - (L.u - aL));
    const double tm_raw = 2.0 * ((R.u + aR) - (L.u + aL));
    const bool tp_cl = tp_raw <= 0.0, tm_cl = tm_raw <= 0.0;
    const double tp = tp_cl ? 1e-8 : tp_raw, tm = tm_cl ? 1e-8 : tm_raw;
    const bool h1 = fabs(l1_0) < tp, h4 = fabs(l4_0) < tm;
    const double l1 = h1 ? 0.5 * ((l1_0 * l1_0) / tp + tp) : l1_0;
    const double l4 = h4 ? 0.5 * ((l4_0 * l4_0) / tm + tm) : l4_0;
    const double Ek = 0.5 * (us * us + vs * vs);
    const double H = gm * ps / rs + Ek;
    const double ua = us * as;
    const double ia = 1.0 / as, ia2 = ia * ia, h = 0.5 * ia2, r2a = 0.5 * rs * ia;
    const double d0 = R.r - L.r, d1 = R.u - L.u, d2 = R.v - L.v, d3 = R.p - L.p;
    const double y0 = (-r2a) * d1 + h * d3, y1 = d0 + (-ia2) * d3, y2 = d2, y3 = r2a * d1 + h * d3;
    const double al1 = fabs(l1), aus = fabs(us), al4 = fabs(l4);
    const double z0 = al1 * y0, z1 = aus * y1, z2 = aus * y2, z3 = al4 * y3;
    // ---------------- reverse
    double bFL[4], bFR[4], bs[4];
    for (int k = 0; k < 4; ++k) {
        bFL[k] = 0.5 * lamF.v[k];
        bFR[k] = 0.5 * lamF.v[k];
        bs[k] = -0.5 * lamF.v[k];
    }
    // s3 = (H-ua) z0 + Ek z1 + vs z2 + (H+ua) z3
    double bH = bs[3] * (z0 + z3), bua = bs[3] * (z3 - z0), bEk = bs[3] * z1, bvs = bs[3] * z2;
    double bz0 = bs[3] * (H - ua), bz1 = bs[3] * Ek, bz2 = bs[3] * vs, bz3 = bs[3] * (H + ua);
    // s2 = vs z0 + vs z1 + z2 + vs z3
    bvs += bs[2] * (z0 + z1 + z3);
    bz0 += bs[2] * vs; bz1 += bs[2] * vs; bz2 += bs[2]; bz3 += bs[2] * vs;
    // s1 = l1 z0 + us z1 + l4 z3
    double bl1 = bs[1] * z0, bus = bs[1] * z1, bl4 = bs[1] * z3;
    bz0 += bs[1] * l1; bz1 += bs[1] * us; bz3 += bs[1] * l4;
    // s0 = z0 + z1 + z3
    bz0 += bs[0]; bz1 += bs[0]; bz3 += bs[0];
    // z = |lambda| y
    bl1 += bz0 * y0 * sgn(l1);
    bus += (bz1 * y1 + bz2 * y2) * sgn(us);
    bl4 += bz3 * y3 * sgn(l4);
    const double by0 = bz0 * al1, by1 = bz1 * aus, by2 = bz2 * aus, by3 = bz3 * al4;
    // y's
    const double br2a = (by3 - by0) * d1;
    const double bh = (by0 + by3) * d3;
    double bia2 = -by1 * d3;
    const double bd0 = by1, bd1 = (by3 - by0) * r2a, bd2 = by2, bd3 = (by0 + by3) * h - by1 * ia2;
    // h = 0.5 ia2; ia2 = ia*ia; r2a = 0.5 rs ia; ia = 1/as
    bia2 += 0.5 * bh;
    double bia = 2.0 * ia * bia2 + 0.5 * rs * br2a;
    double brs = 0.5 * ia * br2a;
    double bas = -ia * ia * bia;
    // ua = us*as
    bus += as * bua;
    bas += us * bua;
    // H = gm ps/rs + Ek
    double bps = gm / rs * bH;
    brs += -gm * ps / (rs * rs) * bH;
    bEk += bH;
    // Ek = 0.5(us^2 + vs^2)
    bus += us * bEk;
    bvs += vs * bEk;
    // Harten
    double bl1_0 = bl1, bl4_0 = bl4, btp = 0.0, btm = 0.0;
    if (h1) { bl1_0 = bl1 * l1_0 / tp; btp = bl1 * 0.5 * (1.0 - (l1_0 * l1_0) / (tp * tp)); }
    if (h4) { bl4_0 = bl4 * l4_0 / tm; btm = bl4 * 0.5 * (1.0 - (l4_0 * l4_0) / (tm * tm)); }
    if (tp_cl) btp = 0.0;
    if (tm_cl) btm = 0.0;
    double bRu = 2.0 * (btp + btm), bLu = -2.0 * (btp + btm);
    double baR = 2.0 * (btm - btp), baL = 2.0 * (btp - btm);
    // l1_0 = us - as, l4_0 = us + as
    bus += bl1_0 + bl4_0;
    bas += bl4_0 - bl1_0;
    // as = sqrt(g ps / rs)
    {
        const double barg = bas * 0.5 / as;
        bps += barg * g / rs;
        brs += -barg * g * ps / (rs * rs);
    }
    double bLr = 0.0, bLp = 0.0, bRr = 0.0, bRp = 0.0, bLv = 0.0, bRv = 0.0;
    {
        const double bargL = baL * 0.5 / aL, bargR = baR * 0.5 / aR;
        bLp += bargL * g / L.r; bLr += -bargL * g * L.p / (L.r * L.r);
        bRp += bargR * g / R.r; bRr += -bargR * g * R.p / (R.r * R.r);
    }
    // us, vs, ps = n*inv
    double binv = bus * nu + bvs * nv + bps * np_;
    const double bnu = bus * inv, bnv = bvs * inv, bnp = bps * inv;
    double bsL = bnu * L.u + bnv * L.v + bnp * L.p;
    double bsR = bnu * R.u + bnv * R.v + bnp * R.p;
    bLu += bnu * sL; bRu += bnu * sR;
    bLv += bnv * sL; bRv += bnv * sR;
    bLp += bnp * sL; bRp += bnp * sR;
    // inv = 1/(sL + sR)
    {
        const double bsum = -inv * inv * binv;
        bsL += bsum;
        bsR += bsum;
    }
    // rs = sqrt(L.r R.r)
    {
        const double bprod = brs * 0.5 / rs;
        bLr += bprod * R.r;
        bRr += bprod * L.r;
    }
    bLr += bsL * 0.5 / sL;
    bRr += bsR * 0.5 / sR;
    // d = R - L
    bRr += bd0; bLr -= bd0;
    bRu += bd1; bLu -= bd1;
    bRv += bd2; bLv -= bd2;
    bRp += bd3; bLp -= bd3;
    // F from primitives (PrimitiveState._F_from_prim_JIT)
    PrimBar tL{bLr, bLu, bLv, bLp}, tR{bRr, bRu, bRv, bRp};
    prim_flux_vjp(L, bFL, ph, tL);
    prim_flux_vjp(R, bFR, ph, tR);
    bLr = tL.r; bLu = tL.u; bLv = tL.v; bLp = tL.p;
    bRr = tR.r; bRu = tR.u; bRv = tR.v; bRp = tR.p;
    bL = PrimBar{bLr, bLu, bLv, bLp};
    bR = PrimBar{bRr, bRu, bRv, bRp};
}

// Adjoint of hlll_flux (flux/HLLL.py:25-49,70-103 as restated in hlll_flux): lamF -> (bL, bR).
__host__ __device__ inline void hlll_flux_vjp(const Prim& L, const Prim& R, const Q4& lamF, const Phys& ph,
                                              PrimBar& bL, PrimBar& bR) {
    const double g = ph.gamma;
    // ---------------- primal recompute
    const double sL = sqrt(L.r), sR = sqrt(R.r);
    const double inv = 1.0 / (sL + sR);
    const double rs = sqrt(L.r * R.r);
    const double nu = L.u * sL + R.u * sR, np_ = L.p * sL + R.p * sR;
    const double us = nu * inv, ps = np_ * inv;
    const double as = sqrt(g * ps / rs);
    const double aL = sqrt(g * L.p / L.r), aR = sqrt(g * R.p / R.r);
    const double Lslow = L.u - aL, Lfast = L.u + aL, Rslow = R.u - aR, Rfast = R.u + aR;
    const double l1_0 = us - as, l4_0 = us + as;
    const double tp_raw = 2.0 * (Rslow - Lslow), tm_raw = 2.0 * (Rfast - Lfast);
    const bool tp_cl = tp_raw <= 0.0, tm_cl = tm_raw <= 0.0;
    const double tp = tp_cl ? 1e-8 : tp_raw, tm = tm_cl ? 1e-8 : tm_raw;
    const bool h1 = fabs(l1_0) < tp, h4 = fabs(l4_0) < tm;
    const double l1 = h1 ? 0.5 * ((l1_0 * l1_0) / tp + tp) : l1_0;
    const double l4 = h4 ? 0.5 * ((l4_0 * l4_0) / tm + tm) : l4_0;
    const bool plus_fast = Rfast >= l4, minus_slow = Lslow <= l1;
    const double Lplus = plus_fast ? Rfast : l4;
    const double Lminus = minus_slow ? Lslow : l1;
    PrimBar tL{0.0, 0.0, 0.0, 0.0}, tR{0.0, 0.0, 0.0, 0.0};
    if (Lminus >= 0.0) {
        prim_flux_vjp(L, lamF.v, ph, tL);
        bL = tL;
        bR = tR;
        return;
    }
    if (Lplus <= 0.0) {
        prim_flux_vjp(R, lamF.v, ph, tR);
        bL = tL;
        bR = tR;
        return;
    }
    const double ekL = (0.5 * L.r) * (L.u * L.u + L.v * L.v);
    const double ekR = (0.5 * R.r) * (R.u * R.u + R.v * R.v);
    const double UL[4] = {L.r, L.r * L.u, L.r * L.v, L.p / ph.gm1 + ekL};
    const double UR[4] = {R.r, R.r * R.u, R.r * R.v, R.p / ph.gm1 + ekR};
    const double FL[4] = {UL[1], UL[1] * L.u + L.p, UL[1] * L.v, L.u * (UL[3] + L.p)};
    const double FR[4] = {UR[1], UR[1] * R.u + R.p, UR[1] * R.v, R.u * (UR[3] + R.p)};
    double dU[4], dF[4], m[4];
    for (int k = 0; k < 4; ++k) {
        dU[k] = UR[k] - UL[k];
        dF[k] = FR[k] - FL[k];
        m[k] = dF[k] - us * dU[k];
    }
    const double N1 = nrm4(dU[0], dU[1], dU[2], dU[3]);
    const double kk = as * N1;
    const double n = nrm4(m[0], m[1], m[2], m[3]);
    const double den = (kk < 1e-16) ? (kk + 1e-14) : kk;
    const double alpha_raw = 1.0 - n / den;
    const bool a_cl = 0.0 >= alpha_raw;
    const double alpha = a_cl ? 0.0 : alpha_raw;
    const double ra = us / Lminus, rb = us / Lplus;
    const bool pick_a = ra >= rb;
    const double mx = pick_a ? ra : rb;
    const double T = 1.0 - alpha * (1.0 - mx);
    const double coef = Lminus * Lplus * T;
    const double w = Lplus - Lminus;
    double G[4], F[4];
    for (int k = 0; k < 4; ++k) {
        G[k] = Lplus * FL[k] - Lminus * FR[k] + coef * dU[k];
        F[k] = G[k] / w;
    }
    // ---------------- reverse
    double bG[4], bFL[4], bFR[4], bdU[4], bUL[4], bUR[4];
    double bw = 0.0, bLplus = 0.0, bLminus = 0.0, bcoef = 0.0;
    for (int k = 0; k < 4; ++k) {
        bG[k] = lamF.v[k] / w;
        bw -= lamF.v[k] * F[k] / w;
        bLplus += bG[k] * FL[k];
        bLminus -= bG[k] * FR[k];
        bFL[k] = bG[k] * Lplus;
        bFR[k] = -bG[k] * Lminus;
        bcoef += bG[k] * dU[k];
        bdU[k] = bG[k] * coef;
    }
    bLplus += bw;
    bLminus -= bw;
    bLminus += bcoef * Lplus * T;
    bLplus += bcoef * Lminus * T;
    const double bT = bcoef * Lminus * Lplus;
    const double balpha = a_cl ? 0.0 : -bT * (1.0 - mx);
    const double bmx = bT * alpha;
    double bus = 0.0, bas = 0.0;
    if (pick_a) {
        bus += bmx / Lminus;
        bLminus -= bmx * us / (Lminus * Lminus);
    } else {
        bus += bmx / Lplus;
        bLplus -= bmx * us / (Lplus * Lplus);
    }
    // alpha = 1 - n/den, den = k (+1e-14), k = as*N1
    const double bn = -balpha / den;
    const double bk = balpha * n / (den * den);
    bas += bk * N1;
    const double bN1 = bk * as;
    double bdF[4];
    for (int k = 0; k < 4; ++k) {
        if (N1 > 0.0) bdU[k] += bN1 * dU[k] / N1;
        const double bm = (n > 0.0) ? bn * m[k] / n : 0.0;
        bdF[k] = bm;
        bus -= bm * dU[k];
        bdU[k] -= bm * us;
    }
    for (int k = 0; k < 4; ++k) {
        bFR[k] += bdF[k];
        bFL[k] -= bdF[k];
        bUR[k] = bdU[k];
        bUL[k] = -bdU[k];
    }
    prim_flux_vjp(L, bFL, ph, tL);
    prim_flux_vjp(R, bFR, ph, tR);
    prim_cons_vjp(L, bUL, ph, tL);
    prim_cons_vjp(R, bUR, ph, tR);
    // Lplus, Lminus selections
    double bRfast = 0.0, bLslow = 0.0, bl1 = 0.0, bl4 = 0.0;
    if (plus_fast) bRfast += bLplus; else bl4 += bLplus;
    if (minus_slow) bLslow += bLminus; else bl1 += bLminus;
    // Harten
    double bl1_0 = bl1, bl4_0 = bl4, btp = 0.0, btm = 0.0;
    if (h1) { bl1_0 = bl1 * l1_0 / tp; btp = bl1 * 0.5 * (1.0 - (l1_0 * l1_0) / (tp * tp)); }
    if (h4) { bl4_0 = bl4 * l4_0 / tm; btm = bl4 * 0.5 * (1.0 - (l4_0 * l4_0) / (tm * tm)); }
    if (tp_cl) btp = 0.0;
    if (tm_cl) btm = 0.0;
    const double bRslow = 2.0 * btp;
    bLslow -= 2.0 * btp;
    bRfast += 2.0 * btm;
    const double bLfast = -2.0 * btm;
    bus += bl1_0 + bl4_0;
    bas += bl4_0 - bl1_0;
    // slow/fast speeds
    tR.u += bRslow + bRfast;
    const double baR = bRfast - bRslow;
    tL.u += bLslow + bLfast;
    const double baL = bLfast - bLslow;
    {
        const double bargL = baL * 0.5 / aL, bargR = baR * 0.5 / aR;
        tL.p += bargL * g / L.r; tL.r -= bargL * g * L.p / (L.r * L.r);
        tR.p += bargR * g / R.r; tR.r -= bargR * g * R.p / (R.r * R.r);
    }
    // as = sqrt(g ps/rs)
    double bps = 0.0, brs = 0.0;
    {
        const double barg = bas * 0.5 / as;
        bps += barg * g / rs;
        brs -= barg * g * ps / (rs * rs);
    }
    // us, ps
    const double binv = bus * nu + bps * np_;
    const double bnu = bus * inv, bnp = bps * inv;
    double bsL = bnu * L.u + bnp * L.p, bsR = bnu * R.u + bnp * R.p;
    tL.u += bnu * sL; tR.u += bnu * sR;
    tL.p += bnp * sL; tR.p += bnp * sR;
    bsL -= inv * inv * binv;
    bsR -= inv * inv * binv;
    {
        const double bprod = brs * 0.5 / rs;
        tL.r += bprod * R.r;
        tR.r += bprod * L.r;
    }
    tL.r += bsL * 0.5 / sL;
    tR.r += bsR * 0.5 / sR;
    bL = tL;
    bR = tR;
}

// Adjoint of face_flux<FLUX> on conservative face states (cons_to_prim, then the flux).
template <int FLUX>
__host__ __device__ inline void face_flux_vjp(const Q4& qL, const Q4& qR, const Q4& lamF, const Phys& ph, Q4& bqL,
                                              Q4& bqR) {
    const Prim L = cons_to_prim(qL, ph), R = cons_to_prim(qR, ph);
    PrimBar bL, bR;
    if (FLUX == 0) roe_flux_vjp(L, R, lamF, ph, bL, bR);
    else hlll_flux_vjp(L, R, lamF, ph, bL, bR);
    bqL = cons_to_prim_vjp(qL, bL, ph);
    bqR = cons_to_prim_vjp(qR, bR, ph);
}

// face_flux_vjp on face states in the reconstruction variables (face_flux_recon): with primitive
// reconstruction the adjoints are those of the primitive states.
template <int FLUX>
__host__ __device__ inline void face_flux_vjp_recon(const Q4& qL, const Q4& qR, const Q4& lamF, const Phys& ph,
                                                    Q4& bqL, Q4& bqR) {
    if (!ph.prim) {
        face_flux_vjp<FLUX>(qL, qR, lamF, ph, bqL, bqR);
        return;
    }
    PrimBar bL, bR;
    if (FLUX == 0) roe_flux_vjp(as_prim(qL), as_prim(qR), lamF, ph, bL, bR);
    else hlll_flux_vjp(as_prim(qL), as_prim(qR), lamF, ph, bL, bR);
    bqL.v[0] = bL.r; bqL.v[1] = bL.u; bqL.v[2] = bL.v; bqL.v[3] = bL.p;
    bqR.v[0] = bR.r; bqR.v[1] = bR.u; bqR.v[2] = bR.v; bqR.v[3] = bR.p;
}

// Adjoint of reconstruct() for one cell: given the adjoints of the four limited face states,
// the adjoints of the cell and of its four neighbour values (each a 4-vector).
struct ReconBar {
    Q4 c, w, e, s, n;
};

__host__ __device__ inline double venk_prime(double s) {
    const double den = s * s + s + 2.0;
    return (-s * s + 4.0 * s + 4.0) / (den * den);
}

// Derivative of limiter_fn(lim, s) with respect to s (s >= 0 on every reachable branch).
__host__ __device__ inline double limiter_prime(int lim, double s) {
    switch (lim) {
        case 1: return (s < 1.0) ? 1.0 : 0.0;                                        // min(1, s)
        case 2: { const double den = s * s + 1.0; return (-s * s + 2.0 * s + 1.0) / (den * den); }
        case 3: return (s >= 0.0) ? 2.0 / ((s + 1.0) * (s + 1.0)) : 0.0;            // 2s/(s+1)
        default: return venk_prime(s);
    }
}

// Adjoint of reconstruct_var for one conserved variable: bq = adjoints of the limited E, W, N, S
// values; out = adjoints of (cell, W, E, S, N) values of that variable.
__host__ __device__ inline void reconstruct_vjp_var(double u, double uw, double ue, double us, double un,
                                                    const CellGeo& g, const double bq[4], int lim, double out[5]) {
    const double vals[5] = {u, uw, ue, us, un};
    const double e = 0.5 * (u + ue), w = 0.5 * (uw + u);
    const double n = 0.5 * (u + un), s = 0.5 * (us + u);
    const double gx = (e * g.xlE + w * g.xlW + n * g.xlN + s * g.xlS) * g.ainv;
    const double gy = (e * g.ylE + w * g.ylW + n * g.ylN + s * g.ylS) * g.ainv;
    const double hx[4] = {g.dxE, g.dxW, g.dxN, g.dxS};
    const double hy[4] = {g.dyE, g.dyW, g.dyN, g.dyS};
    double hf[4];
    for (int f = 0; f < 4; ++f) hf[f] = gx * hx[f] + gy * hy[f];
    int imax = 0, imin = 0;
    double mx = vals[0], mn = vals[0];
    for (int q = 1; q < 5; ++q) {
        if (vals[q] > mx) { mx = vals[q]; imax = q; }
        if (vals[q] < mn) { mn = vals[q]; imin = q; }
    }
    const double dmax = mx - u, dmin = mn - u;
    // the active face (first minimum of the four limiter values) and its d and slope; the arrays
    // are indexed by loop constants only, so that they stay in registers on the device
    int fa = 0;
    double pa = 0.0, da = 0.0, sa = 0.0;
    for (int f = 0; f < 4; ++f) {
        const double d = (u + hf[f]) - u;
        const double sl = face_slope(dmax, dmin, d);
        const double p = limiter_fn(lim, sl);
        if (f == 0 || p < pa) { pa = p; da = d; sa = sl; fa = f; }
    }
    const double phi_raw = pa;
    const double phi = (phi_raw < 0.0) ? 0.0 : phi_raw;
    // Q_f = u + phi h_f
    double bu = bq[0] + bq[1] + bq[2] + bq[3];
    double bphi = 0.0;
    for (int f = 0; f < 4; ++f) bphi += bq[f] * hf[f];
    double bval[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double bd = 0.0;
    bool active = false;
    if (!(phi_raw < 0.0) && (da > 0.0 || da < 0.0)) {
        const double bsv = bphi * limiter_prime(lim, sa);
        const double bnum = bsv / da;
        bd = -bsv * sa / da;
        active = true;
        const int iv = (da > 0.0) ? imax : imin;
        for (int t = 0; t < 5; ++t)
            if (t == iv) bval[t] += bnum;
        bu -= bnum;
    }
    double bgx = 0.0, bgy = 0.0;
    for (int f = 0; f < 4; ++f) {
        double bh = phi * bq[f];
        if (active && f == fa) bh += bd;
        bgx += bh * hx[f];
        bgy += bh * hy[f];
    }
    const double be = (bgx * g.xlE + bgy * g.ylE) * g.ainv;
    const double bw = (bgx * g.xlW + bgy * g.ylW) * g.ainv;
    const double bn = (bgx * g.xlN + bgy * g.ylN) * g.ainv;
    const double bs = (bgx * g.xlS + bgy * g.ylS) * g.ainv;
    bu += 0.5 * (be + bw + bn + bs);
    out[0] = bu + bval[0];
    out[1] = 0.5 * bw + bval[1];
    out[2] = 0.5 * be + bval[2];
    out[3] = 0.5 * bs + bval[3];
    out[4] = 0.5 * bn + bval[4];
}

// reconstruct_vjp_var with nq Gauss points per face (reconstruct_q_var): bq[f][i] = adjoint of the
// limited value at point i of face f (E, W, N, S); phi is the minimum over the 4*nq points in the
// order E points, W points, N points, S points, and the first minimum is the active one.  Only the
// active point's slope data is kept; the high-order terms are recomputed in the reverse sweep.
__host__ __device__ inline void reconstruct_vjp_var_q(double u, double uw, double ue, double us, double un,
                                                      const CellGeo& g, const QGeo& q, int nq,
                                                      const double bq[4][QMAX], int lim, double out[5]) {
    const double vals[5] = {u, uw, ue, us, un};
    const double e = 0.5 * (u + ue), w = 0.5 * (uw + u);
    const double n = 0.5 * (u + un), s = 0.5 * (us + u);
    const double gx = (e * g.xlE + w * g.xlW + n * g.xlN + s * g.xlS) * g.ainv;
    const double gy = (e * g.ylE + w * g.ylW + n * g.ylN + s * g.ylS) * g.ainv;
    int imax = 0, imin = 0;
    double mx = vals[0], mn = vals[0];
    for (int t = 1; t < 5; ++t) {
        if (vals[t] > mx) { mx = vals[t]; imax = t; }
        if (vals[t] < mn) { mn = vals[t]; imin = t; }
    }
    const double dmax = mx - u, dmin = mn - u;
    double pa = 0.0, da = 0.0, sa = 0.0, xa = 0.0, ya = 0.0;  // active point: phi, d, slope, offset
    for (int f = 0; f < 4; ++f)
        for (int i = 0; i < QMAX; ++i) {
            if (i >= nq) continue;
            const double h = gx * q.dx[f][i] + gy * q.dy[f][i];
            const double d = (u + h) - u;
            const double sl = face_slope(dmax, dmin, d);
            const double p = limiter_fn(lim, sl);
            if ((f == 0 && i == 0) || p < pa) { pa = p; da = d; sa = sl; xa = q.dx[f][i]; ya = q.dy[f][i]; }
        }
    const double phi = (pa < 0.0) ? 0.0 : pa;
    // Q_fi = u + phi h_fi
    double bu = 0.0, bphi = 0.0, bgx = 0.0, bgy = 0.0;
    for (int f = 0; f < 4; ++f)
        for (int i = 0; i < QMAX; ++i) {
            if (i >= nq) continue;
            const double h = gx * q.dx[f][i] + gy * q.dy[f][i];
            bu += bq[f][i];
            bphi += bq[f][i] * h;
            const double bh = phi * bq[f][i];
            bgx += bh * q.dx[f][i];
            bgy += bh * q.dy[f][i];
        }
    double bval[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (!(pa < 0.0) && (da > 0.0 || da < 0.0)) {
        const double bsv = bphi * limiter_prime(lim, sa);
        const double bnum = bsv / da;
        const double bd = -bsv * sa / da;
        const int iv = (da > 0.0) ? imax : imin;
        for (int t = 0; t < 5; ++t)
            if (t == iv) bval[t] += bnum;
        bu -= bnum;
        bgx += bd * xa;
        bgy += bd * ya;
    }
    const double be = (bgx * g.xlE + bgy * g.ylE) * g.ainv;
    const double bw = (bgx * g.xlW + bgy * g.ylW) * g.ainv;
    const double bn = (bgx * g.xlN + bgy * g.ylN) * g.ainv;
    const double bs = (bgx * g.xlS + bgy * g.ylS) * g.ainv;
    bu += 0.5 * (be + bw + bn + bs);
    out[0] = bu + bval[0];
    out[1] = 0.5 * bw + bval[1];
    out[2] = 0.5 * be + bval[2];
    out[3] = 0.5 * bs + bval[3];
    out[4] = 0.5 * bn + bval[4];
}

__host__ __device__ inline ReconBar reconstruct_vjp(const Q4& U, const Q4& Uw, const Q4& Ue, const Q4& Us,
                                                    const Q4& Un, const CellGeo& g, const Recon& bQ, int lim = 0) {
    ReconBar out;
    for (int k = 0; k < 4; ++k) {
        const double bq[4] = {bQ.E.v[k], bQ.W.v[k], bQ.N.v[k], bQ.S.v[k]};
        double o[5];
        reconstruct_vjp_var(U.v[k], Uw.v[k], Ue.v[k], Us.v[k], Un.v[k], g, bq, lim, o);
        out.c.v[k] = o[0];
        out.w.v[k] = o[1];
        out.e.v[k] = o[2];
        out.s.v[k] = o[3];
        out.n.v[k] = o[4];
    }
    return out;
}

}  // namespace phx
