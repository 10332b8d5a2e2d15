// phx_numerics.hpp -- per-cell and per-face arithmetic of pyHype's second-order MUSCL scheme.
//
// Every function reproduces the reference's floating-point evaluation order exactly (left-to-right
// Python/Numba association, the COO storage order of the Roe eigen-system mat-vecs, no fused
// multiply-add).  The translation unit must be compiled with -ffp-contract=off; division and
// square root are the IEEE correctly-rounded operations.  The functions are __host__ __device__ so
// that the same source is checked against the CPU oracle on the host.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include <type_traits>

namespace phx {

// A state or flux of 4 components.  The arithmetic below is written once for a scalar type T:
// T = double is the path every kernel runs; the tangent-linear kernels instantiate it with a
// value-and-derivative pair whose value part performs exactly the same double operations.
template <class T>
struct Q4T {
    T v[4];
};
using Q4 = Q4T<double>;

// Fluid constants, evaluated on the host exactly as the reference evaluates them:
// gm1 = gamma - 1 (states/conservative.py:110), one_over_gm1 = 1.0/(gamma - 1.0) and
// g_over_gm1 = gamma/(gamma - 1.0) (fluids/base.py:58-64).
struct Phys {
    double gamma, gm1, one_over_gm1, g_over_gm1, cfl;
    int limiter;  // PHX_LIMITER_* (0 = Venkatakrishnan)
    int prim;     // reconstruction in primitive variables (SolverConfig.reconstruction_type)
    int nq;       // Gauss points per face (SolverConfig.fvm_num_quadrature_points); 0 or 1: midpoint
};

template <class T>
struct PrimT {
    T r, u, v, p;
};
using Prim = PrimT<double>;

__host__ __device__ inline Q4 q4zero() {
    Q4 q;
    q.v[0] = q.v[1] = q.v[2] = q.v[3] = 0.0;
    return q;
}

// IEEE square root of a double.  The compiler's sequence scales x by 2^256 when x < 2^-767, runs
// v_rsq_f64 and a Newton-Raphson refinement (g = x*y, h = y/2, r = 1/2 - h*g, g += g*r, h += h*r,
// then twice g += (x - g*g)*h), scales back by 2^-128, and returns x itself for +-0 and +inf.  For
// 2^-767 <= x <= DBL_MAX the scalings and that select are identities, so the refinement alone gives
// the same bits: sqrt_rn evaluates it and takes the library sqrt only for other x (a branch that
// waves skip when no lane needs it), 12 instead of 17 instructions with the range test.
__host__ __device__ inline double sqrt_rn(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    double d = __builtin_fma(-g, g, x);
    h = __builtin_fma(h, r, h);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    if (!(x >= 0x1p-767 && x <= 0x1.fffffffffffffp+1023)) {
        // an empty volatile asm keeps this a branch: the compiler would otherwise evaluate the
        // library sqrt for every lane and select
        double xs = x;
        asm volatile("" : "+v"(xs));
        g = sqrt(xs);
    }
    return g;
#else
    return sqrt(x);
#endif
}
// sqrt of the flux arithmetic: sqrt_rn for doubles, the type's own sqrt otherwise
template <class T>
__host__ __device__ inline T sqrt_f(T x) {
    return sqrt(x);
}
__host__ __device__ inline double sqrt_f(double x) { return sqrt_rn(x); }

// ConservativeConverter.to_primitive (states/converter/concrete_defs.py:85-100), with
// u = rhou/rho, v = rhov/rho (conservative.py:85-99), ek = rho*Ek, Ek = 0.5*(u*u+v*v)
// (conservative.py:118-124), p = (gamma-1)*(e - ek) (conservative.py:107-110).
template <class T>
__host__ __device__ inline PrimT<T> cons_to_prim(const Q4T<T>& q, const Phys& ph) {
    PrimT<T> w;
    w.r = q.v[0];
    w.u = q.v[1] / w.r;
    w.v = q.v[2] / w.r;
    const T ek = w.r * (0.5 * (w.u * w.u + w.v * w.v));
    w.p = ph.gm1 * (q.v[3] - ek);
    return w;
}

// rotate_JIT / unrotate_JIT (utils/utils.py:75-85,148-158) with c = cos(theta), s = sin(theta).
template <class T>
__host__ __device__ inline Q4T<T> rotate(Q4T<T> q, double c, double s) {
    const T u = q.v[1], v = q.v[2];
    q.v[1] = u * c + v * s;
    q.v[2] = v * c - u * s;
    return q;
}
template <class T>
__host__ __device__ inline Q4T<T> unrotate(Q4T<T> q, double c, double s) {
    const T u = q.v[1], v = q.v[2];
    q.v[1] = u * c - v * s;
    q.v[2] = v * c + u * s;
    return q;
}
// rotate90 / unrotate90 (utils/utils.py:88-112,161-185).
template <class T>
__host__ __device__ inline Q4T<T> rotate90(Q4T<T> q) {
    const T u = q.v[2], v = q.v[1];
    q.v[1] = u;
    q.v[2] = -v;
    return q;
}
template <class T>
__host__ __device__ inline Q4T<T> unrotate90(Q4T<T> q) {
    const T u = q.v[2], v = q.v[1];
    q.v[1] = -u;
    q.v[2] = v;
    return q;
}
// BoundaryConditionFunctions.reflection (boundary_conditions/funcs.py:27-43).
template <class T>
__host__ __device__ inline Q4T<T> reflect(Q4T<T> q, double c, double s) {
    q = rotate(q, c, s);
    q.v[1] = -q.v[1];
    return unrotate(q, c, s);
}

// Harten entropy fix on the Roe wave speeds (flux/base.py:119-146).  Inputs are the slow (u-a)
// and fast (u+a) speeds of the right, left and Roe states; l1/l4 are corrected in place.
template <class T>
__host__ __device__ inline void harten(T Rslow, T Rfast, T Lslow, T Lfast, T& l1, T& l4) {
    T tp = 2.0 * (Rslow - Lslow);
    T tm = 2.0 * (Rfast - Lfast);
    tp = (tp <= 0.0) ? T(1e-8) : tp;
    tm = (tm <= 0.0) ? T(1e-8) : tm;
    if (fabs(l1) < tp) l1 = 0.5 * ((l1 * l1) / tp + tp);
    if (fabs(l4) < tm) l4 = 0.5 * ((l4 * l4) / tm + tm);
}

// A face of a flat region: every jump dW = W_R - W_L is +0 (all bits zero), with a density and a
// pressure in [2^-300, 2^300] and speeds of at most 2^300.  There every intermediate of the Roe
// dissipation below is finite and each term of Rc|Lambda|Lp dW is a finite number times +0, so the
// four row sums are signed zeros: rows 0, 2 and 3 are +0 (each contains a term 1.0*z or Ek*z1 that
// is +0), and row 1's sign does not matter because 0.5*(F_L1 + F_R1) >= p > 0.  The flux is then
// exactly 0.5*(F_L + F_R), and roe_flux skips the square roots and divisions of the dissipation.
__host__ __device__ inline bool roe_flat(double dr, double du, double dv, double dp, const PrimT<double>& L) {
    const unsigned long long z = __builtin_bit_cast(unsigned long long, dr) | __builtin_bit_cast(unsigned long long, du) |
                                 __builtin_bit_cast(unsigned long long, dv) | __builtin_bit_cast(unsigned long long, dp);
    const double lo = 0x1p-300, hi = 0x1p300;
    return z == 0ull && L.r >= lo && L.r <= hi && L.p >= lo && L.p <= hi && fabs(L.u) <= hi && fabs(L.v) <= hi;
}

// FluxRoe.compute_flux (flux/Roe.py:262-304) for one face, primitive left/right states.
//  * Roe state: RoePrimitiveState._roe_state_from_prim_JIT (states/primitive.py:301-320) --
//    note the sqrt(rho)-weighted pressure.
//  * F from primitives: PrimitiveState._F_from_prim_JIT (states/primitive.py:223-237).
//  * Dissipation Rc |Lambda| Lp dW with the SciPy COO accumulation order of the reference's
//    block layout (flux/eigen_system.py:115-158,205-242; values flux/Roe.py:175-260).  The
//    leading "0 +" of each COO row sum only changes the sign of a zero and is omitted.
template <class T>
__host__ __device__ inline Q4T<T> roe_flux(const PrimT<T>& L, const PrimT<T>& R, const Phys& ph) {
    const T ekL = (0.5 * L.r) * (L.u * L.u + L.v * L.v);
    const T ekR = (0.5 * R.r) * (R.u * R.u + R.v * R.v);
    const T ruL = L.r * L.u, ruR = R.r * R.u;
    const T FL0 = ruL, FL1 = ruL * L.u + L.p, FL2 = ruL * L.v;
    const T FL3 = L.u * ((ph.one_over_gm1 * L.p + ekL) + L.p);
    const T FR0 = ruR, FR1 = ruR * R.u + R.p, FR2 = ruR * R.v;
    const T FR3 = R.u * ((ph.one_over_gm1 * R.p + ekR) + R.p);
    if constexpr (std::is_same<T, double>::value) {
        // flat region: the dissipation is exactly zero (roe_flat); the derivative of the dissipation
        // is not, so the tangent-linear instantiation always evaluates it
        if (roe_flat(R.r - L.r, R.u - L.u, R.v - L.v, R.p - L.p, L)) {
            Q4T<T> F;
            F.v[0] = 0.5 * (FL0 + FR0);
            F.v[1] = 0.5 * (FL1 + FR1);
            F.v[2] = 0.5 * (FL2 + FR2);
            F.v[3] = 0.5 * (FL3 + FR3);
            return F;
        }
    }

    const T sL = sqrt_f(L.r), sR = sqrt_f(R.r);
    const T inv = 1.0 / (sL + sR);
    const T rs = sqrt_f(L.r * R.r);
    const T us = (L.u * sL + R.u * sR) * inv;
    const T vs = (L.v * sL + R.v * sR) * inv;
    const T ps = (L.p * sL + R.p * sR) * inv;

    const T aL = sqrt_f(ph.gamma * L.p / L.r);
    const T aR = sqrt_f(ph.gamma * R.p / R.r);
    const T as = sqrt_f(ph.gamma * ps / rs);
    T l1 = us - as, l4 = us + as;
    harten(R.u - aR, R.u + aR, L.u - aL, L.u + aL, l1, l4);

    const T Ek = 0.5 * (us * us + vs * vs);           // Ek_JIT
    const T H = ph.g_over_gm1 * ps / rs + Ek;         // H_given_Ek_JIT
    const T ua = us * as;
    const T ia = 1.0 / as;                            // compute_Lp
    const T ia2 = ia * ia;
    const T h = 0.5 * ia2;
    const T r2a = 0.5 * rs * ia;

    const T d0 = R.r - L.r, d1 = R.u - L.u, d2 = R.v - L.v, d3 = R.p - L.p;
    // Lp * dW: rows in COO order (3,1),(1,0),(2,2),(3,3),(0,1),(1,3),(0,3)
    const T y0 = (-r2a) * d1 + h * d3;
    const T y1 = 1.0 * d0 + (-ia2) * d3;
    const T y2 = 1.0 * d2;
    const T y3 = r2a * d1 + h * d3;
    // |Lambda| (diagonal)
    const T z0 = fabs(l1) * y0, z1 = fabs(us) * y1, z2 = fabs(us) * y2, z3 = fabs(l4) * y3;
    // Rc * z: COO order (3,0),(2,0),(3,1),(1,0),(2,1),(3,2),(0,0),(1,1),(2,2),(3,3),(0,1),(2,3),(1,3),(0,3)
    const T s0 = (1.0 * z0 + 1.0 * z1) + 1.0 * z3;
    const T s1 = (l1 * z0 + us * z1) + l4 * z3;
    const T s2 = ((vs * z0 + vs * z1) + 1.0 * z2) + vs * z3;
    const T s3 = (((H - ua) * z0 + Ek * z1) + vs * z2) + (H + ua) * z3;

    Q4T<T> F;
    F.v[0] = 0.5 * (FL0 + FR0) - 0.5 * s0;
    F.v[1] = 0.5 * (FL1 + FR1) - 0.5 * s1;
    F.v[2] = 0.5 * (FL2 + FR2) - 0.5 * s2;
    F.v[3] = 0.5 * (FL3 + FR3) - 0.5 * s3;
    return F;
}

// ---- Euclidean norm of HLLL's alpha (flux/HLLL.py:86-93) ----
//
// Numba lowers np.linalg.norm of a 1-D float64 array to the BLAS routine dnrm2 of the SciPy build
// it runs against.  SciPy's OpenBLAS (0.3.28 with scipy 1.15.3) uses the x87 kernel on x86-64:
// for n < 8 it accumulates 0 + x0^2 + x1^2 + x2^2 + x3^2 in order in the x87 registers (64-bit
// significand, round to nearest even after every multiply and add), takes fsqrt in the same
// format and returns the result rounded to double.  That is 1 ulp away from the double-precision
// formula sqrt(((x0^2+x1^2)+x2^2)+x3^2) on about one vector in five, so the extended format is
// emulated exactly here with 64-bit integer arithmetic.  Non-negative operands only (squares and
// their sums); no overflow or underflow exists in the extended range for squares of doubles.

// A non-negative extended value m * 2^e: m has bit 63 set, or m = 0 for zero.
struct Ext {
    unsigned long long m;
    int e;
};

__host__ __device__ inline unsigned long long umulhi64(unsigned long long a, unsigned long long b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (unsigned long long)(((unsigned __int128)a * b) >> 64);
#endif
}

__host__ __device__ inline int clz64(unsigned long long x) { return __builtin_clzll(x); }

__host__ __device__ inline unsigned long long dbits(double x) { return __builtin_bit_cast(unsigned long long, x); }

// Rounds the 128-bit significand (hi, lo) * 2^(e - 64), hi bit 63 set, to 64 bits, nearest even;
// `sticky` flags non-zero bits below lo.
__host__ __device__ inline Ext ext_round(unsigned long long hi, unsigned long long lo, bool sticky, int e) {
    const bool half = (lo >> 63) != 0;
    const bool rest = ((lo << 1) != 0) || sticky;
    if (half && (rest || (hi & 1ull))) {
        ++hi;
        if (hi == 0) {
            hi = 1ull << 63;
            ++e;
        }
    }
    return Ext{hi, e};
}

// x*x rounded to the extended format (x87 fmul of two doubles loaded exactly).  Finite x.
__host__ __device__ inline Ext ext_sq(double x) {
    const unsigned long long b = dbits(x) & 0x7fffffffffffffffull;
    if (b == 0) return Ext{0ull, 0};
    const int be = (int)(b >> 52);
    unsigned long long M = b & 0x000fffffffffffffull;
    int E;
    if (be == 0) {
        E = -1074;  // subnormal: M * 2^-1074
    } else {
        M |= 1ull << 52;
        E = be - 1075;
    }
    const int lz = clz64(M);
    M <<= lz;  // x = M * 2^(E - lz), M in [2^63, 2^64)
    E -= lz;
    unsigned long long hi = umulhi64(M, M), lo = M * M;  // x^2 = (hi*2^64 + lo) * 2^(2E)
    int e = 2 * E + 64;
    if ((hi >> 63) == 0) {
        hi = (hi << 1) | (lo >> 63);
        lo <<= 1;
        e -= 1;
    }
    return ext_round(hi, lo, false, e);
}

// a + b rounded to the extended format (x87 fadd), a, b >= 0.
__host__ __device__ inline Ext ext_add(Ext a, Ext b) {
    if (b.m == 0) return a;
    if (a.m == 0) return b;
    if (a.e < b.e) {
        const Ext t = a;
        a = b;
        b = t;
    }
    const int d = a.e - b.e;
    // b < 2^(b.e + 64) <= 2^(a.e - 2), below half an ulp of a: a + b rounds to a.
    if (d > 65) return a;
    // a = (a.m * 2^64) * 2^(a.e - 64); b aligned to the same unit as (bh, bl) plus a sticky bit
    unsigned long long bh, bl;
    bool sticky = false;
    if (d == 0) {
        bh = b.m;
        bl = 0;
    } else if (d < 64) {
        bh = b.m >> d;
        bl = b.m << (64 - d);
    } else if (d == 64) {
        bh = 0;
        bl = b.m;
    } else {  // d == 65
        bh = 0;
        bl = b.m >> 1;
        sticky = (b.m & 1ull) != 0;
    }
    unsigned long long lo = bl;                 // a's low word is zero
    unsigned long long hi = a.m + bh;
    int e = a.e;
    if (hi < a.m) {  // carry out of bit 127: shift the 129-bit sum right by one
        sticky = sticky || (lo & 1ull);
        lo = (lo >> 1) | (hi << 63);
        hi = (hi >> 1) | (1ull << 63);
        e += 1;
    }
    return ext_round(hi, lo, sticky, e);
}

// 128-bit helpers for the square root (value = hi * 2^64 + lo).
__host__ __device__ inline bool u128_gt(unsigned long long ah, unsigned long long al, unsigned long long bh,
                                        unsigned long long bl) {
    return ah > bh || (ah == bh && al > bl);
}

// RN_double(RN_ext(sqrt(s))) for an extended s (x87 fsqrt, then the store to double).
__host__ __device__ inline double ext_sqrt_to_double(Ext s) {
    if (s.m == 0) return 0.0;
    // N = s.m * 2^k with s.e - k even, k in {63, 64}: N in [2^126, 2^128), sqrt(N) in [2^63, 2^64)
    const int k = ((s.e - 64) & 1) ? 63 : 64;
    const unsigned long long nh = (k == 64) ? s.m : (s.m >> 1);
    const unsigned long long nl = (k == 64) ? 0ull : (s.m << 63);
    const int E = (s.e - k) / 2;  // sqrt(s) = sqrt(N) * 2^E  (s.e - k is even)
    // q = floor(sqrt(N)): a double estimate (|error| < 2^13), one Newton step, then an exact fix-up
    const double dn = (double)nh * 18446744073709551616.0 + (double)nl;
    const double sq = sqrt(dn);
    unsigned long long q = (sq >= 18446744073709551616.0) ? 0xffffffffffffffffull : (unsigned long long)sq;
    {
        // rem = N - q^2 (|rem| < 2^79, two's complement over 128 bits)
        const unsigned long long qh = umulhi64(q, q), ql = q * q;
        const unsigned long long rl = nl - ql;
        const long long rh = (long long)(nh - qh - (nl < ql ? 1ull : 0ull));
        const double drem = (double)rh * 18446744073709551616.0 + (double)rl;
        const double step = drem / (2.0 * (double)q);
        const long long dq = (long long)(step + (step >= 0.0 ? 0.5 : -0.5));  // |dq| <= 2^13
        if (dq >= 0) {
            const unsigned long long up = (unsigned long long)dq;
            q = (q > 0xffffffffffffffffull - up) ? 0xffffffffffffffffull : q + up;
        } else {
            q -= (unsigned long long)(-dq);
        }
    }
    // now |q - sqrt(N)| < 1: q^2 > N -> q - 1; (q+1)^2 <= N -> q + 1 (twice each, for margin)
    for (int it = 0; it < 2; ++it) {
        const unsigned long long qh = umulhi64(q, q), ql = q * q;
        if (u128_gt(qh, ql, nh, nl)) --q;
    }
    for (int it = 0; it < 2; ++it) {
        if (q == 0xffffffffffffffffull) break;
        const unsigned long long p = q + 1;
        const unsigned long long ph = umulhi64(p, p), pl = p * p;
        if (!u128_gt(ph, pl, nh, nl)) q = p;
    }
    // round to 64 bits: up iff N - q^2 > q (sqrt(N) is never exactly q + 1/2)
    const unsigned long long qh = umulhi64(q, q), ql = q * q;
    const unsigned long long rl = nl - ql;
    const unsigned long long rh = nh - qh - (nl < ql ? 1ull : 0ull);  // 0 <= rem <= 2q < 2^65
    int e64 = E;                                                       // r64 * 2^e64
    unsigned long long r64 = q;
    if (rh != 0 || rl > q) {
        ++r64;
        if (r64 == 0) {
            r64 = 1ull << 63;
            ++e64;
        }
    }
    // round the 64-bit significand to the double's precision (53 bits, fewer if subnormal)
    int shift = 11;
    const int top = e64 + 63;  // exponent of the leading bit
    if (top < -1022) shift += -1022 - top;
    if (shift > 64) return 0.0;  // below half the smallest subnormal
    unsigned long long r53, rem_bits, halfv;
    if (shift == 64) {
        r53 = 0;
        rem_bits = r64;
        halfv = 1ull << 63;
    } else {
        r53 = r64 >> shift;
        rem_bits = r64 & ((1ull << shift) - 1ull);
        halfv = 1ull << (shift - 1);
    }
    if (rem_bits > halfv || (rem_bits == halfv && (r53 & 1ull))) ++r53;
    // r53 * 2^(e64 + shift) is exact in double (r53 <= 2^53); ldexp only scales
    return ldexp((double)r53, e64 + shift);
}

// The extended-precision norm by the integer emulation above (finite components).
__host__ __device__ inline double nrm4_ext(double a, double b, double c, double d) {
    Ext s = ext_sq(a);
    s = ext_add(s, ext_sq(b));
    s = ext_add(s, ext_sq(c));
    s = ext_add(s, ext_sq(d));
    return ext_sqrt_to_double(s);
}

// ---- fast path: the same extended-precision arithmetic carried out in pairs of doubles ----
// A value v >= 0 with a 64-bit significand is held as hi + lo (doubles, |lo| <= ulp(hi)/2, lo on
// v's 64-bit grid).  It covers components of magnitude 0 or in [2^-400, 2^400] (the low parts of the
// squares do not underflow, nothing overflows).  Each rounding to 64 bits is decided from a pair sum
// accurate to about 2^-104 of the value; where the value lies within 2^-30 of a grid unit of a
// midpoint of the 64-bit grid (exact ties included), or the square root within one 64-bit unit of a
// midpoint between two doubles (only there can the rounding to 64 bits change the rounding to
// double), the result is reported (ok = false) and taken from the integer emulation.
struct DD {
    double hi, lo;
};

// hi + lo = a + b exactly (Knuth's TwoSum)
__host__ __device__ inline DD two_sum(double a, double b) {
    const double s = a + b;
    const double bv = s - a;
    const double av = s - bv;
    return DD{s, (a - av) + (b - bv)};
}

// RN64(hi + lo), hi > 0, |lo| <= ulp(hi)/2: lo rounded to the 64-bit grid of the value's binade,
// 2^(E-63) for 2^E <= hi < 2^(E+1) (half that when hi = 2^E and lo < 0: the value is then in the
// binade below).  Adding and subtracting C = 1.5 * 2^52 * grid rounds lo to that grid, ties to even:
// C's last place is the grid unit and C / grid is even, so the parity is that of the 64-bit
// significand.  When lo carries a rounding error (exact = false), ok is cleared where lo lies within
// 2^-30 of a grid unit of a midpoint of the grid: the side of the midpoint is then not known.
__host__ __device__ inline DD rn64(double hi, double lo, bool exact, bool& ok) {
    int ex;
    const double m = frexp(hi, &ex);  // hi = m 2^ex, m in [1/2, 1): E = ex - 1
    double C = ldexp(0.75, ex - 11);  // 1.5 * 2^(E - 11) = 1.5 * 2^52 * 2^(E - 63)
    double half = ldexp(1.0, ex - 65);  // grid / 2
    if (m == 0.5 && lo < 0.0) {
        C *= 0.5;
        half *= 0.5;
    }
    const double l = (lo + C) - C;
    if (!exact && fabs(fabs(lo - l) - half) <= half * 0x1p-29) ok = false;
    return DD{hi, l};
}

// RN64(x^2), |x| = 0 or in [2^-400, 2^400]: x^2 = p + e exactly (fma)
__host__ __device__ inline DD sq64(double x, bool& ok) {
    const double p = x * x;
    if (p == 0.0) return DD{0.0, 0.0};
    return rn64(p, fma(x, x, -p), true, ok);
}

// RN64(A + B) of two 64-bit values >= 0 (x87 fadd).  The exact sum is s.hi + s.lo + A.lo + B.lo;
// the low terms are summed with their rounding errors e1, e2 tracked, so a sum carried out exactly
// (every case where the four terms span fewer than about 50 binades) decides ties exactly.
__host__ __device__ inline DD add64(DD A, DD B, bool& ok) {
    if (B.hi == 0.0) return A;
    if (A.hi == 0.0) return B;
    const DD s = two_sum(A.hi, B.hi);
    const DD t = two_sum(s.lo, A.lo);
    const DD w = two_sum(t.hi, B.lo);
    const double h = s.hi + w.hi;  // renormalised (Fast2Sum, |s.hi| >= |w.hi|): h + (w.hi - (h - s.hi)) exact
    return rn64(h, w.hi - (h - s.hi), t.lo == 0.0 && w.lo == 0.0, ok);
}

// RN53(RN64(sqrt(S))) of a 64-bit value S > 0 (x87 fsqrt, then the store to double).  z = sqrt(S)
// lies within one double unit of y = RN53(sqrt(S.hi)); z - y is known to about 2^-100 of y.  The
// rounding to 64 bits matters only where z is within half a 64-bit unit of a midpoint between two
// doubles: it then lands on the midpoint, and the store rounds that tie to the even neighbour.
__host__ __device__ inline double sqrt64(DD S, bool& ok) {
    const double y = sqrt(S.hi);
    const double p2 = y * y, e2 = fma(y, y, -p2);
    // z^2 - y^2 for z = sqrt(S) (S.hi - p2 is exact by Sterbenz), then z - y
    const double r = ((S.hi - p2) - e2) + S.lo;
    const double dlt = r / (2.0 * y);
    int ey;
    (void)frexp(y, &ey);                   // y in [2^(ey-1), 2^ey)
    const double ulp = ldexp(1.0, ey - 53);
    const double G = ldexp(1.0, ey - 64);  // 64-bit unit of y's binade
    const bool pow2 = (y == ldexp(0.5, ey));
    const double ulp_dn = pow2 ? 0.5 * ulp : ulp;
    const double g_dn = pow2 ? 0.25 * G : 0.5 * G;  // half the 64-bit unit at the lower midpoint
    const double up = dlt - 0.5 * ulp, dn = dlt + 0.5 * ulp_dn;  // z minus the midpoints
    const double mu = fabs(up), md = fabs(dn);
    if (fabs(mu - 0.5 * G) <= G * 0x1p-20 || fabs(md - g_dn) <= G * 0x1p-20) {  // not decided
        ok = false;
        return y;
    }
    const bool odd = (dbits(y) & 1ull) != 0ull;
    if (mu < 0.5 * G) return odd ? y + ulp : y;    // RN64(z) is the upper midpoint
    if (md < g_dn) return odd ? y - ulp_dn : y;     // RN64(z) is the lower midpoint
    if (up > 0.0) return y + ulp;
    if (dn < 0.0) return y - ulp_dn;
    return y;
}

__host__ __device__ inline bool nrm4_in_range(double x) {
    const double t = fabs(x);
    return t == 0.0 || (t >= 0x1p-400 && t <= 0x1p400);
}

// Euclidean norm of a 4-vector as dnrm2 computes it (see above).  Non-finite components take the
// double formula (inf or NaN, as the extended computation gives).  Components within [2^-400, 2^400]
// take the pair-of-doubles evaluation, which decides all but about one in 10^5 of them; the rest, and
// components outside that range, take the integer emulation.  Force-inlined: the kernels make no
// device-function calls.
__host__ __device__ __forceinline__ double nrm4(double a, double b, double c, double d) {
    // all components +-0 (the jumps across a face of a flat region): the extended sum is 0 and so
    // is the norm; the emulation below is skipped by whole waves there
    if (((dbits(a) | dbits(b) | dbits(c) | dbits(d)) & 0x7fffffffffffffffull) == 0ull) return 0.0;
    const unsigned long long ex = 0x7ff0000000000000ull;
    if ((dbits(a) & ex) == ex || (dbits(b) & ex) == ex || (dbits(c) & ex) == ex || (dbits(d) & ex) == ex)
        return sqrt(((a * a + b * b) + c * c) + d * d);
    bool ok = nrm4_in_range(a) && nrm4_in_range(b) && nrm4_in_range(c) && nrm4_in_range(d);
    if (ok) {
        DD s = sq64(a, ok);
        s = add64(s, sq64(b, ok), ok);
        s = add64(s, sq64(c, ok), ok);
        s = add64(s, sq64(d, ok), ok);
        if (ok) {
            const double r = sqrt64(s, ok);
            if (ok) return r;
        }
    }
    return nrm4_ext(a, b, c, d);
}

// FluxHLLL.compute_flux + _HLLL_flux_JIT (flux/HLLL.py:25-49,70-103) for one face.
template <class T>
__host__ __device__ inline Q4T<T> hlll_flux(const PrimT<T>& L, const PrimT<T>& R, const Phys& ph) {
    const T sL = sqrt_f(L.r), sR = sqrt_f(R.r);
    const T inv = 1.0 / (sL + sR);
    const T rs = sqrt_f(L.r * R.r);
    const T us = (L.u * sL + R.u * sR) * inv;
    const T ps = (L.p * sL + R.p * sR) * inv;
    const T as = sqrt_f(ph.gamma * ps / rs);
    const T aL = sqrt_f(ph.gamma * L.p / L.r);
    const T aR = sqrt_f(ph.gamma * R.p / R.r);
    const T Lslow = L.u - aL, Lfast = L.u + aL;
    const T Rslow = R.u - aR, Rfast = R.u + aR;
    T l1 = us - as, l4 = us + as;
    harten(Rslow, Rfast, Lslow, Lfast, l1, l4);
    const T Lplus = (Rfast >= l4) ? Rfast : l4;      // np.maximum(R_m, Lm)
    const T Lminus = (Lslow <= l1) ? Lslow : l1;     // np.minimum(L_p, Lp)

    // ConservativeState(state=W): PrimitiveConverter.to_conservative (concrete_defs.py:126-141)
    const T ekL = (0.5 * L.r) * (L.u * L.u + L.v * L.v);
    const T ekR = (0.5 * R.r) * (R.u * R.u + R.v * R.v);
    const T UL0 = L.r, UL1 = L.r * L.u, UL2 = L.r * L.v, UL3 = L.p / ph.gm1 + ekL;
    const T UR0 = R.r, UR1 = R.r * R.u, UR2 = R.r * R.v, UR3 = R.p / ph.gm1 + ekR;
    // PrimitiveState.F(U=...) (states/primitive.py:203-209)
    const T FL0 = UL1, FL1 = UL1 * L.u + L.p, FL2 = UL1 * L.v, FL3 = L.u * (UL3 + L.p);
    const T FR0 = UR1, FR1 = UR1 * R.u + R.p, FR2 = UR1 * R.v, FR3 = R.u * (UR3 + R.p);
    Q4T<T> F;
    if (Lminus >= 0.0) {
        F.v[0] = FL0; F.v[1] = FL1; F.v[2] = FL2; F.v[3] = FL3;
    } else if (Lplus <= 0.0) {
        F.v[0] = FR0; F.v[1] = FR1; F.v[2] = FR2; F.v[3] = FR3;
    } else {
        const T dU0 = UR0 - UL0, dU1 = UR1 - UL1, dU2 = UR2 - UL2, dU3 = UR3 - UL3;
        const T dF0 = FR0 - FL0, dF1 = FR1 - FL1, dF2 = FR2 - FL2, dF3 = FR3 - FL3;
        const T k = as * nrm4(dU0, dU1, dU2, dU3);
        const T n = nrm4(dF0 - us * dU0, dF1 - us * dU1, dF2 - us * dU2, dF3 - us * dU3);
        const T den = (k < 1e-16) ? T(k + 1e-14) : k;
        T alpha = 1.0 - n / den;
        alpha = (0.0 >= alpha) ? T(0.0) : alpha;              // np.maximum(0, ...)
        const T ra = us / Lminus, rb = us / Lplus;
        const T mx = (ra >= rb) ? ra : rb;            // np.maximum(u/Lm, u/Lp)
        const T coef = Lminus * Lplus * (1.0 - alpha * (1.0 - mx));
        const T w = Lplus - Lminus;
        F.v[0] = (Lplus * FL0 - Lminus * FR0 + coef * dU0) / w;
        F.v[1] = (Lplus * FL1 - Lminus * FR1 + coef * dU1) / w;
        F.v[2] = (Lplus * FL2 - Lminus * FR2 + coef * dU2) / w;
        F.v[3] = (Lplus * FL3 - Lminus * FR3 + coef * dU3) / w;
    }
    return F;
}

template <int FLUX, class T>
__host__ __device__ inline Q4T<T> face_flux(const Q4T<T>& qL, const Q4T<T>& qR, const Phys& ph) {
    const PrimT<T> L = cons_to_prim(qL, ph);
    const PrimT<T> R = cons_to_prim(qR, ph);
    if (FLUX == 0) return roe_flux(L, R, ph);
    return hlll_flux(L, R, ph);
}

template <class T>
__host__ __device__ inline PrimT<T> as_prim(const Q4T<T>& q) { return PrimT<T>{q.v[0], q.v[1], q.v[2], q.v[3]}; }
template <class T>
__host__ __device__ inline Q4T<T> prim_q4(const PrimT<T>& w) {
    Q4T<T> q;
    q.v[0] = w.r; q.v[1] = w.u; q.v[2] = w.v; q.v[3] = w.p;
    return q;
}

// face_flux on face states in the reconstruction variables: with PrimitiveState reconstruction the
// states are primitive already and go to the flux as they are (fvm/base.py:286-289).
template <int FLUX, class T>
__host__ __device__ inline Q4T<T> face_flux_recon(const Q4T<T>& qL, const Q4T<T>& qR, const Phys& ph) {
    if (!ph.prim) return face_flux<FLUX>(qL, qR, ph);
    if (FLUX == 0) return roe_flux(as_prim(qL), as_prim(qR), ph);
    return hlll_flux(as_prim(qL), as_prim(qR), ph);
}

// Geometry of one cell as seen by the reconstruction: GG face weights x/yl = L*n (per face),
// 1/A, and the quadrature-point offsets from the centroid (dxq = xq - xc, dyq = yq - yc).
struct CellGeo {
    double xlE, xlW, xlN, xlS, ylE, ylW, ylN, ylS;
    double ainv;
    double dxE, dxW, dxN, dxS, dyE, dyW, dyN, dyS;
};

template <class T>
struct ReconT {
    Q4T<T> E, W, N, S;  // limited face states (SecondOrderMUSCL.limited_solution_at_quadrature_point)
};
using Recon = ReconT<double>;

// np.minimum of two scalars: a NaN operand is returned (NaN-propagating, unlike fmin).  Written as
// three selects of plain values: nested conditionals become branches, and the compiler then moves
// the evaluation of b (a whole limiter chain) into the branch where a is not NaN, serialising the
// faces' divisions.
template <class T>
__host__ __device__ inline T np_minimum(T a, T b) {
    const bool a_nan = a != a, b_nan = b != b;
    const T lo = (a <= b) ? a : b;
    const T r = b_nan ? b : lo;
    return a_nan ? a : r;
}

// np.minimum.reduce of the four faces' limiter values in the order E, W, N, S (the first NaN if
// any).  Venkatakrishnan values (GEN = false) are never -0: the numerator s*s + 2*s is +0 where it
// sums to zero (a zero sum of two doubles rounds to +0), the denominator is positive, and the
// flat-face value is 3/4.  Without a NaN the chain's result is then the minimum, which fmin returns
// with the same bits; a lane with a NaN takes the chain (a branch that waves without one skip).
template <bool GEN, class T>
__host__ __device__ inline T limiter_min4(T pE, T pW, T pN, T pS) {
    if constexpr (!GEN && std::is_same<T, double>::value) {
        const bool any_nan = (pE != pE) | (pW != pW) | (pN != pN) | (pS != pS);
        T m = fmin(fmin(pE, pW), fmin(pN, pS));
        if (any_nan) m = np_minimum(np_minimum(np_minimum(pE, pW), pN), pS);
        return m;
    } else {
        return np_minimum(np_minimum(np_minimum(pE, pW), pN), pS);
    }
}

// Slope of one face, SlopeLimiter._compute_slope (limiters/base.py:189-221): dmax/d for d > 0,
// dmin/d for d < 0, 1 otherwise.
template <class T>
__host__ __device__ inline T face_slope(T dmax, T dmin, T d) {
    const T num = (d > 0.0) ? dmax : dmin;
    return (d > 0.0 || d < 0.0) ? num / d : T(1.0);
}

// Limiter functions of limiters/limiters.py:25-62 (slope**2 is s*s).
template <class T>
__host__ __device__ inline T venkatakrishnan_fn(T s) {
    const T s2 = s * s;
    return (s2 + 2.0 * s) / (s2 + s + 2.0);
}
template <class T>
__host__ __device__ inline T limiter_fn(int lim, T s) {
    switch (lim) {
        case 1: return np_minimum(T(1.0), s);                                 // BarthJespersen
        case 2: { const T s2 = s * s; return (s2 + s) / (s2 + 1.0); }        // VanAlbada
        case 3: return (fabs(s) + s) / (s + 1.0);                             // VanLeer
        default: return venkatakrishnan_fn(s);                                 // Venkatakrishnan
    }
}

// Limiter value of one face.  GEN = false: Venkatakrishnan (the default, compiled without the
// limiter switch); GEN = true: the limiter `lim`.
// Limiter value at the default slope s = 1 (d = 0): Venkatakrishnan 3/4, the other three 1.
__host__ __device__ inline double limiter_at_one(bool gen, int lim) {
    return (gen && lim >= 1 && lim <= 3) ? 1.0 : 0.75;
}

// Where d = 0 the slope is the default 1 and the limiter value the constant limiter_fn(lim, 1)
// (3/4 = (1 + 2)/(1 + 1 + 2) exactly, or 1), so neither division is evaluated there: a wave whose
// faces all have d = 0 (a flat region) skips both.
template <bool GEN, class T>
__host__ __device__ inline T limit_face(T dmax, T dmin, T d, int lim) {
    if (!(d > 0.0 || d < 0.0)) return T(limiter_at_one(GEN, lim));
    const T num = (d > 0.0) ? dmax : dmin;
    const T s = num / d;
    return GEN ? limiter_fn(lim, s) : venkatakrishnan_fn(s);
}

// Venkatakrishnan._venkata (limiters/limiters.py:38-49) on the slope of one face.
__host__ __device__ inline double venkat_face(double dmax, double dmin, double d) {
    return limit_face<false>(dmax, dmin, d, 0);
}

// One cell of the second-order reconstruction (reconstruct_var: one conserved variable):
//  * interface averages (blocks/quad_block.py:181-218),
//  * Green-Gauss gradient (gradients/greengauss.py:134-155),
//  * high-order term gx*dx + gy*dy (blocks/base.py:282-289),
//  * slope limiter with the 5-point min/max incl. ghosts (limiters/base.py:47-187): Venkatakrishnan,
//    or with GEN the limiter `lim` (BarthJespersen, VanAlbada, VanLeer),
//  * limited face values U + phi*hot (fvm/SecondOrderMUSCL.py:106-126).
// U is the cell, Uw/Ue/Us/Un its west/east/south/north neighbours (ghost values at block edges).
template <bool GEN = false, class T>
__host__ __device__ inline void reconstruct_var(T u, T uw, T ue, T us, T un, const CellGeo& g, int lim, T& qE,
                                                T& qW, T& qN, T& qS) {
    const T e = 0.5 * (u + ue);
    const T w = 0.5 * (uw + u);
    const T n = 0.5 * (u + un);
    const T s = 0.5 * (us + u);
    const T gx = (e * g.xlE + w * g.xlW + n * g.xlN + s * g.xlS) * g.ainv;
    const T gy = (e * g.ylE + w * g.ylW + n * g.ylN + s * g.ylS) * g.ainv;
    const T hE = gx * g.dxE + gy * g.dyE;
    const T hW = gx * g.dxW + gy * g.dyW;
    const T hN = gx * g.dxN + gy * g.dyN;
    const T hS = gx * g.dxS + gy * g.dyS;
    // np.maximum.reduce((U, UW, UE, US, UN)) - U
    const T mx = fmax(fmax(fmax(fmax(u, uw), ue), us), un);
    const T mn = fmin(fmin(fmin(fmin(u, uw), ue), us), un);
    const T dmax = mx - u;
    const T dmin = mn - u;
    const T pE = limit_face<GEN>(dmax, dmin, (u + hE) - u, lim);
    const T pW = limit_face<GEN>(dmax, dmin, (u + hW) - u, lim);
    const T pN = limit_face<GEN>(dmax, dmin, (u + hN) - u, lim);
    const T pS = limit_face<GEN>(dmax, dmin, (u + hS) - u, lim);
    T phi = limiter_min4<GEN>(pE, pW, pN, pS);
    phi = (phi < 0.0) ? T(0.0) : phi;                  // np.where(phi < 0, 0, phi)
    qE = u + phi * hE;
    qW = u + phi * hW;
    qN = u + phi * hN;
    qS = u + phi * hS;
}

template <bool GEN = false, class T>
__host__ __device__ inline ReconT<T> reconstruct(const Q4T<T>& U, const Q4T<T>& Uw, const Q4T<T>& Ue,
                                                 const Q4T<T>& Us, const Q4T<T>& Un, const CellGeo& g, int lim = 0) {
    ReconT<T> r;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        reconstruct_var<GEN>(U.v[k], Uw.v[k], Ue.v[k], Us.v[k], Un.v[k], g, lim, r.E.v[k], r.W.v[k], r.N.v[k],
                             r.S.v[k]);
    return r;
}

// ---------------------------------------------------------------------------------------------
// nq-point Gauss faces (mesh/quadratures.py:33-95, fvm/base.py:168-190)
// ---------------------------------------------------------------------------------------------
constexpr int QMAX = 3;

// Offsets of the Gauss points of one cell from its centroid, faces E, W, N, S, points in rule order.
struct QGeo {
    double dx[4][QMAX], dy[4][QMAX];
};

// Reconstruction of one cell with nq points per face: the cell value (reconstruction variables),
// the Green-Gauss gradient and the limiter value; the limited value at a point is u + phi*hot.
struct ReconQ {
    Q4 u, gx, gy, phi;
};

// Weight of point i of the nq-point rule (_QUAD_2 = {-1/sqrt(3): 1, 1/sqrt(3): 1},
// _QUAD_3 = {-sqrt(3/5): 5/9, 0: 8/9, sqrt(3/5): 5/9}).
__host__ __device__ inline double gauss_w(int nq, int i) {
    if (nq == 3) return i == 1 ? 8.0 / 9.0 : 5.0 / 9.0;
    return nq == 2 ? 1.0 : 2.0;
}

// One variable of reconstruct_var with nq points per face: interface averages, Green-Gauss, and
// phi = max(0, min over E points, W points, N points, S points of the limiter of each point's slope)
// (limiters/base.py:47-187: np.minimum.reduce over the faces' point lists in that order).
__host__ __device__ inline void reconstruct_q_var(double u, double uw, double ue, double us, double un,
                                                  const CellGeo& g, const QGeo& q, int nq, int lim, double& gx,
                                                  double& gy, double& phi) {
    const double e = 0.5 * (u + ue);
    const double w = 0.5 * (uw + u);
    const double n = 0.5 * (u + un);
    const double s = 0.5 * (us + u);
    gx = (e * g.xlE + w * g.xlW + n * g.xlN + s * g.xlS) * g.ainv;
    gy = (e * g.ylE + w * g.ylW + n * g.ylN + s * g.ylS) * g.ainv;
    const double mx = fmax(fmax(fmax(fmax(u, uw), ue), us), un);
    const double mn = fmin(fmin(fmin(fmin(u, uw), ue), us), un);
    const double dmax = mx - u;
    const double dmin = mn - u;
    double p = 0.0;
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
        for (int i = 0; i < QMAX; ++i) {
            if (i >= nq) continue;
            const double h = gx * q.dx[f][i] + gy * q.dy[f][i];
            const double pf = limit_face<true>(dmax, dmin, (u + h) - u, lim);
            p = (f == 0 && i == 0) ? pf : np_minimum(p, pf);
        }
    phi = (p < 0.0) ? 0.0 : p;
}

__host__ __device__ inline ReconQ reconstruct_q(const Q4& U, const Q4& Uw, const Q4& Ue, const Q4& Us, const Q4& Un,
                                                const CellGeo& g, const QGeo& q, int nq, int lim) {
    ReconQ r;
    r.u = U;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        reconstruct_q_var(U.v[k], Uw.v[k], Ue.v[k], Us.v[k], Un.v[k], g, q, nq, lim, r.gx.v[k], r.gy.v[k], r.phi.v[k]);
    return r;
}

// Limited value at point i of face f: U + phi*(gx*dxq + gy*dyq)
// (SecondOrderMUSCL.limited_solution_at_quadrature_point, fvm/SecondOrderMUSCL.py:106-126).
__host__ __device__ inline Q4 qstate(const ReconQ& r, const QGeo& q, int f, int i) {
    Q4 o;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double h = r.gx.v[k] * q.dx[f][i] + r.gy.v[k] * q.dy[f][i];
        o.v[k] = r.u.v[k] + r.phi.v[k] * h;
    }
    return o;
}

// Ideal-gas CFL quantity min(dx/(|u|+a), dy/(|v|+a)) of one cell (blocks/quad_block.py:433-436;
// ConservativeState.u/v/p/a, states/conservative.py:85-137).
__host__ __device__ inline double cell_dt(const Q4& q, double dx, double dy, const Phys& ph) {
    const double rho = q.v[0];
    const double u = q.v[1] / rho;
    const double v = q.v[2] / rho;
    const double p = ph.gm1 * (q.v[3] - rho * (0.5 * (u * u + v * v)));
    const double a = sqrt_rn(ph.gamma * p / rho);
    const double tx = dx / (fabs(u) + a);
    const double ty = dy / (fabs(v) + a);
    return (tx <= ty) ? tx : ty;
}

}  // namespace phx
