// Host-callable wrappers around the device numerics (pyhype_amd/csrc/phx_numerics.hpp) so the CPU
// test suite can compare the exact kernel arithmetic with the oracle.  Test infrastructure only.
#include "../../pyhype_amd/csrc/phx_numerics.hpp"

using namespace phx;

static Phys make_phys(const double* p) { return Phys{p[0], p[1], p[2], p[3], p[4]}; }
static Q4 q4(const double* p) { Q4 q; for (int k = 0; k < 4; ++k) q.v[k] = p[k]; return q; }
static void put(double* p, const Q4& q) { for (int k = 0; k < 4; ++k) p[k] = q.v[k]; }

extern "C" {
// primitive face states (n x 4) -> flux (n x 4); flux 0 = Roe, 1 = HLLL
void shim_prim_flux(int flux, int n, const double* WL, const double* WR, double* F, const double* phys) {
    const Phys ph = make_phys(phys);
    for (int i = 0; i < n; ++i) {
        Prim L{WL[4 * i], WL[4 * i + 1], WL[4 * i + 2], WL[4 * i + 3]};
        Prim R{WR[4 * i], WR[4 * i + 1], WR[4 * i + 2], WR[4 * i + 3]};
        put(F + 4 * i, flux == 0 ? roe_flux(L, R, ph) : hlll_flux(L, R, ph));
    }
}
void shim_cons_to_prim(int n, const double* U, double* W, const double* phys) {
    const Phys ph = make_phys(phys);
    for (int i = 0; i < n; ++i) {
        Prim w = cons_to_prim(q4(U + 4 * i), ph);
        W[4 * i] = w.r; W[4 * i + 1] = w.u; W[4 * i + 2] = w.v; W[4 * i + 3] = w.p;
    }
}
// limiter value of each face (n x 4: E, W, N, S) -> their minimum: the Venkatakrishnan path
// (limiter_min4<false>) and the np.minimum chain (limiter_min4<true>)
void shim_limiter_min4(int n, const double* p, double* fast, double* chain) {
    for (int i = 0; i < n; ++i) {
        const double* q = p + 4 * i;
        fast[i] = limiter_min4<false>(q[0], q[1], q[2], q[3]);
        chain[i] = limiter_min4<true>(q[0], q[1], q[2], q[3]);
    }
}
void shim_venkat_face(int n, const double* dmax, const double* dmin, const double* d, double* out) {
    for (int i = 0; i < n; ++i) out[i] = venkat_face(dmax[i], dmin[i], d[i]);
}
void shim_reflect(int n, const double* U, const double* c, const double* s, double* out) {
    for (int i = 0; i < n; ++i) put(out + 4 * i, reflect(q4(U + 4 * i), c[i], s[i]));
}
// g: 17 doubles per cell in CellGeo order; U, Uw, Ue, Us, Un: n x 4; out: n x 16 (E, W, N, S)
void shim_reconstruct(int n, const double* U, const double* Uw, const double* Ue, const double* Us,
                      const double* Un, const double* g, double* out) {
    for (int i = 0; i < n; ++i) {
        const double* q = g + 17 * i;
        CellGeo G{q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], q[8], q[9], q[10], q[11], q[12], q[13], q[14], q[15], q[16]};
        Recon r = reconstruct(q4(U + 4 * i), q4(Uw + 4 * i), q4(Ue + 4 * i), q4(Us + 4 * i), q4(Un + 4 * i), G);
        put(out + 16 * i, r.E); put(out + 16 * i + 4, r.W); put(out + 16 * i + 8, r.N); put(out + 16 * i + 12, r.S);
    }
}
// as shim_reconstruct with limiter `lim` (0 Venkatakrishnan, 1 BarthJespersen, 2 VanAlbada, 3 VanLeer)
void shim_reconstruct_lim(int lim, int n, const double* U, const double* Uw, const double* Ue, const double* Us,
                          const double* Un, const double* g, double* out) {
    for (int i = 0; i < n; ++i) {
        const double* q = g + 17 * i;
        CellGeo G{q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], q[8], q[9], q[10], q[11], q[12], q[13], q[14], q[15], q[16]};
        Recon r = reconstruct<true>(q4(U + 4 * i), q4(Uw + 4 * i), q4(Ue + 4 * i), q4(Us + 4 * i), q4(Un + 4 * i), G, lim);
        put(out + 16 * i, r.E); put(out + 16 * i + 4, r.W); put(out + 16 * i + 8, r.N); put(out + 16 * i + 12, r.S);
    }
}
// Euclidean norm of n 4-vectors as HLLL's alpha computes it (nrm4)
void shim_nrm4(int n, const double* X, double* out) {
    for (int i = 0; i < n; ++i) out[i] = nrm4(X[4 * i], X[4 * i + 1], X[4 * i + 2], X[4 * i + 3]);
}
// the same norm through the 64-bit integer emulation alone (nrm4's fallback path)
void shim_nrm4_ext(int n, const double* X, double* out) {
    for (int i = 0; i < n; ++i) out[i] = nrm4_ext(X[4 * i], X[4 * i + 1], X[4 * i + 2], X[4 * i + 3]);
}
// nrm4's fast-path sum of squares RN64(RN64(RN64(a^2 + b^2) + c^2) + d^2) as hi + lo, and whether
// the pair arithmetic decided every rounding (ok = 1) or nrm4 would take the emulation (ok = 0)
void shim_sum64(int n, const double* X, double* hi, double* lo, int* okv) {
    for (int i = 0; i < n; ++i) {
        const double* x = X + 4 * i;
        bool ok = nrm4_in_range(x[0]) && nrm4_in_range(x[1]) && nrm4_in_range(x[2]) && nrm4_in_range(x[3]);
        DD s{0.0, 0.0};
        if (ok) {
            s = sq64(x[0], ok);
            for (int k = 1; k < 4; ++k) s = add64(s, sq64(x[k], ok), ok);
        }
        hi[i] = s.hi;
        lo[i] = s.lo;
        okv[i] = ok ? 1 : 0;
    }
}
void shim_cell_dt(int n, const double* U, const double* dx, const double* dy, double* out, const double* phys) {
    const Phys ph = make_phys(phys);
    for (int i = 0; i < n; ++i) out[i] = cell_dt(q4(U + 4 * i), dx[i], dy[i], ph);
}
}
