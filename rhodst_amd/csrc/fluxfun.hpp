// fluxfun.hpp -- the six run-time-selected inviscid face fluxes of rhoDST
// (fluxSchemes/*/ *DST.C computeFluxDST), written once as templates on the
// scalar type: T = double in the face loop, T = D1 (value + one tangent) in
// the exact linearisation of the adjoint.  Masks (foam pos/neg/sign) act on
// the value and carry no derivative.
#pragma once
#include <hip/hip_runtime.h>

#include "rdst_internal.hpp"

// No floating-point contraction (see flux.hip): the reference's rounded
// products are reproduced, so x*y - x*y is an exact zero.
#pragma clang fp contract(off)

namespace rdst {

// forward-mode dual number: value and one directional derivative
struct D1 {
    double v, d;
    __device__ D1() : v(0.0), d(0.0) {}
    __device__ D1(double x) : v(x), d(0.0) {}
    __device__ D1(double x, double dx) : v(x), d(dx) {}
};
__device__ inline D1 operator+(D1 a, D1 b) { return D1(a.v + b.v, a.d + b.d); }
__device__ inline D1 operator-(D1 a, D1 b) { return D1(a.v - b.v, a.d - b.d); }
__device__ inline D1 operator-(D1 a) { return D1(-a.v, -a.d); }
__device__ inline D1 operator*(D1 a, D1 b) { return D1(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ inline D1 operator/(D1 a, D1 b) {
    const double q = a.v / b.v;
    return D1(q, (a.d - q * b.d) / b.v);
}
__device__ inline D1& operator+=(D1& a, D1 b) { a = a + b; return a; }
__device__ inline D1& operator-=(D1& a, D1 b) { a = a - b; return a; }
__device__ inline D1& operator*=(D1& a, D1 b) { a = a * b; return a; }
// the double forms next to the D1 ones (unqualified calls inside namespace rdst
// would otherwise find only the D1 overloads)
__device__ inline double sqrt(double a) { return ::sqrt(a); }
__device__ inline double fabs(double a) { return ::fabs(a); }
__device__ inline double fmax(double a, double b) { return ::fmax(a, b); }
__device__ inline double fmin(double a, double b) { return ::fmin(a, b); }
__device__ inline D1 sqrt(D1 a) {
    const double s = ::sqrt(a.v);
    return D1(s, s > 0.0 ? a.d / (2.0 * s) : 0.0);
}
__device__ inline D1 fabs(D1 a) { return a.v >= 0.0 ? a : -a; }
__device__ inline D1 fmax(D1 a, D1 b) { return a.v >= b.v ? a : b; }
__device__ inline D1 fmin(D1 a, D1 b) { return a.v <= b.v ? a : b; }
__device__ inline double val(double s) { return s; }
__device__ inline double val(const D1& s) { return s.v; }
__device__ inline double dval(double) { return 0.0; }
__device__ inline double dval(const D1& s) { return s.d; }

constexpr double kSMALL = 1.0e-15;
constexpr double kVSMALL = 1.0e-300;

template <class T>
struct SideT {
    T rho, rux, ruy, ruz, rPsi, e, aStar, c;
    T rr, ux, uy, uz;  // 1/rho and the velocity, formed once per side
};
template <class T>
struct Flux5T {
    T phi, fx, fy, fz, fE, upx, upy, upz;
};
using Side = SideT<double>;
using Flux5 = Flux5T<double>;

// foam masks on the value (piecewise constant: no derivative)
template <class T> __device__ inline T fsign(const T& s) { return val(s) >= 0.0 ? T(1.0) : T(-1.0); }
template <class T> __device__ inline T dpos(const T& s) { return val(s) > 0.0 ? T(1.0) : T(0.0); }
template <class T> __device__ inline T dpos0(const T& s) { return val(s) >= 0.0 ? T(1.0) : T(0.0); }
template <class T> __device__ inline T dneg(const T& s) { return val(s) < 0.0 ? T(1.0) : T(0.0); }
template <class T> __device__ inline T dneg0(const T& s) { return val(s) <= 0.0 ? T(1.0) : T(0.0); }
template <class T> __device__ inline T sq(const T& a) { return a * a; }

// Side state from the thermo fields: rPsi = 1/psi, e = h - p/rho
// (AUSMPLDST.C:141-148), critical sound speed a* from
// H = max(rhoE/rho, SMALL) + max(p/rho, SMALL) (AUSMPLDST.C:156-161),
// c = sqrt(Cp/Cv rPsi) (KurganovDST.C:129).
template <class T>
__device__ inline SideT<T> make_side(T rho, T rux, T ruy, T ruz, T rhoE, T p, T h, T psi, double gamma) {
    SideT<T> s;
    s.rho = rho; s.rux = rux; s.ruy = ruy; s.ruz = ruz;
    s.rr = 1.0 / rho;
    s.ux = rux * s.rr; s.uy = ruy * s.rr; s.uz = ruz * s.rr;
    s.rPsi = 1.0 / psi;
    const T pr = p * s.rr;
    s.e = h - pr;
    T H = fmax(rhoE * s.rr, T(kSMALL)) + fmax(pr, T(kSMALL));
    s.aStar = sqrt(2.0 * (gamma - 1.0) / (gamma + 1.0) * H);
    s.c = sqrt(gamma * s.rPsi);
    return s;
}

// the packed face-loop record of cell i (Fields::cell8), from its state
__device__ inline void store_cell8(const Fields& F, int i, double rho, double rux, double ruy, double ruz, double rhoE,
                                   double p, double h, double psi, double gamma) {
    double2* o = reinterpret_cast<double2*>(F.cell8 + 8 * (size_t)i);
    if (F.cell8Kind == CELL8_SIDE) {
        const SideT<double> sd = make_side<double>(rho, rux, ruy, ruz, rhoE, p, h, psi, gamma);
        o[0] = make_double2(sd.rho, sd.ux);
        o[1] = make_double2(sd.uy, sd.uz);
        o[2] = make_double2(sd.rr, sd.rPsi);
        o[3] = make_double2(sd.e, sd.aStar);
    } else {
        o[0] = make_double2(rho, rux);
        o[1] = make_double2(ruy, ruz);
        o[2] = make_double2(rhoE, p);
        o[3] = make_double2(h, psi);
    }
}

// AUSM+ split Mach/pressure polynomials with the foam sign() masks
// (sign(0) = +1, so at |M| = 1 both branches contribute).
template <class T>
__device__ inline void ausm_split(T Ma_own, T Ma_nei, T alpha, double beta, T& Ma4_own, T& P5_own, T& Ma4_nei,
                                  T& P5_nei) {
    T s1o = fsign(1.0 - fabs(Ma_own)), s2o = fsign(-1.0 + fabs(Ma_own));
    T s1n = fsign(1.0 - fabs(Ma_nei)), s2n = fsign(-1.0 + fabs(Ma_nei));
    T m1o = 0.5 * (s1o + fabs(s1o)), m2o = 0.5 * (s2o + fabs(s2o));
    T m1n = 0.5 * (s1n + fabs(s1n)), m2n = 0.5 * (s2n + fabs(s2n));
    Ma4_own = m1o * (0.25 * sq(Ma_own + 1.0)) * (1.0 - 16.0 * beta * (-0.25) * sq(Ma_own - 1.0)) + m2o * fmax(Ma_own, T(0.0));
    P5_own = m1o * 0.25 * sq(Ma_own + 1.0) * (2.0 - Ma_own - 16.0 * alpha * Ma_own * (-0.25) * sq(Ma_own - 1.0)) + m2o * dpos(Ma_own);
    Ma4_nei = m1n * (-0.25) * sq(Ma_nei - 1.0) * (1.0 + 16.0 * beta * 0.25 * sq(Ma_nei + 1.0)) + m2n * fmin(Ma_nei, T(0.0));
    P5_nei = m1n * (-0.25) * sq(Ma_nei - 1.0) * (-2.0 - Ma_nei + 16.0 * alpha * Ma_nei * 0.25 * sq(Ma_nei + 1.0)) + m2n * dneg(Ma_nei);
}

// AUSM+ (AUSMPLDST.C:127-217) and AUSM+up (AUSMPLUPDST.C:209-320)
template <bool UP, class T>
__device__ inline Flux5T<T> flux_ausm(const SideT<T>& L, const SideT<T>& R, double sx, double sy, double sz, double ms,
                                      int lowMach, double Mainf) {
    const double beta = 0.125;
    const T rms = 1.0 / ms;
    T nx = sx * rms, ny = sy * rms, nz = sz * rms;
    T uox = L.ux, uoy = L.uy, uoz = L.uz;
    T unx = R.ux, uny = R.uy, unz = R.uz;
    T Uv_own = uox * nx + uoy * ny + uoz * nz;
    T Uv_nei = unx * nx + uny * ny + unz * nz;
    T rhoE_own = L.rho * (L.e + 0.5 * (uox * uox + uoy * uoy + uoz * uoz));
    T rhoE_nei = R.rho * (R.e + 0.5 * (unx * unx + uny * uny + unz * unz));
    T p_own = L.rho * L.rPsi, p_nei = R.rho * R.rPsi;
    T a12, alpha = 3.0 / 16.0, fa = 0.0, Mmean2 = 0.0;
    if (!UP) {
        a12 = 0.5 * (L.aStar + R.aStar);
    } else {
        T a_own = sq(L.aStar) / fmax(L.aStar, Uv_own);
        T a_nei = sq(R.aStar) / fmax(R.aStar, -Uv_nei);
        a12 = fmin(a_own, a_nei);
        alpha = 0.0;  // only set by the lowMach branch (AUSMPLUPDST.C:118-131, :257-263)
        if (lowMach) {
            Mmean2 = 0.5 * (sq(Uv_own) + sq(Uv_nei)) / sq(a12);
            T M0 = sqrt(fmin(1.0, fmax(Mmean2, sq(Mainf))));
            fa = M0 * (2.0 - M0);
            alpha = 3.0 / 16.0 * (-4.0 + 5.0 * sq(fa));
        }
    }
    T Ma_own = Uv_own / a12, Ma_nei = Uv_nei / a12;
    T Ma4_own, P5_own, Ma4_nei, P5_nei;
    ausm_split(Ma_own, Ma_nei, alpha, beta, Ma4_own, P5_own, Ma4_nei, P5_nei);
    T Ma12 = Ma4_own + Ma4_nei;
    T P12 = P5_own * p_own + P5_nei * p_nei;
    if (UP && lowMach) {
        T Mp = -0.25 / fa * fmax(1.0 - 1.0 * Mmean2, 0.0) * (p_nei - p_own) / (0.5 * (L.rho + R.rho)) / sq(a12);
        T pu = -0.75 * P5_own * P5_nei * (L.rho + R.rho) * a12 * fa * (Uv_own - Uv_nei);
        Ma12 += Mp;
        P12 += pu;
    }
    T ps = dpos(Ma12), ng = dneg0(Ma12);
    Flux5T<T> o;
    o.upx = ps * uox + ng * unx;
    o.upy = ps * uoy + ng * uny;
    o.upz = ps * uoz + ng * unz;
    o.phi = Ma12 * a12 * (ps * L.rho + ng * R.rho) * ms;
    o.fx = o.phi * o.upx + P12 * sx;
    o.fy = o.phi * o.upy + P12 * sy;
    o.fz = o.phi * o.upz + P12 * sz;
    T H_own = (rhoE_own + p_own) * L.rr, H_nei = (rhoE_nei + p_nei) * R.rr;
    o.fE = o.phi * (ps * H_own + ng * H_nei);
    return o;
}

// HLL (HLLDST.C:91-204)
template <class T>
__device__ inline Flux5T<T> flux_hll(const SideT<T>& L, const SideT<T>& R, double sx, double sy, double sz, double ms,
                                     double gamma) {
    T uox = L.ux, uoy = L.uy, uoz = L.uz;
    T unx = R.ux, uny = R.uy, unz = R.uz;
    T E_own = L.e + 0.5 * (uox * uox + uoy * uoy + uoz * uoz);
    T E_nei = R.e + 0.5 * (unx * unx + uny * uny + unz * unz);
    T p_own = L.rho * L.rPsi, p_nei = R.rho * R.rPsi;
    T H_own = E_own + p_own * L.rr, H_nei = E_nei + p_nei * R.rr;
    T coefR = sqrt(fmax(T(kVSMALL), R.rho) / fmax(T(kVSMALL), L.rho));
    T HT = (coefR * H_nei + H_own) / (coefR + 1.0);
    T utx = (coefR * unx + uox) / (coefR + 1.0), uty = (coefR * uny + uoy) / (coefR + 1.0), utz = (coefR * unz + uoz) / (coefR + 1.0);
    T aT = sqrt((gamma - 1.0) * (HT - 0.5 * (utx * utx + uty * uty + utz * utz)));
    const T rms = 1.0 / ms;
    T nx = sx * rms, ny = sy * rms, nz = sz * rms;
    T Uv_own = uox * nx + uoy * ny + uoz * nz, Uv_nei = unx * nx + uny * ny + unz * nz;
    T UvT = utx * nx + uty * ny + utz * nz;
    T S_own = fmin(Uv_own - L.aStar, UvT - aT);
    T S_nei = fmax(Uv_nei + R.aStar, UvT + aT);
    T cf_own = dpos0(S_own), cf_m = dneg(S_own) * dpos0(S_nei), cf_nei = dneg(S_nei) * dneg(S_own);
    T dS = S_nei - S_own;
    T ro = L.rho, rn = R.rho;
    T mo_x = ro * uox * Uv_own + p_own * nx, mo_y = ro * uoy * Uv_own + p_own * ny, mo_z = ro * uoz * Uv_own + p_own * nz;
    T mn_x = rn * unx * Uv_nei + p_nei * nx, mn_y = rn * uny * Uv_nei + p_nei * ny, mn_z = rn * unz * Uv_nei + p_nei * nz;
    Flux5T<T> o;
    o.phi = (cf_own * ro * Uv_own + cf_m * (S_nei * ro * Uv_own - S_own * rn * Uv_nei + S_own * S_nei * (rn - ro)) / dS
             + cf_nei * rn * Uv_nei) * ms;
    o.fx = (cf_own * mo_x + cf_m * ((S_nei * mo_x - S_own * mn_x + S_own * S_nei * (rn * unx - ro * uox)) / dS) + cf_nei * mn_x) * ms;
    o.fy = (cf_own * mo_y + cf_m * ((S_nei * mo_y - S_own * mn_y + S_own * S_nei * (rn * uny - ro * uoy)) / dS) + cf_nei * mn_y) * ms;
    o.fz = (cf_own * mo_z + cf_m * ((S_nei * mo_z - S_own * mn_z + S_own * S_nei * (rn * unz - ro * uoz)) / dS) + cf_nei * mn_z) * ms;
    o.fE = (cf_own * ro * Uv_own * H_own
            + cf_m * (S_nei * ro * Uv_own * H_own - S_own * rn * Uv_nei * H_nei + S_own * S_nei * (rn * E_nei - ro * E_own)) / dS
            + cf_nei * rn * Uv_nei * H_nei) * ms;
    o.upx = cf_own * uox + cf_m * ((S_nei * uox - S_own * unx + Uv_own * uox - Uv_nei * unx) / dS) + cf_nei * unx;
    o.upy = cf_own * uoy + cf_m * ((S_nei * uoy - S_own * uny + Uv_own * uoy - Uv_nei * uny) / dS) + cf_nei * uny;
    o.upz = cf_own * uoz + cf_m * ((S_nei * uoz - S_own * unz + Uv_own * uoz - Uv_nei * unz) / dS) + cf_nei * unz;
    return o;
}

// HLLC (HLLCDST.C:135-272, directWaveSpeed) and HLLCP (HLLCPDST.C:166-317)
template <bool P, class T>
__device__ inline Flux5T<T> flux_hllc(const SideT<T>& L, const SideT<T>& R, double sx, double sy, double sz, double ms,
                                      double gamma, double f) {
    T uox = L.ux, uoy = L.uy, uoz = L.uz;
    T unx = R.ux, uny = R.uy, unz = R.uz;
    T E_own = L.e + 0.5 * (uox * uox + uoy * uoy + uoz * uoz);
    T E_nei = R.e + 0.5 * (unx * unx + uny * uny + unz * unz);
    T p_own = L.rho * L.rPsi, p_nei = R.rho * R.rPsi;
    T H_own = E_own + p_own * L.rr, H_nei = E_nei + p_nei * R.rr;
    const T rms = 1.0 / ms;
    T nx = sx * rms, ny = sy * rms, nz = sz * rms;
    T Uv_own = uox * nx + uoy * ny + uoz * nz, Uv_nei = unx * nx + uny * ny + unz * nz;
    T coefR = sqrt(fmax(T(kVSMALL), R.rho) / fmax(T(kVSMALL), L.rho));
    T HT = (coefR * H_nei + H_own) / (coefR + 1.0);
    T utx = (coefR * unx + uox) / (coefR + 1.0), uty = (coefR * uny + uoy) / (coefR + 1.0), utz = (coefR * unz + uoz) / (coefR + 1.0);
    T UT2 = utx * utx + uty * uty + utz * utz;
    T aT = sqrt((gamma - 1.0) * (HT - 0.5 * UT2));
    T UvT = utx * nx + uty * ny + utz * nz;
    T S_own = fmin(Uv_own - L.aStar, UvT - aT);
    T S_nei = fmax(Uv_nei + R.aStar, UvT + aT);
    T ro = L.rho, rn = R.rho;
    T S_star = (p_nei - p_own + ro * Uv_own * (S_own - Uv_own) - rn * Uv_nei * (S_nei - Uv_nei))
                  / (ro * (S_own - Uv_own) - rn * (S_nei - Uv_nei));
    T cf_own = dpos(S_own);
    T cf_m1 = dneg0(S_own) * dpos(S_star);
    T cf_m2 = dneg0(S_own) * dneg0(S_star) * dpos0(S_nei);
    T cf_nei = dneg0(S_own) * dneg0(S_star) * dneg(S_nei);
    T rhoE_own = ro * E_own, rhoE_nei = rn * E_nei;
    T mo_x = L.rux * Uv_own + p_own * nx, mo_y = L.ruy * Uv_own + p_own * ny, mo_z = L.ruz * Uv_own + p_own * nz;
    T mn_x = R.rux * Uv_nei + p_nei * nx, mn_y = R.ruy * Uv_nei + p_nei * ny, mn_z = R.ruz * Uv_nei + p_nei * nz;
    T eo = (rhoE_own + p_own) * Uv_own, en = (rhoE_nei + p_nei) * Uv_nei;
    T d1 = S_own - S_star, d2 = S_nei - S_star;
    Flux5T<T> o;
    if (!P) {
        T ps_o = p_own + ro * (S_own - Uv_own) * (S_star - Uv_own);
        T ps_n = p_nei + rn * (S_nei - Uv_nei) * (S_star - Uv_nei);
        o.phi = (cf_own * ro * Uv_own + cf_m1 * S_star * ro * (S_own - Uv_own) / d1
                 + cf_m2 * S_star * rn * (S_nei - Uv_nei) / d2 + cf_nei * rn * Uv_nei) * ms;
        o.fx = (cf_own * mo_x + cf_m1 * ((S_star * (S_own * L.rux - mo_x) + S_own * ps_o * nx) / d1)
                + cf_m2 * ((S_star * (S_nei * R.rux - mn_x) + S_nei * ps_n * nx) / d2) + cf_nei * mn_x) * ms;
        o.fy = (cf_own * mo_y + cf_m1 * ((S_star * (S_own * L.ruy - mo_y) + S_own * ps_o * ny) / d1)
                + cf_m2 * ((S_star * (S_nei * R.ruy - mn_y) + S_nei * ps_n * ny) / d2) + cf_nei * mn_y) * ms;
        o.fz = (cf_own * mo_z + cf_m1 * ((S_star * (S_own * L.ruz - mo_z) + S_own * ps_o * nz) / d1)
                + cf_m2 * ((S_star * (S_nei * R.ruz - mn_z) + S_nei * ps_n * nz) / d2) + cf_nei * mn_z) * ms;
        o.fE = (cf_own * eo + cf_m1 * S_star * (S_own * rhoE_own - eo + S_own * ps_o) / d1
                + cf_m2 * S_star * (S_nei * rhoE_nei - en + S_nei * ps_n) / d2 + cf_nei * en) * ms;
    } else {
        T Ma_own = Uv_own / L.aStar, Ma_nei = Uv_nei / R.aStar;
        T MaT = (coefR * Ma_nei + Ma_own) / (coefR + 1.0);
        T c_own = ro * (S_own - Uv_own), c_nei = rn * (S_nei - Uv_nei);
        T p_star = (c_nei * p_own - c_own * p_nei - c_own * c_nei * (Uv_own - Uv_nei)) / (c_nei - c_own);
        T theta = fmin(fmax(fabs(Ma_own), fabs(Ma_nei)), 1.0);
        T p_av = 0.5 * (p_own + p_nei);
        T p_ss = p_star * theta + (1.0 - theta) * p_av;
        T p_sss = f * p_ss + (1.0 - f) * p_star;
        T phip = (f - 1.0) * S_own * S_nei / (S_nei - S_own) / (1.0 + fabs(MaT)) * (p_nei - p_own) / sq(aT);
        o.phi = (cf_own * ro * Uv_own + cf_m1 * (S_star * ro * (S_own - Uv_own) / d1 + phip)
                 + cf_m2 * (S_star * rn * (S_nei - Uv_nei) / d2 + phip) + cf_nei * rn * Uv_nei) * ms;
        o.fx = (cf_own * mo_x + cf_m1 * ((S_star * (S_own * L.rux - mo_x) + S_own * p_sss * nx) / d1 + phip * utx)
                + cf_m2 * ((S_star * (S_nei * R.rux - mn_x) + S_nei * p_sss * nx) / d2 + phip * utx) + cf_nei * mn_x) * ms;
        o.fy = (cf_own * mo_y + cf_m1 * ((S_star * (S_own * L.ruy - mo_y) + S_own * p_sss * ny) / d1 + phip * uty)
                + cf_m2 * ((S_star * (S_nei * R.ruy - mn_y) + S_nei * p_sss * ny) / d2 + phip * uty) + cf_nei * mn_y) * ms;
        o.fz = (cf_own * mo_z + cf_m1 * ((S_star * (S_own * L.ruz - mo_z) + S_own * p_sss * nz) / d1 + phip * utz)
                + cf_m2 * ((S_star * (S_nei * R.ruz - mn_z) + S_nei * p_sss * nz) / d2 + phip * utz) + cf_nei * mn_z) * ms;
        o.fE = (cf_own * eo + cf_m1 * (S_star * (S_own * rhoE_own - eo + S_own * p_star) / d1 + 0.5 * phip * UT2)
                + cf_m2 * (S_star * (S_nei * rhoE_nei - en + S_nei * p_star) / d2 + 0.5 * phip * UT2) + cf_nei * en) * ms;
    }
    o.upx = cf_own * uox + cf_m1 * (S_own - Uv_own) * uox / d1 + cf_m2 * (S_nei - Uv_nei) * unx / d2 + cf_nei * unx;
    o.upy = cf_own * uoy + cf_m1 * (S_own - Uv_own) * uoy / d1 + cf_m2 * (S_nei - Uv_nei) * uny / d2 + cf_nei * uny;
    o.upz = cf_own * uoz + cf_m1 * (S_own - Uv_own) * uoz / d1 + cf_m2 * (S_nei - Uv_nei) * unz / d2 + cf_nei * unz;
    return o;
}

// Kurganov-Tadmor central-upwind (KurganovDST.C:93-167)
template <class T>
__device__ inline Flux5T<T> flux_kt(const SideT<T>& L, const SideT<T>& R, double sx, double sy, double sz, double ms) {
    T uox = L.ux, uoy = L.uy, uoz = L.uz;
    T unx = R.ux, uny = R.uy, unz = R.uz;
    T rhoE_own = L.rho * (L.e + 0.5 * (uox * uox + uoy * uoy + uoz * uoz));
    T rhoE_nei = R.rho * (R.e + 0.5 * (unx * unx + uny * uny + unz * unz));
    T p_own = L.rho * L.rPsi, p_nei = R.rho * R.rPsi;
    T phiv_own = uox * sx + uoy * sy + uoz * sz, phiv_nei = unx * sx + uny * sy + unz * sz;
    T cSf_own = L.c * ms, cSf_nei = R.c * ms;
    T ap = fmax(fmax(phiv_own + cSf_own, phiv_nei + cSf_nei), 0.0);
    T am = fmin(fmin(phiv_own - cSf_own, phiv_nei - cSf_nei), 0.0);
    T a_own = ap / (ap - am);
    T aSf = am * a_own;
    T a_nei = 1.0 - a_own;
    phiv_own *= a_own;
    phiv_nei *= a_nei;
    T aphiv_own = phiv_own - aSf, aphiv_nei = phiv_nei + aSf;
    Flux5T<T> o;
    o.phi = aphiv_own * L.rho + aphiv_nei * R.rho;
    T pf = a_own * p_own + a_nei * p_nei;
    o.fx = (aphiv_own * L.rux + aphiv_nei * R.rux) + pf * sx;
    o.fy = (aphiv_own * L.ruy + aphiv_nei * R.ruy) + pf * sy;
    o.fz = (aphiv_own * L.ruz + aphiv_nei * R.ruz) + pf * sz;
    o.fE = aphiv_own * (rhoE_own + p_own) + aphiv_nei * (rhoE_nei + p_nei) + aSf * p_own - aSf * p_nei;
    o.upx = a_own * uox + a_nei * unx;
    o.upy = a_own * uoy + a_nei * uny;
    o.upz = a_own * uoz + a_nei * unz;
    return o;
}

template <int S, class T>
__device__ inline Flux5T<T> eval_flux(const SideT<T>& L, const SideT<T>& R, double sx, double sy, double sz, double ms,
                                      double gamma, double f, int lowMach, double Mainf) {
    if constexpr (S == RDST_FLUX_AUSMPLUS) return flux_ausm<false>(L, R, sx, sy, sz, ms, 0, 0.0);
    else if constexpr (S == RDST_FLUX_AUSMPLUSUP) return flux_ausm<true>(L, R, sx, sy, sz, ms, lowMach, Mainf);
    else if constexpr (S == RDST_FLUX_HLL) return flux_hll(L, R, sx, sy, sz, ms, gamma);
    else if constexpr (S == RDST_FLUX_HLLC) return flux_hllc<false>(L, R, sx, sy, sz, ms, gamma, 1.0);
    else if constexpr (S == RDST_FLUX_HLLCP) return flux_hllc<true>(L, R, sx, sy, sz, ms, gamma, f);
    else return flux_kt(L, R, sx, sy, sz, ms);
}

}  // namespace rdst
