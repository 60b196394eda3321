// flux.hip -- the face loop: reconstruction, the six run-time-selected
// inviscid fluxes and the residual R_ = div(phi, phiUp, phiEp)/V.
//
// Replaces baseFluxDST::computeFluxDST + calcFields
// (fluxSchemes/*/ *DST.C computeFluxDST, baseFluxDST.C:223-255).
//
// Design (MI355X): one thread per cell walks the cell's faces through a
// coalesced ELL map ([slot][cell] layout, faces in surface-field order).
// Each internal face's flux is evaluated by both of its cells from the same
// operands in the same order, so the two contributions cancel exactly
// (discrete conservation) and every cell sums its faces in the reference's
// face order with no atomics and no face-flux round trip through HBM.
// First-order (upwind) reconstruction is the closed form own_=+1 -> owner
// value, nei_=-1 -> neighbour value (baseFluxDST.C:144-171); boundary faces
// take the patch value on both sides.
#include "rdst_internal.hpp"
#include "reduce.hpp"

// No floating-point contraction: the reference evaluates every product
// rounded, so a*x - a*x cancels exactly (a tangential state on both sides of
// a boundary face gives phi = 0).  Fused into fma(a, x, -(a*x)) it leaves the
// rounding error of a*x instead, and the sign tests that read phi (far-field
// regimes mach <= -1, < 0, >= 0, >= 1; inletOutlet; upwind weights) flip on it.
#pragma clang fp contract(off)

namespace rdst {

namespace {

constexpr double kSMALL = 1.0e-15;
constexpr double kVSMALL = 1.0e-300;

struct Side {
    double rho, rux, ruy, ruz, rPsi, e, aStar, c;
    double rr, ux, uy, uz;  // 1/rho and the velocity, formed once per side
};
struct Flux5 {
    double phi, fx, fy, fz, fE, upx, upy, upz;
};

__device__ inline double fsign(double s) { return s >= 0.0 ? 1.0 : -1.0; }
__device__ inline double dpos(double s) { return s > 0.0 ? 1.0 : 0.0; }
__device__ inline double dpos0(double s) { return s >= 0.0 ? 1.0 : 0.0; }
__device__ inline double dneg(double s) { return s < 0.0 ? 1.0 : 0.0; }
__device__ inline double dneg0(double s) { return s <= 0.0 ? 1.0 : 0.0; }
__device__ inline double sq(double a) { return a * a; }

// Side state from the thermo fields: rPsi = 1/psi, e = h - p/rho
// (AUSMPLDST.C:141-148), critical sound speed a* from
// H = max(rhoE/rho, SMALL) + max(p/rho, SMALL) (AUSMPLDST.C:156-161),
// c = sqrt(Cp/Cv rPsi) (KurganovDST.C:129).
__device__ inline Side make_side(double rho, double rux, double ruy, double ruz, double rhoE, double p,
                                 double h, double psi, double gamma) {
    Side s;
    s.rho = rho; s.rux = rux; s.ruy = ruy; s.ruz = ruz;
    s.rr = 1.0 / rho;
    s.ux = rux * s.rr; s.uy = ruy * s.rr; s.uz = ruz * s.rr;
    s.rPsi = 1.0 / psi;
    const double pr = p * s.rr;
    s.e = h - pr;
    double H = fmax(rhoE * s.rr, kSMALL) + fmax(pr, kSMALL);
    s.aStar = sqrt(2.0 * (gamma - 1.0) / (gamma + 1.0) * H);
    s.c = sqrt(gamma * s.rPsi);
    return s;
}

// AUSM+ split Mach/pressure polynomials with the foam sign() masks
// (sign(0) = +1, so at |M| = 1 both branches contribute).
__device__ inline void ausm_split(double Ma_own, double Ma_nei, double alpha, double beta,
                                  double& Ma4_own, double& P5_own, double& Ma4_nei, double& P5_nei) {
    double s1o = fsign(1.0 - fabs(Ma_own)), s2o = fsign(-1.0 + fabs(Ma_own));
    double s1n = fsign(1.0 - fabs(Ma_nei)), s2n = fsign(-1.0 + fabs(Ma_nei));
    double m1o = 0.5 * (s1o + fabs(s1o)), m2o = 0.5 * (s2o + fabs(s2o));
    double m1n = 0.5 * (s1n + fabs(s1n)), m2n = 0.5 * (s2n + fabs(s2n));
    Ma4_own = m1o * (0.25 * sq(Ma_own + 1.0)) * (1.0 - 16.0 * beta * (-0.25) * sq(Ma_own - 1.0)) + m2o * fmax(Ma_own, 0.0);
    P5_own = m1o * 0.25 * sq(Ma_own + 1.0) * (2.0 - Ma_own - 16.0 * alpha * Ma_own * (-0.25) * sq(Ma_own - 1.0)) + m2o * dpos(Ma_own);
    Ma4_nei = m1n * (-0.25) * sq(Ma_nei - 1.0) * (1.0 + 16.0 * beta * 0.25 * sq(Ma_nei + 1.0)) + m2n * fmin(Ma_nei, 0.0);
    P5_nei = m1n * (-0.25) * sq(Ma_nei - 1.0) * (-2.0 - Ma_nei + 16.0 * alpha * Ma_nei * 0.25 * sq(Ma_nei + 1.0)) + m2n * dneg(Ma_nei);
}

// AUSM+ (AUSMPLDST.C:127-217) and AUSM+up (AUSMPLUPDST.C:209-320)
template <bool UP>
__device__ inline Flux5 flux_ausm(const Side& L, const Side& R, double sx, double sy, double sz, double ms,
                                  int lowMach, double Mainf) {
    const double beta = 0.125;
    const double rms = 1.0 / ms;
    double nx = sx * rms, ny = sy * rms, nz = sz * rms;
    double uox = L.ux, uoy = L.uy, uoz = L.uz;
    double unx = R.ux, uny = R.uy, unz = R.uz;
    double Uv_own = uox * nx + uoy * ny + uoz * nz;
    double Uv_nei = unx * nx + uny * ny + unz * nz;
    double rhoE_own = L.rho * (L.e + 0.5 * (uox * uox + uoy * uoy + uoz * uoz));
    double rhoE_nei = R.rho * (R.e + 0.5 * (unx * unx + uny * uny + unz * unz));
    double p_own = L.rho * L.rPsi, p_nei = R.rho * R.rPsi;
    double a12, alpha = 3.0 / 16.0, fa = 0.0, Mmean2 = 0.0;
    if (!UP) {
        a12 = 0.5 * (L.aStar + R.aStar);
    } else {
        double a_own = sq(L.aStar) / fmax(L.aStar, Uv_own);
        double a_nei = sq(R.aStar) / fmax(R.aStar, -Uv_nei);
        a12 = fmin(a_own, a_nei);
        alpha = 0.0;  // only set by the lowMach branch (AUSMPLUPDST.C:118-131, :257-263)
        if (lowMach) {
            Mmean2 = 0.5 * (sq(Uv_own) + sq(Uv_nei)) / sq(a12);
            double M0 = sqrt(fmin(1.0, fmax(Mmean2, sq(Mainf))));
            fa = M0 * (2.0 - M0);
            alpha = 3.0 / 16.0 * (-4.0 + 5.0 * sq(fa));
        }
    }
    double Ma_own = Uv_own / a12, Ma_nei = Uv_nei / a12;
    double Ma4_own, P5_own, Ma4_nei, P5_nei;
    ausm_split(Ma_own, Ma_nei, alpha, beta, Ma4_own, P5_own, Ma4_nei, P5_nei);
    double Ma12 = Ma4_own + Ma4_nei;
    double P12 = P5_own * p_own + P5_nei * p_nei;
    if (UP && lowMach) {
        double Mp = -0.25 / fa * fmax(1.0 - 1.0 * Mmean2, 0.0) * (p_nei - p_own) / (0.5 * (L.rho + R.rho)) / sq(a12);
        double pu = -0.75 * P5_own * P5_nei * (L.rho + R.rho) * a12 * fa * (Uv_own - Uv_nei);
        Ma12 += Mp;
        P12 += pu;
    }
    double ps = dpos(Ma12), ng = dneg0(Ma12);
    Flux5 o;
    o.upx = ps * uox + ng * unx;
    o.upy = ps * uoy + ng * uny;
    o.upz = ps * uoz + ng * unz;
    o.phi = Ma12 * a12 * (ps * L.rho + ng * R.rho) * ms;
    o.fx = o.phi * o.upx + P12 * sx;
    o.fy = o.phi * o.upy + P12 * sy;
    o.fz = o.phi * o.upz + P12 * sz;
    double H_own = (rhoE_own + p_own) * L.rr, H_nei = (rhoE_nei + p_nei) * R.rr;
    o.fE = o.phi * (ps * H_own + ng * H_nei);
    return o;
}

// HLL (HLLDST.C:91-204)
__device__ inline Flux5 flux_hll(const Side& L, const Side& R, double sx, double sy, double sz, double ms, double gamma) {
    double uox = L.ux, uoy = L.uy, uoz = L.uz;
    double unx = R.ux, uny = R.uy, unz = R.uz;
    double E_own = L.e + 0.5 * (uox * uox + uoy * uoy + uoz * uoz);
    double E_nei = R.e + 0.5 * (unx * unx + uny * uny + unz * unz);
    double p_own = L.rho * L.rPsi, p_nei = R.rho * R.rPsi;
    double H_own = E_own + p_own * L.rr, H_nei = E_nei + p_nei * R.rr;
    double coefR = sqrt(fmax(kVSMALL, R.rho) / fmax(kVSMALL, L.rho));
    double HT = (coefR * H_nei + H_own) / (coefR + 1.0);
    double utx = (coefR * unx + uox) / (coefR + 1.0), uty = (coefR * uny + uoy) / (coefR + 1.0), utz = (coefR * unz + uoz) / (coefR + 1.0);
    double aT = sqrt((gamma - 1.0) * (HT - 0.5 * (utx * utx + uty * uty + utz * utz)));
    const double rms = 1.0 / ms;
    double nx = sx * rms, ny = sy * rms, nz = sz * rms;
    double Uv_own = uox * nx + uoy * ny + uoz * nz, Uv_nei = unx * nx + uny * ny + unz * nz;
    double UvT = utx * nx + uty * ny + utz * nz;
    double S_own = fmin(Uv_own - L.aStar, UvT - aT);
    double S_nei = fmax(Uv_nei + R.aStar, UvT + aT);
    double cf_own = dpos0(S_own), cf_m = dneg(S_own) * dpos0(S_nei), cf_nei = dneg(S_nei) * dneg(S_own);
    double dS = S_nei - S_own;
    double ro = L.rho, rn = R.rho;
    double mo_x = ro * uox * Uv_own + p_own * nx, mo_y = ro * uoy * Uv_own + p_own * ny, mo_z = ro * uoz * Uv_own + p_own * nz;
    double mn_x = rn * unx * Uv_nei + p_nei * nx, mn_y = rn * uny * Uv_nei + p_nei * ny, mn_z = rn * unz * Uv_nei + p_nei * nz;
    Flux5 o;
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
template <bool P>
__device__ inline Flux5 flux_hllc(const Side& L, const Side& R, double sx, double sy, double sz, double ms, double gamma,
                                  double f) {
    double uox = L.ux, uoy = L.uy, uoz = L.uz;
    double unx = R.ux, uny = R.uy, unz = R.uz;
    double E_own = L.e + 0.5 * (uox * uox + uoy * uoy + uoz * uoz);
    double E_nei = R.e + 0.5 * (unx * unx + uny * uny + unz * unz);
    double p_own = L.rho * L.rPsi, p_nei = R.rho * R.rPsi;
    double H_own = E_own + p_own * L.rr, H_nei = E_nei + p_nei * R.rr;
    const double rms = 1.0 / ms;
    double nx = sx * rms, ny = sy * rms, nz = sz * rms;
    double Uv_own = uox * nx + uoy * ny + uoz * nz, Uv_nei = unx * nx + uny * ny + unz * nz;
    double coefR = sqrt(fmax(kVSMALL, R.rho) / fmax(kVSMALL, L.rho));
    double HT = (coefR * H_nei + H_own) / (coefR + 1.0);
    double utx = (coefR * unx + uox) / (coefR + 1.0), uty = (coefR * uny + uoy) / (coefR + 1.0), utz = (coefR * unz + uoz) / (coefR + 1.0);
    double UT2 = utx * utx + uty * uty + utz * utz;
    double aT = sqrt((gamma - 1.0) * (HT - 0.5 * UT2));
    double UvT = utx * nx + uty * ny + utz * nz;
    double S_own = fmin(Uv_own - L.aStar, UvT - aT);
    double S_nei = fmax(Uv_nei + R.aStar, UvT + aT);
    double ro = L.rho, rn = R.rho;
    double S_star = (p_nei - p_own + ro * Uv_own * (S_own - Uv_own) - rn * Uv_nei * (S_nei - Uv_nei))
                  / (ro * (S_own - Uv_own) - rn * (S_nei - Uv_nei));
    double cf_own = dpos(S_own);
    double cf_m1 = dneg0(S_own) * dpos(S_star);
    double cf_m2 = dneg0(S_own) * dneg0(S_star) * dpos0(S_nei);
    double cf_nei = dneg0(S_own) * dneg0(S_star) * dneg(S_nei);
    double rhoE_own = ro * E_own, rhoE_nei = rn * E_nei;
    double mo_x = L.rux * Uv_own + p_own * nx, mo_y = L.ruy * Uv_own + p_own * ny, mo_z = L.ruz * Uv_own + p_own * nz;
    double mn_x = R.rux * Uv_nei + p_nei * nx, mn_y = R.ruy * Uv_nei + p_nei * ny, mn_z = R.ruz * Uv_nei + p_nei * nz;
    double eo = (rhoE_own + p_own) * Uv_own, en = (rhoE_nei + p_nei) * Uv_nei;
    double d1 = S_own - S_star, d2 = S_nei - S_star;
    Flux5 o;
    if (!P) {
        double ps_o = p_own + ro * (S_own - Uv_own) * (S_star - Uv_own);
        double ps_n = p_nei + rn * (S_nei - Uv_nei) * (S_star - Uv_nei);
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
        double Ma_own = Uv_own / L.aStar, Ma_nei = Uv_nei / R.aStar;
        double MaT = (coefR * Ma_nei + Ma_own) / (coefR + 1.0);
        double c_own = ro * (S_own - Uv_own), c_nei = rn * (S_nei - Uv_nei);
        double p_star = (c_nei * p_own - c_own * p_nei - c_own * c_nei * (Uv_own - Uv_nei)) / (c_nei - c_own);
        double theta = fmin(fmax(fabs(Ma_own), fabs(Ma_nei)), 1.0);
        double p_av = 0.5 * (p_own + p_nei);
        double p_ss = p_star * theta + (1.0 - theta) * p_av;
        double p_sss = f * p_ss + (1.0 - f) * p_star;
        double phip = (f - 1.0) * S_own * S_nei / (S_nei - S_own) / (1.0 + fabs(MaT)) * (p_nei - p_own) / sq(aT);
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
__device__ inline Flux5 flux_kt(const Side& L, const Side& R, double sx, double sy, double sz, double ms) {
    double uox = L.ux, uoy = L.uy, uoz = L.uz;
    double unx = R.ux, uny = R.uy, unz = R.uz;
    double rhoE_own = L.rho * (L.e + 0.5 * (uox * uox + uoy * uoy + uoz * uoz));
    double rhoE_nei = R.rho * (R.e + 0.5 * (unx * unx + uny * uny + unz * unz));
    double p_own = L.rho * L.rPsi, p_nei = R.rho * R.rPsi;
    double phiv_own = uox * sx + uoy * sy + uoz * sz, phiv_nei = unx * sx + uny * sy + unz * sz;
    double cSf_own = L.c * ms, cSf_nei = R.c * ms;
    double ap = fmax(fmax(phiv_own + cSf_own, phiv_nei + cSf_nei), 0.0);
    double am = fmin(fmin(phiv_own - cSf_own, phiv_nei - cSf_nei), 0.0);
    double a_own = ap / (ap - am);
    double aSf = am * a_own;
    double a_nei = 1.0 - a_own;
    phiv_own *= a_own;
    phiv_nei *= a_nei;
    double aphiv_own = phiv_own - aSf, aphiv_nei = phiv_nei + aSf;
    Flux5 o;
    o.phi = aphiv_own * L.rho + aphiv_nei * R.rho;
    double pf = a_own * p_own + a_nei * p_nei;
    o.fx = (aphiv_own * L.rux + aphiv_nei * R.rux) + pf * sx;
    o.fy = (aphiv_own * L.ruy + aphiv_nei * R.ruy) + pf * sy;
    o.fz = (aphiv_own * L.ruz + aphiv_nei * R.ruz) + pf * sz;
    o.fE = aphiv_own * (rhoE_own + p_own) + aphiv_nei * (rhoE_nei + p_nei) + aSf * p_own - aSf * p_nei;
    o.upx = a_own * uox + a_nei * unx;
    o.upy = a_own * uoy + a_nei * uny;
    o.upz = a_own * uoz + a_nei * unz;
    return o;
}

template <int S>
__device__ inline Flux5 eval_flux(const Side& L, const Side& R, double sx, double sy, double sz, double ms,
                                  double gamma, double f, int lowMach, double Mainf) {
    if constexpr (S == RDST_FLUX_AUSMPLUS) return flux_ausm<false>(L, R, sx, sy, sz, ms, 0, 0.0);
    else if constexpr (S == RDST_FLUX_AUSMPLUSUP) return flux_ausm<true>(L, R, sx, sy, sz, ms, lowMach, Mainf);
    else if constexpr (S == RDST_FLUX_HLL) return flux_hll(L, R, sx, sy, sz, ms, gamma);
    else if constexpr (S == RDST_FLUX_HLLC) return flux_hllc<false>(L, R, sx, sy, sz, ms, gamma, 1.0);
    else if constexpr (S == RDST_FLUX_HLLCP) return flux_hllc<true>(L, R, sx, sy, sz, ms, gamma, f);
    else return flux_kt(L, R, sx, sy, sz, ms);
}


// ---------------------------------------------------------------- viscous terms
// snGrad() of the U boundary condition (fixedValue dc (U_b - U_I);
// zeroGradient 0; basicSymmetry (transform(I - 2nn, U_I) - U_I) dc/2;
// mixed vf (ref - U_I) dc + (1 - vf) 0)
__device__ inline void sngrad_u_bnd(const BcPatchDev& B, const Fields& F, int b, double nx, double ny, double nz,
                                    double dcb, double uIx, double uIy, double uIz, double& ox, double& oy, double& oz) {
    switch (B.U_type) {
        case RDST_BC_FIXED_VALUE:
            ox = dcb * (F.bUx[b] - uIx); oy = dcb * (F.bUy[b] - uIy); oz = dcb * (F.bUz[b] - uIz);
            return;
        case RDST_BC_SYMMETRY: {
            double nU = nx * uIx + ny * uIy + nz * uIz;
            double tx = uIx - 2.0 * nx * nU, ty = uIy - 2.0 * ny * nU, tz = uIz - 2.0 * nz * nU;
            double h = dcb / 2.0;
            ox = (tx - uIx) * h; oy = (ty - uIy) * h; oz = (tz - uIz) * h;
            return;
        }
        case RDST_BC_INLET_OUTLET: case RDST_BC_FARFIELD: {
            double vf = F.vf_U[b];
            ox = vf * (F.refUx[b] - uIx) * dcb + (1.0 - vf) * 0.0;
            oy = vf * (F.refUy[b] - uIy) * dcb + (1.0 - vf) * 0.0;
            oz = vf * (F.refUz[b] - uIz) * dcb + (1.0 - vf) * 0.0;
            return;
        }
        default: ox = oy = oz = 0.0;
    }
}

// tauMC = muEff dev2(T(grad U)), dev2(A) = A - (2/3) tr(A) I
__device__ inline void tau_mc(double mu, const double* g, double* t) {
    const double tr = g[0] + g[4] + g[8];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) t[3 * i + j] = mu * (g[3 * j + i] - (i == j ? (2.0 / 3.0) * tr : 0.0));
}

// K2: fvc::grad(U), Gauss linear, one thread per cell
__global__ __launch_bounds__(256) void k_grad_u(MeshDev m, Fields F) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const size_t E = (size_t)m.ellW * m.nC;
    const double ux = F.Ux[c], uy = F.Uy[c], uz = F.Uz[c];
    double g[9];
    for (int q = 0; q < 9; ++q) g[q] = 0.0;
    for (int k = 0; k < m.ellW; ++k) {
        const size_t e = (size_t)k * m.nC + c;
        const uint32_t code = m.ellCode[e];
        const uint32_t kind = code >> 30;
        if (kind == ELL_NONE) break;
        const int idx = (int)(code & ELL_IDX_MASK);
        const double S[3] = {m.ellSf[e], m.ellSf[E + e], m.ellSf[2 * E + e]};
        double f[3];
        double sgn = 1.0;
        if (kind == ELL_BOUNDARY) {
            f[0] = F.bUx[idx]; f[1] = F.bUy[idx]; f[2] = F.bUz[idx];
        } else {
            const double w = m.ellWt[e];
            const double ox = F.Ux[idx], oy = F.Uy[idx], oz = F.Uz[idx];
            const bool own = kind == ELL_OWNER;
            const double Px = own ? ux : ox, Py = own ? uy : oy, Pz = own ? uz : oz;
            const double Nx = own ? ox : ux, Ny = own ? oy : uy, Nz = own ? oz : uz;
            f[0] = w * (Px - Nx) + Nx; f[1] = w * (Py - Ny) + Ny; f[2] = w * (Pz - Nz) + Nz;
            if (!own) sgn = -1.0;
        }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const double t = S[i] * f[j];
                if (sgn > 0) g[3 * i + j] += t; else g[3 * i + j] -= t;
            }
    }
    const double V = m.V[c];
    for (int q = 0; q < 9; ++q) F.gradU[(size_t)q * m.nC + c] = g[q] / V;
}

// gaussGrad::correctBoundaryConditions: g_b = g_I + n (snGrad_b - n.g_I)
__global__ void k_grad_u_bnd(MeshDev m, Fields F, const BcPatchDev* bc) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.nB) return;
    const int ci = m.bcell[b], k = m.nI + b;
    const double ms = m.magSf[k];
    const double n[3] = {m.Sf[k] / ms, m.Sf[m.nF + k] / ms, m.Sf[2 * m.nF + k] / ms};
    double gi[9];
    for (int q = 0; q < 9; ++q) gi[q] = F.gradU[(size_t)q * m.nC + ci];
    double sn[3];
    sngrad_u_bnd(bc[m.bpatch[b]], F, b, n[0], n[1], n[2], m.dc[k], F.Ux[ci], F.Uy[ci], F.Uz[ci], sn[0], sn[1], sn[2]);
    for (int j = 0; j < 3; ++j) {
        const double nG = n[0] * gi[j] + n[1] * gi[3 + j] + n[2] * gi[6 + j];
        const double d = sn[j] - nG;
        for (int i = 0; i < 3; ++i) F.bgradU[(size_t)(3 * i + j) * m.nB + b] = gi[3 * i + j] + n[i] * d;
    }
}

__device__ inline Side load_cell(const Fields& F, int i, double gamma) {
    return make_side(F.rho[i], F.rhoUx[i], F.rhoUy[i], F.rhoUz[i], F.rhoE[i], F.p[i], F.h[i], F.psi[i], gamma);
}
__device__ inline Side load_bnd(const Fields& F, int b, double gamma) {
    return make_side(F.brho[b], F.brhoUx[b], F.brhoUy[b], F.brhoUz[b], F.brhoE[b], F.bp[b], F.bh[b], F.bpsi[b], gamma);
}

// K1: face flux + residual, one thread per cell.
template <int S, int STORE, bool VISC>
__global__ __launch_bounds__(256) void k_flux_residual(MeshDev m, Fields F, const BcPatchDev* bc, double gamma, int lowMach,
                                                       double Mainf) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const Side me = load_cell(F, c, gamma);
    const double fHc = (S == RDST_FLUX_HLLCP) ? F.fH[c] : 0.0;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0;
    // viscous accumulators: laplacian(muEff, U), div(tauMC), div(sigmaDotU)
    double lx = 0.0, ly = 0.0, lz = 0.0, tx = 0.0, ty = 0.0, tz = 0.0, sd = 0.0;
    double gC[9], tauC[9], muC = 0.0, uCx = 0.0, uCy = 0.0, uCz = 0.0;
    if (VISC) {
        muC = F.mut[c] + F.mu[c];
        uCx = F.Ux[c]; uCy = F.Uy[c]; uCz = F.Uz[c];
        for (int q = 0; q < 9; ++q) gC[q] = F.gradU[(size_t)q * m.nC + c];
        tau_mc(muC, gC, tauC);
    }
    const size_t E = (size_t)m.ellW * m.nC;
    for (int k = 0; k < m.ellW; ++k) {
        const size_t e = (size_t)k * m.nC + c;
        const uint32_t code = m.ellCode[e];
        const uint32_t kind = code >> 30;
        if (kind == ELL_NONE) break;
        const int idx = (int)(code & ELL_IDX_MASK);
        const double sx = m.ellSf[e], sy = m.ellSf[E + e], sz = m.ellSf[2 * E + e], ms = m.ellMagSf[e];
        Side L, R;
        double f = 1.0;
        if (kind == ELL_BOUNDARY) {
            L = load_bnd(F, idx, gamma);
            R = L;
        } else {
            const Side other = load_cell(F, idx, gamma);
            if (kind == ELL_OWNER) { L = me; R = other; } else { L = other; R = me; }
        }
        if constexpr (S == RDST_FLUX_HLLCP) {
            if (kind != ELL_BOUNDARY) {
                const double w = m.ellWt[e];
                const double fo = (kind == ELL_OWNER) ? fHc : F.fH[idx];
                const double fn = (kind == ELL_OWNER) ? F.fH[idx] : fHc;
                f = w * (fo - fn) + fn;
            }
        }
        const Flux5 o = eval_flux<S>(L, R, sx, sy, sz, ms, gamma, f, lowMach, Mainf);
        if (VISC) {
            // face viscous flux (interp(muEff)|S| snGrad(U) and Sf & interp(tauMC)), then its work on Up
            double lfx, lfy, lfz, tfx, tfy, tfz;
            if (kind == ELL_BOUNDARY) {
                const double rms = 1.0 / ms;
                const double nx = sx * rms, ny = sy * rms, nz = sz * rms;
                double snx, sny, snz;
                sngrad_u_bnd(bc[m.bpatch[idx]], F, idx, nx, ny, nz, m.ellDc[e], uCx, uCy, uCz, snx, sny, snz);
                const double bmuE = F.bmut[idx] + F.bmu[idx];
                const double mg = bmuE * ms;
                lfx = mg * snx; lfy = mg * sny; lfz = mg * snz;
                double gb[9], tb[9];
                for (int q = 0; q < 9; ++q) gb[q] = F.bgradU[(size_t)q * m.nB + idx];
                tau_mc(bmuE, gb, tb);
                tfx = sx * tb[0] + sy * tb[3] + sz * tb[6];
                tfy = sx * tb[1] + sy * tb[4] + sz * tb[7];
                tfz = sx * tb[2] + sy * tb[5] + sz * tb[8];
            } else {
                const bool own = kind == ELL_OWNER;
                const double w = m.ellWt[e];
                const double muO = F.mut[idx] + F.mu[idx];
                double gO[9], tauO[9];
                for (int q = 0; q < 9; ++q) gO[q] = F.gradU[(size_t)q * m.nC + idx];
                tau_mc(muO, gO, tauO);
                const double* gP = own ? gC : gO;
                const double* gN = own ? gO : gC;
                const double* tP = own ? tauC : tauO;
                const double* tN = own ? tauO : tauC;
                const double muP = own ? muC : muO, muN = own ? muO : muC;
                const double oUx = F.Ux[idx], oUy = F.Uy[idx], oUz = F.Uz[idx];
                const double UPx = own ? uCx : oUx, UPy = own ? uCy : oUy, UPz = own ? uCz : oUz;
                const double UNx = own ? oUx : uCx, UNy = own ? oUy : uCy, UNz = own ? oUz : uCz;
                const double muf = w * (muP - muN) + muN;
                double gf[9], tf[9];
                for (int q = 0; q < 9; ++q) {
                    gf[q] = w * (gP[q] - gN[q]) + gN[q];
                    tf[q] = w * (tP[q] - tN[q]) + tN[q];
                }
                const double cx = m.ellCorr[e], cy = m.ellCorr[E + e], cz = m.ellCorr[2 * E + e];
                const double dcf = m.ellDc[e];
                const double snx = dcf * (UNx - UPx) + (cx * gf[0] + cy * gf[3] + cz * gf[6]);
                const double sny = dcf * (UNy - UPy) + (cx * gf[1] + cy * gf[4] + cz * gf[7]);
                const double snz = dcf * (UNz - UPz) + (cx * gf[2] + cy * gf[5] + cz * gf[8]);
                const double mg = muf * ms;
                lfx = mg * snx; lfy = mg * sny; lfz = mg * snz;
                tfx = sx * tf[0] + sy * tf[3] + sz * tf[6];
                tfy = sx * tf[1] + sy * tf[4] + sz * tf[7];
                tfz = sx * tf[2] + sy * tf[5] + sz * tf[8];
            }
            const double sdf = (lfx + tfx) * o.upx + (lfy + tfy) * o.upy + (lfz + tfz) * o.upz;
            if (kind == ELL_NEIGH) {
                lx -= lfx; ly -= lfy; lz -= lfz; tx -= tfx; ty -= tfy; tz -= tfz; sd -= sdf;
            } else {
                lx += lfx; ly += lfy; lz += lfz; tx += tfx; ty += tfy; tz += tfz; sd += sdf;
            }
        }
        if (kind == ELL_NEIGH) {
            a0 -= o.phi; a1 -= o.fx; a2 -= o.fy; a3 -= o.fz; a4 -= o.fE;
        } else {
            a0 += o.phi; a1 += o.fx; a2 += o.fy; a3 += o.fz; a4 += o.fE;
            if (STORE == STORE_ALL) {
                const int fi = m.ellFace[e];
                F.phi[fi] = o.phi; F.phiUx[fi] = o.fx; F.phiUy[fi] = o.fy; F.phiUz[fi] = o.fz; F.phiE[fi] = o.fE;
                F.Upx[fi] = o.upx; F.Upy[fi] = o.upy; F.Upz[fi] = o.upz;
            } else if (STORE == STORE_PHI) {
                F.phi[m.ellFace[e]] = o.phi;  // the turbulence convection reads phi on every face
            } else if (kind == ELL_BOUNDARY) {
                F.phi[m.nI + idx] = o.phi;  // boundary mass flux feeds the BC updates
            }
        }
    }
    const double V = m.V[c];
    const int n = m.nC;
    double r1 = a1 / V, r2 = a2 / V, r3 = a3 / V, r4 = a4 / V;
    if (VISC) {
        // R1 -= laplacian + div(tauMC); R4 -= div(sigmaDotU)
        r1 -= lx / V + tx / V;
        r2 -= ly / V + ty / V;
        r3 -= lz / V + tz / V;
        r4 -= sd / V;
    }
    F.R[c] = a0 / V;
    F.R[n + c] = r1;
    F.R[2 * n + c] = r2;
    F.R[3 * n + c] = r3;
    F.R[4 * n + c] = r4;
}

// HLLCP pressure sensor (HLLCPDST.C:131-164): fCells = min over cellCells of
// min(p_i/p_j, p_j/p_i), starting from 1; store pow3.
__global__ void k_hllcp_sensor(MeshDev m, Fields F) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const double pc = F.p[c];
    double fc = 1.0;
    for (int k = 0; k < m.ellW; ++k) {
        const uint32_t code = m.ellCode[(size_t)k * m.nC + c];
        const uint32_t kind = code >> 30;
        if (kind == ELL_NONE) break;
        if (kind == ELL_BOUNDARY) continue;
        const double pj = F.p[code & ELL_IDX_MASK];
        fc = fmin(fc, fmin(pc / pj, pj / pc));
    }
    F.fH[c] = fc * fc * fc;
}

// forceCoeffs on one patch: sum (p_b - pRef) Sf and sum Sf & devRhoReff_b,
// devRhoReff = -muEff dev(twoSymm(grad U)) from the corrected boundary
// gradient; one block, fixed strided order (deterministic)
__global__ __launch_bounds__(RED_BS) void k_forces(MeshDev m, Fields F, int b0, int b1, double pRef, int visc,
                                                   double* out) {
    __shared__ double sh[16];
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (int b = b0 + (int)threadIdx.x; b < b1; b += blockDim.x) {
        const int k = m.nI + b;
        const double S[3] = {m.Sf[k], m.Sf[m.nF + k], m.Sf[2 * m.nF + k]};
        const double dp = F.bp[b] - pRef;
        for (int d = 0; d < 3; ++d) a[d] += S[d] * dp;
        if (visc) {
            double G[9];
            for (int q = 0; q < 9; ++q) G[q] = F.bgradU[(size_t)q * m.nB + b];
            const double tr = G[0] + G[4] + G[8];
            const double muE = F.bmut[b] + F.bmu[b];
            for (int j = 0; j < 3; ++j) {
                double v = 0.0;
                for (int i = 0; i < 3; ++i) {
                    const double ts = G[3 * i + j] + G[3 * j + i];
                    v += S[i] * (-muE * (ts - (i == j ? (2.0 / 3.0) * tr : 0.0)));
                }
                a[3 + j] += v;
            }
        }
    }
    for (int q = 0; q < 6; ++q) {
        const double t = block_sum(a[q], sh);
        if (threadIdx.x == 0) out[q] = t;
    }
}

template <int S, int STORE>
void launch_flux_mode(Case& c) {
    MeshDev md = c.meshDev();
    const int bs = 256;
    const int nb = blocks_for(md.nC, bs);
    const BcPatchDev* bc = c.d_bc.p;
    if (c.st.viscous)
        hipLaunchKernelGGL((k_flux_residual<S, STORE, true>), dim3(nb), dim3(bs), 0, c.stream, md, c.f, bc, c.th.gamma,
                           c.st.low_mach, c.st.Mainf);
    else
        hipLaunchKernelGGL((k_flux_residual<S, STORE, false>), dim3(nb), dim3(bs), 0, c.stream, md, c.f, bc, c.th.gamma,
                           c.st.low_mach, c.st.Mainf);
}

template <int S>
void launch_flux(Case& c, int store) {
    if (c.st.viscous) grad_u_launch(c);
    switch (store) {
        case STORE_ALL: launch_flux_mode<S, STORE_ALL>(c); break;
        case STORE_PHI: launch_flux_mode<S, STORE_PHI>(c); break;
        default: launch_flux_mode<S, STORE_BPHI>(c);
    }
}

}  // namespace

void grad_u_launch(Case& c) {
    MeshDev md = c.meshDev();
    const int bs = 256;
    hipLaunchKernelGGL(k_grad_u, dim3(blocks_for(md.nC, bs)), dim3(bs), 0, c.stream, md, c.f);
    if (md.nB) hipLaunchKernelGGL(k_grad_u_bnd, dim3(blocks_for(md.nB, bs)), dim3(bs), 0, c.stream, md, c.f, c.d_bc.p);
}

void flux_launch(Case& c, int store) {
    MeshDev md = c.meshDev();
    if (c.st.flux_scheme == RDST_FLUX_HLLCP)
        hipLaunchKernelGGL(k_hllcp_sensor, dim3(blocks_for(md.nC, 256)), dim3(256), 0, c.stream, md, c.f);
    switch (c.st.flux_scheme) {
        case RDST_FLUX_AUSMPLUS: launch_flux<RDST_FLUX_AUSMPLUS>(c, store); break;
        case RDST_FLUX_AUSMPLUSUP: launch_flux<RDST_FLUX_AUSMPLUSUP>(c, store); break;
        case RDST_FLUX_HLL: launch_flux<RDST_FLUX_HLL>(c, store); break;
        case RDST_FLUX_HLLC: launch_flux<RDST_FLUX_HLLC>(c, store); break;
        case RDST_FLUX_HLLCP: launch_flux<RDST_FLUX_HLLCP>(c, store); break;
        case RDST_FLUX_KURGANOV: launch_flux<RDST_FLUX_KURGANOV>(c, store); break;
        default: throw ArgError("unknown flux scheme", RDST_ERR_UNKNOWN_NAME);
    }
    RDST_HIP(hipGetLastError());
}

void flux_residual(Case& c, int storeFaceFluxes) {
    if (!c.hasState) throw ArgError("rdst_flux_residual: no state (call rdst_case_set_state)", RDST_ERR_STATE);
    RDST_HIP(hipEventRecord(c.ev0, c.stream));
    flux_launch(c, storeFaceFluxes);
    RDST_HIP(hipEventRecord(c.ev1, c.stream));
    RDST_HIP(hipStreamSynchronize(c.stream));
    RDST_HIP(hipEventElapsedTime(&c.lastFluxMs, c.ev0, c.ev1));
}

double flux_time(Case& c, int n) {
    if (!c.hasState) throw ArgError("rdst_time_flux: no state", RDST_ERR_STATE);
    flux_launch(c, STORE_BPHI);  // warm-up
    RDST_HIP(hipEventRecord(c.ev0, c.stream));
    for (int i = 0; i < n; ++i) flux_launch(c, STORE_BPHI);
    RDST_HIP(hipEventRecord(c.ev1, c.stream));
    RDST_HIP(hipStreamSynchronize(c.stream));
    float ms = 0.f;
    RDST_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    return (double)ms / n;
}

void forces(Case& c, int patch, const double* dragDir, const double* liftDir, double pRef, double rhoInf, double magUInf,
            double aRef, double* out) {
    const Mesh& m = *c.mesh;
    // boundary faces of the patch are contiguous in the compacted boundary order
    int b0 = -1, b1 = -1;
    for (int b = 0; b < m.nB; ++b)
        if (m.bPatch[b] == patch) {
            if (b0 < 0) b0 = b;
            b1 = b + 1;
        }
    double F[6] = {0, 0, 0, 0, 0, 0};
    if (b0 >= 0) {
        if (c.st.viscous) grad_u_launch(c);
        DBuf<double> d;
        d.alloc(6);
        hipLaunchKernelGGL(k_forces, dim3(1), dim3(RED_BS), 0, c.stream, c.meshDev(), c.f, b0, b1, pRef,
                           c.st.viscous, d.p);
        RDST_HIP(hipGetLastError());
        RDST_HIP(hipMemcpyAsync(F, d.p, sizeof(F), hipMemcpyDeviceToHost, c.stream));
        RDST_HIP(hipStreamSynchronize(c.stream));
    }
    const double Ft[3] = {F[0] + F[3], F[1] + F[4], F[2] + F[5]};
    const double q = 0.5 * rhoInf * magUInf * magUInf * aRef;
    out[0] = (Ft[0] * dragDir[0] + Ft[1] * dragDir[1] + Ft[2] * dragDir[2]) / q;
    out[1] = (Ft[0] * liftDir[0] + Ft[1] * liftDir[1] + Ft[2] * liftDir[2]) / q;
    out[2] = Ft[0]; out[3] = Ft[1]; out[4] = Ft[2];
    out[5] = F[3]; out[6] = F[4]; out[7] = F[5];
}

}  // namespace rdst
