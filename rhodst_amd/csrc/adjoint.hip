// adjoint.hip -- discrete adjoint of the first-order inviscid residual for
// force-coefficient sensitivities.
//
// The residual is R_c(W) = (1/V_c) sum_f s_f F_f(W_L, W_R), the steady form of
// the face loop (baseFluxDST.C:223-255 without viscous terms), with the cell
// primitives retrieved from W as updateFields.H:40-94 does and the boundary
// states from the boundary conditions evaluated at the same state (the old
// and new values of the far-field updateCoeffs coincide at a fixed point).
// The linearisation is exact: every face flux is differentiated in forward
// mode (dual numbers, fluxfun.hpp) through the flux scheme, the
// reconstruction, the W -> primitive chain and the boundary conditions; only
// the piecewise-constant masks (foam pos/neg/sign, the far-field regimes and
// inletOutlet valueFractions read from the stored mass flux) carry no
// derivative.  This is not the Roe-Jacobian of the implicit operator
// (Roe.C:152-379), which approximates the same matrix.
//
// J = dR/dW is block-CSR on the mesh's slots (one 5x5 block per cell pair of
// a face and the diagonal); its transpose is the same structure with the
// upper/lower slots swapped and every block transposed.  The adjoint system
// J^T lambda = dJ/dW is solved by GMRES + block ILU(0) of J^T (FP64), or by
// FP32 GMRES cycles with FP32 transposed ILU inside an FP64 defect correction
// on the exact FP64 J^T (mixed).  HLLCP is excluded: its pressure sensor
// couples second neighbours, outside the block stencil.
#include <algorithm>
#include <cmath>
#include <vector>

#include "rdst_internal.hpp"
#include "reduce.hpp"
#include "fluxfun.hpp"

namespace rdst {

namespace {

template <class T>
__device__ inline T clampT(const T& x, double lo, double hi) {
    if (val(x) < lo) return T(lo);
    if (val(x) > hi) return T(hi);
    return x;
}

template <class T>
struct Prim {
    T rho, rux, ruy, ruz, rhoE, ux, uy, uz, h, Tt, psi, p;
};

struct Bounds {
    double rhoMin, rhoMax, hMin, hMax, pMin, pMax;
};

// updateFields.H:40-94 on one cell: rho bounded, U = rhoU/rho,
// h = (Cp/Cv)(rhoE/rho - |U|^2/2) bounded, T = h/Cp, psi = 1/(R T),
// p = rho/psi bounded (the bounds are inactive at a regular state)
template <class T>
__device__ inline Prim<T> prim_from_w(T w0, T w1, T w2, T w3, T w4, const Thermo& th, const Bounds& bd) {
    Prim<T> q;
    q.rho = clampT(w0, bd.rhoMin, bd.rhoMax);
    q.rux = w1; q.ruy = w2; q.ruz = w3; q.rhoE = w4;
    q.ux = w1 / q.rho; q.uy = w2 / q.rho; q.uz = w3 / q.rho;
    T h = (th.Cp / th.Cv) * (w4 / q.rho - 0.5 * (q.ux * q.ux + q.uy * q.uy + q.uz * q.uz));
    q.h = clampT(h, bd.hMin, bd.hMax);
    q.Tt = q.h / th.Cp;
    q.psi = 1.0 / (th.R * q.Tt);
    q.p = clampT(q.rho / q.psi, bd.pMin, bd.pMax);
    return q;
}

template <class T>
struct BState {
    T rho, rux, ruy, ruz, rhoE, p, h, psi;
};

__device__ inline double mpos0(double s) { return s >= 0.0 ? 1.0 : 0.0; }
__device__ inline double mneg(double s) { return s < 0.0 ? 1.0 : 0.0; }
__device__ inline double mneg0(double s) { return s <= 0.0 ? 1.0 : 0.0; }

// the boundary state of face b from its cell (update.hip k_update_bnd_pre /
// k_update_bnd_post with old = new); masks from the stored phi, rho_b, psi_b
template <class T>
__device__ inline BState<T> bc_state(const BcPatchDev& B, const Fields& F, int b, int k, double nx, double ny, double nz,
                                     double ms, const Prim<T>& I, const Thermo& th, const Bounds& bd) {
    const double gam = th.Cp / th.Cv;
    const double phib = F.phi[k];
    // far-field regime masks (P.C:174-191), frozen
    double supIn = 0, supOut = 0, subIn = 0, subOut = 0;
    if (B.U_type == RDST_BC_FARFIELD || B.T_type == RDST_BC_FARFIELD || B.p_type == RDST_BC_FARFIELD) {
        const double cb = ::sqrt(gam / F.bpsi[b]);
        const double mach = phib / F.brho[b] / ms / cb;
        supIn = mneg0(mach + 1); supOut = mpos0(mach - 1);
        subIn = mneg(mach) - supIn; subOut = mpos0(mach) - supOut;
    }
    const T cI = sqrt(gam / I.psi);
    const T rho0c0 = I.rho * cI;
    // velocity
    T ux, uy, uz;
    switch (B.U_type) {
        case RDST_BC_FIXED_VALUE: ux = T(B.U_value[0]); uy = T(B.U_value[1]); uz = T(B.U_value[2]); break;
        case RDST_BC_SYMMETRY: {
            const T nU = nx * I.ux + ny * I.uy + nz * I.uz;
            const T tx = I.ux - 2.0 * nx * nU, ty = I.uy - 2.0 * ny * nU, tz = I.uz - 2.0 * nz * nU;
            ux = (tx + I.ux) / 2.0; uy = (ty + I.uy) / 2.0; uz = (tz + I.uz) / 2.0;
            break;
        }
        case RDST_BC_INLET_OUTLET: {
            const double vf = 1.0 - mpos0(phib);
            ux = vf * B.U_value[0] + (1.0 - vf) * I.ux;
            uy = vf * B.U_value[1] + (1.0 - vf) * I.uy;
            uz = vf * B.U_value[2] + (1.0 - vf) * I.uz;
            break;
        }
        case RDST_BC_FARFIELD: {
            const T dUn = nx * (B.U_inf[0] - I.ux) + ny * (B.U_inf[1] - I.uy) + nz * (B.U_inf[2] - I.uz);
            const T pb = 0.5 * (B.p_inf + I.p - rho0c0 * dUn);
            const T a = (B.p_inf - pb) / rho0c0, o = (I.p - pb) / rho0c0;
            const T rx = supIn * B.U_inf[0] + subIn * (B.U_inf[0] - nx * a) + subOut * (nx * o);
            const T ry = supIn * B.U_inf[1] + subIn * (B.U_inf[1] - ny * a) + subOut * (ny * o);
            const T rz = supIn * B.U_inf[2] + subIn * (B.U_inf[2] - nz * a) + subOut * (nz * o);
            const double vf = 1 - supOut - subOut;
            ux = vf * rx + (1.0 - vf) * I.ux;
            uy = vf * ry + (1.0 - vf) * I.uy;
            uz = vf * rz + (1.0 - vf) * I.uz;
            break;
        }
        default: ux = I.ux; uy = I.uy; uz = I.uz;
    }
    // enthalpy from the temperature BC
    T hb;
    switch (B.T_type) {
        case RDST_BC_FIXED_VALUE: hb = T(th.Cp * B.T_value); break;
        case RDST_BC_INLET_OUTLET: {
            const double vf = 1.0 - mpos0(phib);
            hb = vf * (th.Cp * B.T_value) + (1.0 - vf) * I.h;
            break;
        }
        case RDST_BC_FARFIELD: {
            const T dUn = nx * (B.U_inf[0] - I.ux) + ny * (B.U_inf[1] - I.uy) + nz * (B.U_inf[2] - I.uz);
            const T pb = 0.5 * (B.p_inf + I.p - rho0c0 * dUn);
            const T RI = I.p / I.rho / I.Tt;
            const T rhoa = B.p_inf / B.T_inf / RI;
            const T cI2 = cI * cI;
            const T subsIn = pb / (rhoa + (pb - B.p_inf) / cI2) / RI;
            const T subsOutVF = pb / (I.p + RI * I.Tt * (pb - I.p) / cI2);
            const T TrefH = supIn * B.T_inf + subIn * subsIn;
            const T vf = 1 - supOut - subsOutVF * subOut;
            hb = vf * (th.Cp * TrefH) + (1.0 - vf) * I.h;
            break;
        }
        default: hb = I.h;
    }
    T Tb;
    if (B.T_type == RDST_BC_FIXED_VALUE) { Tb = T(B.T_value); hb = T(th.Cp * B.T_value); }
    else Tb = hb / th.Cp;
    const T psib = 1.0 / (th.R * Tb);
    // pressure
    T pb;
    switch (B.p_type) {
        case RDST_BC_FIXED_VALUE: pb = T(B.p_value); break;
        case RDST_BC_INLET_OUTLET: {
            const double vf = 1.0 - mpos0(phib);
            pb = vf * B.p_value + (1.0 - vf) * I.p;
            break;
        }
        case RDST_BC_FARFIELD: {
            const T dUn = nx * (B.U_inf[0] - I.ux) + ny * (B.U_inf[1] - I.uy) + nz * (B.U_inf[2] - I.uz);
            const T pbf = B.p_inf - rho0c0 * dUn;
            const T ref = supIn * B.p_inf + subIn * pbf + subOut * B.p_inf;
            const double vf = 1 - supOut - 0.5 * subIn;
            pb = vf * ref + (1.0 - vf) * I.p;
            break;
        }
        default: pb = I.p;
    }
    pb = clampT(pb, bd.pMin, bd.pMax);
    BState<T> s;
    s.p = pb; s.h = hb; s.psi = psib;
    s.rho = pb * psib;
    s.rux = s.rho * ux; s.ruy = s.rho * uy; s.ruz = s.rho * uz;
    s.rhoE = s.rho * (hb + 0.5 * (ux * ux + uy * uy + uz * uz)) - pb;
    return s;
}

template <class T>
__device__ inline SideT<T> side_of(const Prim<T>& q, double gamma) {
    return make_side(q.rho, q.rux, q.ruy, q.ruz, q.rhoE, q.p, q.h, q.psi, gamma);
}
template <class T>
__device__ inline SideT<T> side_of(const BState<T>& q, double gamma) {
    return make_side(q.rho, q.rux, q.ruy, q.ruz, q.rhoE, q.p, q.h, q.psi, gamma);
}

__device__ inline Prim<D1> seeded(const double* W, int i, int k, const Thermo& th, const Bounds& bd) {
    D1 w[5];
    for (int q = 0; q < 5; ++q) w[q] = D1(W[5 * (size_t)i + q], q == k ? 1.0 : 0.0);
    return prim_from_w(w[0], w[1], w[2], w[3], w[4], th, bd);
}

// one thread per cell: its row of J = dR/dW (diagonal block and the block of
// every internal face, at this row's slots).  NEWTON: the row of the Newton
// operator instead, V_c dR_c/dW + rDeltaTDST V_c I (the face sums not divided
// by V), stored as TJ (FP32 for the mixed path), and the Newton source -R_c V_c
// into src (the ddt part of the steady-state source vanishes).  The Newton
// operator feeds only the SpMV, so it is written in the sliced-ELL layout
// (entry q of slot s at sellOfSlot[s] + 64 q) rather than block-CSR.
// VISC (Newton operator only): the viscous face terms of calcFields
// (baseFluxDST.C:229-250) are added with their compact linearisation: the
// interp(muEff)|S| deltaCoeffs (U_N - U_P) part of the corrected snGrad (and
// its boundary form) and the work term's dependence on it and on Up are
// differentiated; the gradient-borne parts (non-orthogonal correction,
// Sf & interp(tauMC)) and muEff (the turbulence frozen) keep their current
// values.  Their full linearisation would reach second neighbours, outside
// the block stencil.
template <class T>
__device__ inline void sngrad_u_bnd_t(const BcPatchDev& B, const Fields& F, int b, double nx, double ny, double nz,
                                      double dcb, const T& uIx, const T& uIy, const T& uIz, T& ox, T& oy, T& oz) {
    switch (B.U_type) {
        case RDST_BC_FIXED_VALUE:
            ox = dcb * (F.bUx[b] - uIx); oy = dcb * (F.bUy[b] - uIy); oz = dcb * (F.bUz[b] - uIz);
            return;
        case RDST_BC_SYMMETRY: {
            const T nU = nx * uIx + ny * uIy + nz * uIz;
            const T tx = uIx - 2.0 * nx * nU, ty = uIy - 2.0 * ny * nU, tz = uIz - 2.0 * nz * nU;
            const double h = dcb / 2.0;
            ox = (tx - uIx) * h; oy = (ty - uIy) * h; oz = (tz - uIz) * h;
            return;
        }
        case RDST_BC_INLET_OUTLET: case RDST_BC_FARFIELD: {
            const double vf = F.vf_U[b];
            ox = vf * (F.refUx[b] - uIx) * dcb; oy = vf * (F.refUy[b] - uIy) * dcb; oz = vf * (F.refUz[b] - uIz) * dcb;
            return;
        }
        default: ox = T(0.0); oy = T(0.0); oz = T(0.0);
    }
}

__device__ inline void tau_mc_a(double mu, const double* g, double* t) {
    const double tr = g[0] + g[4] + g[8];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) t[3 * i + j] = mu * (g[3 * j + i] - (i == j ? (2.0 / 3.0) * tr : 0.0));
}

template <int S, class TJ, bool NEWTON, bool VISC>
__global__ __launch_bounds__(256) void k_adj_jacobian(MeshDev m, Fields F, const BcPatchDev* bc, Thermo th, Bounds bd,
                                                      int lowMach, double Mainf, TJ* J, double* src,
                                                      const int64_t* sellOfSlot, int heat) {
    // value index of entry q of block-CSR slot s
    auto at = [&](int s, int q) -> size_t {
        if constexpr (NEWTON) return (size_t)sellOfSlot[s] + 64 * (size_t)q;
        else return 25 * (size_t)s + q;
    };
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const size_t E = (size_t)m.ellW * m.nC;
    const double rV = NEWTON ? 1.0 : 1.0 / m.V[c];
    const double gamma = th.gamma;
    double Jcc[25];
    for (int q = 0; q < 25; ++q) Jcc[q] = 0.0;
    for (int ke = 0; ke < m.ellW; ++ke) {
        const size_t e = (size_t)ke * m.nC + c;
        const uint32_t code = m.ellCode[e];
        const uint32_t kind = code >> 30;
        if (kind == ELL_NONE) break;
        const int idx = (int)(code & ELL_IDX_MASK);
        const double sx = m.ellSf[e], sy = m.ellSf[E + e], sz = m.ellSf[2 * E + e], ms = m.ellMagSf[e];
        if (kind == ELL_BOUNDARY) {
            const int k = m.nI + idx;
            const double nx = sx / ms, ny = sy / ms, nz = sz / ms;
            // frozen viscous boundary data: muEff_b |S| and Sf & tauMC_b
            double mgb = 0.0, tbx = 0.0, tby = 0.0, tbz = 0.0;
            if constexpr (VISC) {
                const double bmuE = F.bmut[idx] + F.bmu[idx];
                mgb = bmuE * ms;
                double gb[9], tb[9];
                for (int q = 0; q < 9; ++q) gb[q] = F.bgradU[(size_t)q * m.nB + idx];
                tau_mc_a(bmuE, gb, tb);
                tbx = sx * tb[0] + sy * tb[3] + sz * tb[6];
                tby = sx * tb[1] + sy * tb[4] + sz * tb[7];
                tbz = sx * tb[2] + sy * tb[5] + sz * tb[8];
            }
            for (int kk = 0; kk < 5; ++kk) {
                const Prim<D1> I = seeded(F.W, c, kk, th, bd);
                const BState<D1> bs = bc_state(bc[m.bpatch[idx]], F, idx, k, nx, ny, nz, ms, I, th, bd);
                const SideT<D1> L = side_of(bs, gamma);
                const Flux5T<D1> o = eval_flux<S>(L, L, sx, sy, sz, ms, gamma, 1.0, lowMach, Mainf);
                D1 vx(0.0), vy(0.0), vz(0.0), ve(0.0);
                if constexpr (VISC) {
                    D1 snx, sny, snz;
                    sngrad_u_bnd_t(bc[m.bpatch[idx]], F, idx, nx, ny, nz, m.ellDc[e], I.ux, I.uy, I.uz, snx, sny, snz);
                    vx = mgb * snx; vy = mgb * sny; vz = mgb * snz;
                    ve = (vx + tbx) * o.upx + (vy + tby) * o.upy + (vz + tbz) * o.upz;
                    if (heat)  // conduction extension: alphaEff_b |S| dc (h_b - h_P), alphaEff frozen
                        ve = ve + ((F.balphat[idx] + F.balpha[idx]) * ms * m.ellDc[e]) * (bs.h - I.h);
                }
                Jcc[0 * 5 + kk] += o.phi.d * rV;
                Jcc[1 * 5 + kk] += (o.fx.d - vx.d) * rV;
                Jcc[2 * 5 + kk] += (o.fy.d - vy.d) * rV;
                Jcc[3 * 5 + kk] += (o.fz.d - vz.d) * rV;
                Jcc[4 * 5 + kk] += (o.fE.d - ve.d) * rV;
            }
            continue;
        }
        const bool own = kind == ELL_OWNER;
        const double sgn = own ? rV : -rV;
        const int fi = m.ellFace[e];
        const int sco = own ? m.slotUpper[fi] : m.slotLower[fi];
        // frozen viscous face data: interp(muEff)|S| deltaCoeffs, the
        // correction part of snGrad(U) and Sf & interp(tauMC)
        double mgd = 0.0, cgx = 0.0, cgy = 0.0, cgz = 0.0, tfx = 0.0, tfy = 0.0, tfz = 0.0, mg = 0.0, ahd = 0.0;
        if constexpr (VISC) {
            const double w = m.ellWt[e];
            const double muC = F.mut[c] + F.mu[c], muO = F.mut[idx] + F.mu[idx];
            const double muP = own ? muC : muO, muN = own ? muO : muC;
            const double muf = w * (muP - muN) + muN;
            double gP[9], gN[9], tP[9], tN[9], gf[9], tf[9];
            for (int q = 0; q < 9; ++q) {
                gP[q] = F.gradU[(size_t)q * m.nC + (own ? c : idx)];
                gN[q] = F.gradU[(size_t)q * m.nC + (own ? idx : c)];
            }
            tau_mc_a(muP, gP, tP);
            tau_mc_a(muN, gN, tN);
            for (int q = 0; q < 9; ++q) {
                gf[q] = w * (gP[q] - gN[q]) + gN[q];
                tf[q] = w * (tP[q] - tN[q]) + tN[q];
            }
            const double cx = m.ellCorr[e], cy = m.ellCorr[E + e], cz = m.ellCorr[2 * E + e];
            cgx = cx * gf[0] + cy * gf[3] + cz * gf[6];
            cgy = cx * gf[1] + cy * gf[4] + cz * gf[7];
            cgz = cx * gf[2] + cy * gf[5] + cz * gf[8];
            mg = muf * ms;
            mgd = m.ellDc[e];
            tfx = sx * tf[0] + sy * tf[3] + sz * tf[6];
            tfy = sx * tf[1] + sy * tf[4] + sz * tf[7];
            tfz = sx * tf[2] + sy * tf[5] + sz * tf[8];
            if (heat) {  // conduction extension: interp(alphaEff) |S| dc, frozen
                const double alC = F.alphat[c] + F.alpha[c], alO = F.alphat[idx] + F.alpha[idx];
                const double alP = own ? alC : alO, alN = own ? alO : alC;
                ahd = (w * (alP - alN) + alN) * ms * mgd;
            }
        }
        // derivative w.r.t. this cell's W, then w.r.t. the other cell's W
        for (int side = 0; side < 2; ++side) {
            for (int kk = 0; kk < 5; ++kk) {
                const Prim<D1> Pc = seeded(F.W, c, side == 0 ? kk : -1, th, bd);
                const Prim<D1> Po = seeded(F.W, idx, side == 1 ? kk : -1, th, bd);
                const SideT<D1> sc = side_of(Pc, gamma), so = side_of(Po, gamma);
                const Flux5T<D1> o = own ? eval_flux<S>(sc, so, sx, sy, sz, ms, gamma, 1.0, lowMach, Mainf)
                                         : eval_flux<S>(so, sc, sx, sy, sz, ms, gamma, 1.0, lowMach, Mainf);
                double d[5] = {o.phi.d, o.fx.d, o.fy.d, o.fz.d, o.fE.d};
                if constexpr (VISC) {
                    const Prim<D1>& QP = own ? Pc : Po;
                    const Prim<D1>& QN = own ? Po : Pc;
                    const D1 lx = mg * (mgd * (QN.ux - QP.ux) + cgx);
                    const D1 ly = mg * (mgd * (QN.uy - QP.uy) + cgy);
                    const D1 lz = mg * (mgd * (QN.uz - QP.uz) + cgz);
                    const D1 sdf = (lx + tfx) * o.upx + (ly + tfy) * o.upy + (lz + tfz) * o.upz;
                    d[1] -= lx.d; d[2] -= ly.d; d[3] -= lz.d; d[4] -= sdf.d;
                    if (heat) d[4] -= (ahd * (QN.h - QP.h)).d;
                }
                if (side == 0)
                    for (int r = 0; r < 5; ++r) Jcc[r * 5 + kk] += sgn * d[r];
                else
                    for (int r = 0; r < 5; ++r) J[at(sco, r * 5 + kk)] = (TJ)(sgn * d[r]);
            }
        }
    }
    const int sd = m.slotDiag[c];
    const double dd = NEWTON ? F.rDeltaT[c] * m.V[c] : 0.0;
    for (int q = 0; q < 25; ++q) J[at(sd, q)] = (TJ)(Jcc[q] + ((q % 6 == 0) ? dd : 0.0));
    if (NEWTON) {
        const double V = m.V[c];
        for (int q = 0; q < 5; ++q) src[5 * (size_t)c + q] = 0.0 - F.R[(size_t)q * m.nC + c] * V;
    }
}

// the residual itself from W by the same chain (the value parts of the duals)
template <int S>
__global__ __launch_bounds__(256) void k_adj_residual(MeshDev m, Fields F, const BcPatchDev* bc, Thermo th, Bounds bd,
                                                      int lowMach, double Mainf, double* R) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const size_t E = (size_t)m.ellW * m.nC;
    const double gamma = th.gamma;
    const Prim<double> Pc = prim_from_w(F.W[5 * (size_t)c], F.W[5 * (size_t)c + 1], F.W[5 * (size_t)c + 2],
                                        F.W[5 * (size_t)c + 3], F.W[5 * (size_t)c + 4], th, bd);
    const SideT<double> sc = side_of(Pc, gamma);
    double a[5] = {0, 0, 0, 0, 0};
    for (int ke = 0; ke < m.ellW; ++ke) {
        const size_t e = (size_t)ke * m.nC + c;
        const uint32_t code = m.ellCode[e];
        const uint32_t kind = code >> 30;
        if (kind == ELL_NONE) break;
        const int idx = (int)(code & ELL_IDX_MASK);
        const double sx = m.ellSf[e], sy = m.ellSf[E + e], sz = m.ellSf[2 * E + e], ms = m.ellMagSf[e];
        Flux5T<double> o;
        double sg = 1.0;
        if (kind == ELL_BOUNDARY) {
            const BState<double> bs = bc_state(bc[m.bpatch[idx]], F, idx, m.nI + idx, sx / ms, sy / ms, sz / ms, ms, Pc,
                                               th, bd);
            const SideT<double> L = side_of(bs, gamma);
            o = eval_flux<S>(L, L, sx, sy, sz, ms, gamma, 1.0, lowMach, Mainf);
        } else {
            const double* Wo = F.W + 5 * (size_t)idx;
            const Prim<double> Po = prim_from_w(Wo[0], Wo[1], Wo[2], Wo[3], Wo[4], th, bd);
            const SideT<double> so = side_of(Po, gamma);
            if (kind == ELL_OWNER) o = eval_flux<S>(sc, so, sx, sy, sz, ms, gamma, 1.0, lowMach, Mainf);
            else { o = eval_flux<S>(so, sc, sx, sy, sz, ms, gamma, 1.0, lowMach, Mainf); sg = -1.0; }
        }
        a[0] += sg * o.phi; a[1] += sg * o.fx; a[2] += sg * o.fy; a[3] += sg * o.fz; a[4] += sg * o.fE;
    }
    const double V = m.V[c];
    for (int r = 0; r < 5; ++r) R[5 * (size_t)c + r] = a[r] / V;
}

// dJ/dW of J = sum_{b in patch} (p_b - pRef)(Sf_b . dir)/q
__global__ __launch_bounds__(256) void k_adj_objective(MeshDev m, Fields F, const BcPatchDev* bc, Thermo th, Bounds bd,
                                                       int patch, double dx, double dy, double dz, double rq, double* g) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const size_t E = (size_t)m.ellW * m.nC;
    double gc[5] = {0, 0, 0, 0, 0};
    for (int ke = 0; ke < m.ellW; ++ke) {
        const size_t e = (size_t)ke * m.nC + c;
        const uint32_t code = m.ellCode[e];
        const uint32_t kind = code >> 30;
        if (kind == ELL_NONE) break;
        if (kind != ELL_BOUNDARY) continue;
        const int idx = (int)(code & ELL_IDX_MASK);
        if (m.bpatch[idx] != patch) continue;
        const double sx = m.ellSf[e], sy = m.ellSf[E + e], sz = m.ellSf[2 * E + e], ms = m.ellMagSf[e];
        const double w = (sx * dx + sy * dy + sz * dz) * rq;
        for (int kk = 0; kk < 5; ++kk) {
            const Prim<D1> I = seeded(F.W, c, kk, th, bd);
            const BState<D1> bs = bc_state(bc[patch], F, idx, m.nI + idx, sx / ms, sy / ms, sz / ms, ms, I, th, bd);
            gc[kk] += bs.p.d * w;
        }
    }
    for (int kk = 0; kk < 5; ++kk) g[5 * (size_t)c + kk] = gc[kk];
}

// J^T on the same slots: (r, c) <- (c, r)^T
__global__ void k_transpose_faces(int nI, const int32_t* slotUpper, const int32_t* slotLower, const double* J,
                                  double* JT) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nI) return;
    const double* U = J + 25 * (size_t)slotUpper[f];
    const double* L = J + 25 * (size_t)slotLower[f];
    double* TU = JT + 25 * (size_t)slotUpper[f];
    double* TL = JT + 25 * (size_t)slotLower[f];
    for (int r = 0; r < 5; ++r)
        for (int q = 0; q < 5; ++q) {
            TU[r * 5 + q] = L[q * 5 + r];
            TL[r * 5 + q] = U[q * 5 + r];
        }
}
__global__ void k_transpose_diag(int n, const int32_t* slotDiag, const double* J, double* JT) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double* D = J + 25 * (size_t)slotDiag[c];
    double* T = JT + 25 * (size_t)slotDiag[c];
    for (int r = 0; r < 5; ++r)
        for (int q = 0; q < 5; ++q) T[r * 5 + q] = D[q * 5 + r];
}

Bounds bounds_of(const Case& c) {
    Bounds b;
    b.rhoMin = c.st.rho_min; b.rhoMax = c.st.rho_max;
    b.hMin = c.th.Cp * c.st.T_min; b.hMax = c.th.Cp * c.st.T_max;
    b.pMin = c.st.p_min; b.pMax = c.st.p_max;
    return b;
}

template <int S, class TJ, bool NEWTON>
void launch_jac(Case& c, TJ* J) {
    const int nb = blocks_for(c.mesh->nC, 256);
    if (NEWTON && c.st.viscous)
        hipLaunchKernelGGL((k_adj_jacobian<S, TJ, NEWTON, true>), dim3(nb), dim3(256), 0, c.stream, c.meshDev(), c.f,
                           c.d_bc.p, c.th, bounds_of(c), c.st.low_mach, c.st.Mainf, J, c.src.p, c.mesh->d_sellOfSlot.p,
                           c.st.heat_conduction);
    else
        hipLaunchKernelGGL((k_adj_jacobian<S, TJ, NEWTON, false>), dim3(nb), dim3(256), 0, c.stream, c.meshDev(), c.f,
                           c.d_bc.p, c.th, bounds_of(c), c.st.low_mach, c.st.Mainf, J,
                           NEWTON ? c.src.p : (double*)nullptr, c.mesh->d_sellOfSlot.p, 0);
}
template <class TJ, bool NEWTON>
void launch_jac_scheme(Case& c, TJ* J) {
    switch (c.st.flux_scheme) {
        case RDST_FLUX_AUSMPLUS: launch_jac<RDST_FLUX_AUSMPLUS, TJ, NEWTON>(c, J); break;
        case RDST_FLUX_AUSMPLUSUP: launch_jac<RDST_FLUX_AUSMPLUSUP, TJ, NEWTON>(c, J); break;
        case RDST_FLUX_HLL: launch_jac<RDST_FLUX_HLL, TJ, NEWTON>(c, J); break;
        case RDST_FLUX_HLLC: launch_jac<RDST_FLUX_HLLC, TJ, NEWTON>(c, J); break;
        default: launch_jac<RDST_FLUX_KURGANOV, TJ, NEWTON>(c, J); break;
    }
    RDST_HIP(hipGetLastError());
}
template <int S>
void launch_res(Case& c, double* R) {
    const int nb = blocks_for(c.mesh->nC, 256);
    hipLaunchKernelGGL((k_adj_residual<S>), dim3(nb), dim3(256), 0, c.stream, c.meshDev(), c.f, c.d_bc.p, c.th,
                       bounds_of(c), c.st.low_mach, c.st.Mainf, R);
}

}  // namespace

void adjoint_check_scheme(const Case& c) {
    if (c.st.flux_scheme == RDST_FLUX_HLLCP)
        throw ArgError("adjoint: HLLCP's pressure sensor couples second neighbours; use AUSM+, AUSM+up, HLL, HLLC or "
                       "Kurganov", RDST_ERR_UNKNOWN_NAME);
    if (c.st.reconstruction != RDST_RECON_UPWIND)
        throw ArgError("adjoint: the exact linearisation covers first-order (upwind) reconstruction");
    if (c.st.viscous || c.sa.on)
        throw ArgError("adjoint: the exact linearisation covers the inviscid residual (settings.viscous = 0, laminar)");
}

void adjoint_jacobian(Case& c) {
    adjoint_check_scheme(c);
    const Mesh& m = *c.mesh;
    const size_t nn = 25 * (size_t)m.nnzb;
    if (c.adjJ.n != nn) c.adjJ.alloc(nn);
    c.adjJ.zero(c.stream);
    launch_jac_scheme<double, false>(c, c.adjJ.p);
}

void newton_operator(Case& c) {
    if (c.st.flux_scheme == RDST_FLUX_HLLCP || c.st.reconstruction != RDST_RECON_UPWIND)
        throw ArgError("exact Jacobian: first-order (upwind) reconstruction and a flux other than HLLCP");
    const Mesh& m = *c.mesh;
    const size_t nn = 25 * (size_t)m.nnzb;
    // every slot is written (a cell's diagonal and its side of each face);
    // the padding blocks of the slices stay zero from the allocation.  The
    // block-CSR buffers keep their identity for the SpMV lookup and receive
    // the operator only when the matrix is read back (newton_operator_csr).
    const size_t nv = 25 * (size_t)m.sellBlocks;
    if (c.st.precision == RDST_PREC_MIXED) {
        if (c.A32.n != nn) c.A32.alloc(nn);
        if (c.S32.n != nv) {
            c.S32.alloc(nv);
            c.S32.zero(c.stream);
        }
        launch_jac_scheme<float, true>(c, c.S32.p);
    } else {
        if (c.A64.n != nn) c.A64.alloc(nn);
        if (c.S64.n != nv) {
            c.S64.alloc(nv);
            c.S64.zero(c.stream);
        }
        launch_jac_scheme<double, true>(c, c.S64.p);
    }
    c.sellFresh = true;
    c.csrStale = true;
}

namespace {
template <class T>
__global__ void k_from_sell(long nnzb, const int64_t* sellOfSlot, const T* S, T* A) {
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < nnzb; j += (long)gridDim.x * blockDim.x) {
        const int64_t base = sellOfSlot[j];
        for (int e = 0; e < 25; ++e) A[25 * (size_t)j + e] = S[base + 64 * e];
    }
}
}  // namespace

// the FP64 Newton operator in block-CSR (A64), for a read-back
void newton_operator_csr(Case& c) {
    if (!c.csrStale || c.st.precision == RDST_PREC_MIXED) return;
    const Mesh& m = *c.mesh;
    hipLaunchKernelGGL((k_from_sell<double>), dim3(std::min(4096, blocks_for(m.nnzb, 256))), dim3(256), 0, c.stream,
                       (long)m.nnzb, m.d_sellOfSlot.p, c.S64.p, c.A64.p);
    RDST_HIP(hipGetLastError());
    RDST_HIP(hipStreamSynchronize(c.stream));
    c.csrStale = false;
}

void adjoint_linearise(Case& c) {
    adjoint_jacobian(c);
    const Mesh& m = *c.mesh;
    const size_t nn = 25 * (size_t)m.nnzb;
    if (c.adjJT.n != nn) c.adjJT.alloc(nn);
    if (m.nI)
        hipLaunchKernelGGL(k_transpose_faces, dim3(blocks_for(m.nI, 256)), dim3(256), 0, c.stream, m.nI, m.d_slotUpper.p,
                           m.d_slotLower.p, c.adjJ.p, c.adjJT.p);
    hipLaunchKernelGGL(k_transpose_diag, dim3(blocks_for(m.nC, 256)), dim3(256), 0, c.stream, m.nC, m.d_slotDiag.p,
                       c.adjJ.p, c.adjJT.p);
    RDST_HIP(hipGetLastError());
    c.hasAdjJ = true;
}

void adjoint_residual(Case& c, double* R) {
    adjoint_check_scheme(c);
    switch (c.st.flux_scheme) {
        case RDST_FLUX_AUSMPLUS: launch_res<RDST_FLUX_AUSMPLUS>(c, R); break;
        case RDST_FLUX_AUSMPLUSUP: launch_res<RDST_FLUX_AUSMPLUSUP>(c, R); break;
        case RDST_FLUX_HLL: launch_res<RDST_FLUX_HLL>(c, R); break;
        case RDST_FLUX_HLLC: launch_res<RDST_FLUX_HLLC>(c, R); break;
        default: launch_res<RDST_FLUX_KURGANOV>(c, R); break;
    }
    RDST_HIP(hipGetLastError());
}

void adjoint_objective_gradient(Case& c, int patch, const double* dir, double q, double* g) {
    adjoint_check_scheme(c);
    const Mesh& m = *c.mesh;
    hipLaunchKernelGGL(k_adj_objective, dim3(blocks_for(m.nC, 256)), dim3(256), 0, c.stream, c.meshDev(), c.f, c.d_bc.p,
                       c.th, bounds_of(c), patch, dir[0], dir[1], dir[2], 1.0 / q, g);
    RDST_HIP(hipGetLastError());
}

}  // namespace rdst
