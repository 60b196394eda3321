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
#include <stdexcept>
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
    T ux, uy, uz, Tb;  // the boundary velocity and temperature themselves
};

__device__ inline double mpos0(double s) { return s >= 0.0 ? 1.0 : 0.0; }
__device__ inline double mneg(double s) { return s < 0.0 ? 1.0 : 0.0; }
__device__ inline double mneg0(double s) { return s <= 0.0 ? 1.0 : 0.0; }

// the boundary state of face b from its cell (update.hip k_update_bnd_pre /
// k_update_bnd_post with old = new); masks from the stored phi, rho_b, psi_b
// refU (optional): the refValue of the mixed U conditions (inletOutlet: the
// inlet value; far field: the characteristic state), which their snGrad reads
template <class T>
__device__ inline BState<T> bc_state(const BcPatchDev& B, const Fields& F, int b, int k, double nx, double ny, double nz,
                                     double ms, const Prim<T>& I, const Thermo& th, const Bounds& bd,
                                     T* refU = nullptr) {
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
            if (refU) { refU[0] = T(B.U_value[0]); refU[1] = T(B.U_value[1]); refU[2] = T(B.U_value[2]); }
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
            if (refU) { refU[0] = rx; refU[1] = ry; refU[2] = rz; }
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
    s.ux = ux; s.uy = uy; s.uz = uz; s.Tb = Tb;
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

// each cell's share of J = sum_{b in patch} (p_b - pRef)(Sf_b . dir)/q (values)
__global__ __launch_bounds__(256) void k_adj_objective_cells(MeshDev m, Fields F, const BcPatchDev* bc, Thermo th,
                                                             Bounds bd, int patch, double dx, double dy, double dz,
                                                             double rq, double pRef, double* o) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const size_t E = (size_t)m.ellW * m.nC;
    const Prim<double> I = prim_from_w(F.W[5 * (size_t)c], F.W[5 * (size_t)c + 1], F.W[5 * (size_t)c + 2],
                                       F.W[5 * (size_t)c + 3], F.W[5 * (size_t)c + 4], th, bd);
    double s = 0.0;
    for (int ke = 0; ke < m.ellW; ++ke) {
        const size_t e = (size_t)ke * m.nC + c;
        const uint32_t code = m.ellCode[e];
        const uint32_t kind = code >> 30;
        if (kind == ELL_NONE) break;
        if (kind != ELL_BOUNDARY) continue;
        const int idx = (int)(code & ELL_IDX_MASK);
        if (m.bpatch[idx] != patch) continue;
        const double sx = m.ellSf[e], sy = m.ellSf[E + e], sz = m.ellSf[2 * E + e], ms = m.ellMagSf[e];
        const BState<double> bs = bc_state(bc[patch], F, idx, m.nI + idx, sx / ms, sy / ms, sz / ms, ms, I, th, bd);
        s += (bs.p - pRef) * ((sx * dx + sy * dy + sz * dz) * rq);
    }
    o[c] = s;
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

// ---------------------------------------------------------------- viscous residual
// The viscous face terms of calcFields (baseFluxDST.C:229-250) as the face
// loop forms them (the laplacian of muEff U with its non-orthogonal
// correction, Sf & interp(tauMC) with tauMC = muEff dev2(T(grad U)), their
// work on the interface velocity, and the conduction extension), with
// fvc::grad(U) (Gauss linear) and its corrected boundary value, on a scalar
// type: the laminar viscosity and conductivity follow the cell or boundary
// temperature (constant or Sutherland); the turbulent ones (mut, alphat and
// their boundary values) are frozen at their stored values.  The cell's
// residual then depends on its second face neighbours (through their
// gradients), so J is formed on the second-neighbour pattern by forward-mode
// passes: one pass per colour of a distance-4 colouring and per component
// seeds that component of every cell of the colour; row r's block of column
// j is read from row r's tangent, j being the one column of that colour in
// row r.
template <class T>
__device__ inline T mu_lam(const Thermo& th, const T& Tt) {
    if (th.transport == 1) return th.As * sqrt(Tt) / (1.0 + th.Ts / Tt);
    return T(th.mu);
}
template <class T>
__device__ inline T alpha_lam(const Thermo& th, const T& Tt) {
    if (th.transport == 1) return mu_lam(th, Tt) * th.Cv * (1.32 + 1.77 * th.R / th.Cv) / th.Cp;
    return T(th.mu / th.Pr);
}
template <class T>
__device__ inline void tau_mc_t(const T& mu, const T* g, T* t) {
    const T tr = g[0] + g[4] + g[8];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) t[3 * i + j] = mu * (g[3 * j + i] - (i == j ? (2.0 / 3.0) * tr : T(0.0)));
}
template <class T>
__device__ inline T sngrad_limiter_t(double psi, const T& u2, const T& c2) {
    if (psi < 0.0) return T(1.0);
    return fmin(psi * sqrt(u2) / ((1.0 - psi) * sqrt(c2) + 1.0e-15), T(1.0));
}
// snGrad of the U condition with the refValue of the mixed conditions as a
// function of the state (ref from bc_state)
template <class T>
__device__ inline void sngrad_u_bnd_r(const BcPatchDev& B, const Fields& F, int b, double nx, double ny, double nz,
                                      double dcb, const T& uIx, const T& uIy, const T& uIz, const T* ref, T* o) {
    switch (B.U_type) {
        case RDST_BC_FIXED_VALUE:
            o[0] = dcb * (F.bUx[b] - uIx); o[1] = dcb * (F.bUy[b] - uIy); o[2] = dcb * (F.bUz[b] - uIz);
            return;
        case RDST_BC_SYMMETRY: {
            const T nU = nx * uIx + ny * uIy + nz * uIz;
            const T tx = uIx - 2.0 * nx * nU, ty = uIy - 2.0 * ny * nU, tz = uIz - 2.0 * nz * nU;
            const double h = dcb / 2.0;
            o[0] = (tx - uIx) * h; o[1] = (ty - uIy) * h; o[2] = (tz - uIz) * h;
            return;
        }
        case RDST_BC_INLET_OUTLET: case RDST_BC_FARFIELD: {
            const double vf = F.vf_U[b];
            o[0] = vf * (ref[0] - uIx) * dcb; o[1] = vf * (ref[1] - uIy) * dcb; o[2] = vf * (ref[2] - uIz) * dcb;
            return;
        }
        default: o[0] = T(0.0); o[1] = T(0.0); o[2] = T(0.0);
    }
}

__device__ inline int seed_of(const int32_t* color, int i, int col, int comp) {
    return (col >= 0 && color[i] == col) ? comp : -1;
}

// fvc::grad(U) of every cell, value and tangent: gd[q n + c], gd[(9 + q) n + c]
__global__ __launch_bounds__(256) void k_av_grad(MeshDev m, Fields F, const BcPatchDev* bc, Thermo th, Bounds bd,
                                                 const int32_t* color, int col, int comp, double* gd) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const int n = m.nC;
    const size_t E = (size_t)m.ellW * m.nC;
    const Prim<D1> Pc = seeded(F.W, c, seed_of(color, c, col, comp), th, bd);
    D1 g[9];
    for (int ke = 0; ke < m.ellW; ++ke) {
        const size_t e = (size_t)ke * m.nC + c;
        const uint32_t code = m.ellCode[e];
        const uint32_t kind = code >> 30;
        if (kind == ELL_NONE) break;
        const int idx = (int)(code & ELL_IDX_MASK);
        const double S[3] = {m.ellSf[e], m.ellSf[E + e], m.ellSf[2 * E + e]};
        D1 f[3];
        bool neg = false;
        if (kind == ELL_BOUNDARY) {
            const double ms = m.ellMagSf[e];
            const BState<D1> bs =
                bc_state(bc[m.bpatch[idx]], F, idx, m.nI + idx, S[0] / ms, S[1] / ms, S[2] / ms, ms, Pc, th, bd);
            f[0] = bs.ux; f[1] = bs.uy; f[2] = bs.uz;
        } else {
            const Prim<D1> Po = seeded(F.W, idx, seed_of(color, idx, col, comp), th, bd);
            const double w = m.ellWt[e];
            const bool own = kind == ELL_OWNER;
            const Prim<D1>& QP = own ? Pc : Po;
            const Prim<D1>& QN = own ? Po : Pc;
            f[0] = w * (QP.ux - QN.ux) + QN.ux;
            f[1] = w * (QP.uy - QN.uy) + QN.uy;
            f[2] = w * (QP.uz - QN.uz) + QN.uz;
            neg = !own;
        }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const D1 t = S[i] * f[j];
                if (neg) g[3 * i + j] -= t; else g[3 * i + j] += t;
            }
    }
    const double V = m.V[c];
    for (int q = 0; q < 9; ++q) {
        const D1 r = g[q] / V;
        gd[(size_t)q * n + c] = r.v;
        gd[(size_t)(9 + q) * n + c] = r.d;
    }
}

// the residual (inviscid flux + viscous terms) of every cell, value Rv and
// tangent Jv [n][5], and the tangent of the cell's share of the force
// objective (its faces on the patch: pressure and viscous force . dir / q)
template <int S>
__global__ __launch_bounds__(256) void k_av_resid(MeshDev m, Fields F, const BcPatchDev* bc, Thermo th, Bounds bd,
                                                  int lowMach, double Mainf, double lapPsi, double snPsi, int heat,
                                                  const int32_t* color, int col, int comp, const double* gd,
                                                  double* Rv, double* Jv, int patch, double dx, double dy, double dz,
                                                  double rq, double pRef, double* objd, double* objv) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const int n = m.nC;
    const size_t E = (size_t)m.ellW * m.nC;
    const double gamma = th.gamma;
    const Prim<D1> Pc = seeded(F.W, c, seed_of(color, c, col, comp), th, bd);
    const SideT<D1> sc = side_of(Pc, gamma);
    D1 gC[9], tauC[9];
    for (int q = 0; q < 9; ++q) gC[q] = D1(gd[(size_t)q * n + c], gd[(size_t)(9 + q) * n + c]);
    const D1 muC = F.mut[c] + mu_lam(th, Pc.Tt);
    tau_mc_t(muC, gC, tauC);
    const D1 alC = heat ? F.alphat[c] + alpha_lam(th, Pc.Tt) : D1(0.0);
    D1 a[5], lx, ly, lz, tx, ty, tz, sd, hq, ob;
    for (int ke = 0; ke < m.ellW; ++ke) {
        const size_t e = (size_t)ke * m.nC + c;
        const uint32_t code = m.ellCode[e];
        const uint32_t kind = code >> 30;
        if (kind == ELL_NONE) break;
        const int idx = (int)(code & ELL_IDX_MASK);
        const double sx = m.ellSf[e], sy = m.ellSf[E + e], sz = m.ellSf[2 * E + e], ms = m.ellMagSf[e];
        Flux5T<D1> o;
        D1 lf[3], sg[3], tf[3], qh;
        if (kind == ELL_BOUNDARY) {
            const BcPatchDev& B = bc[m.bpatch[idx]];
            D1 ref[3];
            const BState<D1> bs = bc_state(B, F, idx, m.nI + idx, sx / ms, sy / ms, sz / ms, ms, Pc, th, bd, ref);
            const SideT<D1> L = side_of(bs, gamma);
            o = eval_flux<S>(L, L, sx, sy, sz, ms, gamma, 1.0, lowMach, Mainf);
            const double rms = 1.0 / ms;
            D1 sn[3];
            sngrad_u_bnd_r(B, F, idx, sx * rms, sy * rms, sz * rms, m.ellDc[e], Pc.ux, Pc.uy, Pc.uz, ref, sn);
            const D1 bmuE = F.bmut[idx] + mu_lam(th, bs.Tb);
            const D1 mg = bmuE * ms;
            for (int j = 0; j < 3; ++j) { lf[j] = mg * sn[j]; sg[j] = lf[j]; }
            // gaussGrad::correctBoundaryConditions: g_b = g_I + n (snGrad_b - n.g_I)
            const double bn[3] = {sx / ms, sy / ms, sz / ms};
            D1 gb[9], tb[9];
            for (int j = 0; j < 3; ++j) {
                const D1 nG = bn[0] * gC[j] + bn[1] * gC[3 + j] + bn[2] * gC[6 + j];
                const D1 d = sn[j] - nG;
                for (int i = 0; i < 3; ++i) gb[3 * i + j] = gC[3 * i + j] + bn[i] * d;
            }
            tau_mc_t(bmuE, gb, tb);
            tf[0] = sx * tb[0] + sy * tb[3] + sz * tb[6];
            tf[1] = sx * tb[1] + sy * tb[4] + sz * tb[7];
            tf[2] = sx * tb[2] + sy * tb[5] + sz * tb[8];
            if (heat) qh = (F.balphat[idx] + alpha_lam(th, bs.Tb)) * ms * (m.ellDc[e] * (bs.h - Pc.h));
            if (m.bpatch[idx] == patch) {
                // forceCoeffs: (p_b - p_ref) Sf + Sf & devRhoReff_b, devRhoReff = -muEff dev(twoSymm(grad U))
                ob += (bs.p - pRef) * ((sx * dx + sy * dy + sz * dz) * rq);
                const D1 tr = gb[0] + gb[4] + gb[8];
                const double S3[3] = {sx, sy, sz}, dir[3] = {dx, dy, dz};
                for (int j = 0; j < 3; ++j) {
                    D1 v(0.0);
                    for (int i = 0; i < 3; ++i) {
                        const D1 ts = gb[3 * i + j] + gb[3 * j + i];
                        v += S3[i] * (-bmuE * (ts - (i == j ? (2.0 / 3.0) * tr : D1(0.0))));
                    }
                    ob += v * (dir[j] * rq);
                }
            }
        } else {
            const bool own = kind == ELL_OWNER;
            const Prim<D1> Po = seeded(F.W, idx, seed_of(color, idx, col, comp), th, bd);
            const SideT<D1> so = side_of(Po, gamma);
            o = own ? eval_flux<S>(sc, so, sx, sy, sz, ms, gamma, 1.0, lowMach, Mainf)
                    : eval_flux<S>(so, sc, sx, sy, sz, ms, gamma, 1.0, lowMach, Mainf);
            const double w = m.ellWt[e];
            const D1 muO = F.mut[idx] + mu_lam(th, Po.Tt);
            D1 gO[9], tauO[9];
            for (int q = 0; q < 9; ++q) gO[q] = D1(gd[(size_t)q * n + idx], gd[(size_t)(9 + q) * n + idx]);
            tau_mc_t(muO, gO, tauO);
            const D1* gP = own ? gC : gO;
            const D1* gN = own ? gO : gC;
            const D1* tP = own ? tauC : tauO;
            const D1* tN = own ? tauO : tauC;
            const D1 muP = own ? muC : muO, muN = own ? muO : muC;
            const Prim<D1>& QP = own ? Pc : Po;
            const Prim<D1>& QN = own ? Po : Pc;
            const D1 muf = w * (muP - muN) + muN;
            D1 gf[9], tfc[9];
            for (int q = 0; q < 9; ++q) {
                gf[q] = w * (gP[q] - gN[q]) + gN[q];
                tfc[q] = w * (tP[q] - tN[q]) + tN[q];
            }
            const double cx = m.ellCorr[e], cy = m.ellCorr[E + e], cz = m.ellCorr[2 * E + e];
            const double dcf = m.ellDc[e];
            const D1 ux = dcf * (QN.ux - QP.ux), uy = dcf * (QN.uy - QP.uy), uz = dcf * (QN.uz - QP.uz);
            const D1 qx = cx * gf[0] + cy * gf[3] + cz * gf[6];
            const D1 qy = cx * gf[1] + cy * gf[4] + cz * gf[7];
            const D1 qz = cx * gf[2] + cy * gf[5] + cz * gf[8];
            D1 kL(1.0), kS(1.0);
            if (lapPsi >= 0.0 || snPsi >= 0.0) {
                const D1 u2 = ux * ux + uy * uy + uz * uz, q2 = qx * qx + qy * qy + qz * qz;
                kL = sngrad_limiter_t(lapPsi, u2, q2);
                kS = sngrad_limiter_t(snPsi, u2, q2);
            }
            const D1 mg = muf * ms;
            lf[0] = mg * (ux + kL * qx); lf[1] = mg * (uy + kL * qy); lf[2] = mg * (uz + kL * qz);
            sg[0] = mg * (ux + kS * qx); sg[1] = mg * (uy + kS * qy); sg[2] = mg * (uz + kS * qz);
            tf[0] = sx * tfc[0] + sy * tfc[3] + sz * tfc[6];
            tf[1] = sx * tfc[1] + sy * tfc[4] + sz * tfc[7];
            tf[2] = sx * tfc[2] + sy * tfc[5] + sz * tfc[8];
            if (heat) {
                const D1 alO = F.alphat[idx] + alpha_lam(th, Po.Tt);
                const D1 alP = own ? alC : alO, alN = own ? alO : alC;
                qh = (w * (alP - alN) + alN) * ms * (dcf * (QN.h - QP.h));
            }
        }
        const D1 sdf = (sg[0] + tf[0]) * o.upx + (sg[1] + tf[1]) * o.upy + (sg[2] + tf[2]) * o.upz;
        if (kind == ELL_NEIGH) {
            a[0] -= o.phi; a[1] -= o.fx; a[2] -= o.fy; a[3] -= o.fz; a[4] -= o.fE;
            lx -= lf[0]; ly -= lf[1]; lz -= lf[2]; tx -= tf[0]; ty -= tf[1]; tz -= tf[2]; sd -= sdf; hq -= qh;
        } else {
            a[0] += o.phi; a[1] += o.fx; a[2] += o.fy; a[3] += o.fz; a[4] += o.fE;
            lx += lf[0]; ly += lf[1]; lz += lf[2]; tx += tf[0]; ty += tf[1]; tz += tf[2]; sd += sdf; hq += qh;
        }
    }
    const double V = m.V[c];
    D1 r[5];
    r[0] = a[0] / V;
    r[1] = a[1] / V - (lx / V + tx / V);
    r[2] = a[2] / V - (ly / V + ty / V);
    r[3] = a[3] / V - (lz / V + tz / V);
    r[4] = a[4] / V - sd / V;
    if (heat) r[4] -= hq / V;
    for (int q = 0; q < 5; ++q) {
        if (Rv) Rv[5 * (size_t)c + q] = r[q].v;
        if (Jv) Jv[5 * (size_t)c + q] = r[q].d;
    }
    if (objd) objd[c] = ob.d;
    if (objv) objv[c] = ob.v;
}

// column comp of the blocks of this pass's colour: row r's tangent is the
// block (r, j) of the one column j of the colour in row r
__global__ void k_av_scatter(int n, const int32_t* rowPtr, const int32_t* cols, const int32_t* color, int col,
                             int comp, const double* Jv, double* Jw) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    for (int e = rowPtr[r]; e < rowPtr[r + 1]; ++e)
        if (color[cols[e]] == col) {
            for (int q = 0; q < 5; ++q) Jw[25 * (size_t)e + 5 * q + comp] = Jv[5 * (size_t)r + q];
            return;
        }
}

// dJ/dW of the cells of this pass's colour: the objective share of a cell
// depends on the cell and its face neighbours, so cell j collects the
// tangents of itself and of its neighbours
__global__ void k_av_obj_gather(MeshDev m, const int32_t* color, int col, int comp, const double* objd, double* g) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m.nC || color[j] != col) return;
    double s = objd[j];
    for (int ke = 0; ke < m.ellW; ++ke) {
        const size_t e = (size_t)ke * m.nC + j;
        const uint32_t code = m.ellCode[e];
        const uint32_t kind = code >> 30;
        if (kind == ELL_NONE) break;
        if (kind == ELL_BOUNDARY) continue;
        s += objd[code & ELL_IDX_MASK];
    }
    g[5 * (size_t)j + comp] = s;
}

// J^T on the second-neighbour pattern: block e of J, transposed, at wideT[e]
__global__ void k_wide_transpose(long nnz, const int32_t* wideT, const double* J, double* JT) {
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (long)gridDim.x * blockDim.x) {
        const double* a = J + 25 * (size_t)e;
        double* t = JT + 25 * (size_t)wideT[e];
        for (int r = 0; r < 5; ++r)
            for (int q = 0; q < 5; ++q) t[5 * r + q] = a[5 * q + r];
    }
}
// the block-CSR slots of a second-neighbour matrix
__global__ void k_wide_to_slots(long nnzb, const int32_t* wideOfSlot, const double* W, double* A) {
    for (long s = (long)blockIdx.x * blockDim.x + threadIdx.x; s < nnzb; s += (long)gridDim.x * blockDim.x)
        for (int q = 0; q < 25; ++q) A[25 * (size_t)s + q] = W[25 * (size_t)wideOfSlot[s] + q];
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

template <int S>
void launch_av_resid(Case& c, int col, int comp, double* Rv, double* Jv, int patch, const double* dir, double rq,
                     double* objd, double pRef = 0.0, double* objv = nullptr) {
    const Mesh& m = *c.mesh;
    const int nb = blocks_for(m.nC, 256);
    hipLaunchKernelGGL((k_av_resid<S>), dim3(nb), dim3(256), 0, c.stream, c.meshDev(), c.f, c.d_bc.p, c.th, bounds_of(c),
                       c.st.low_mach, c.st.Mainf, sngrad_psi(c.st.laplacian_sngrad, c.st.laplacian_limit),
                       sngrad_psi(c.st.sngrad, c.st.sngrad_limit), c.st.heat_conduction, m.d_wideColor.p, col, comp,
                       c.avGrad.p, Rv, Jv, patch, dir ? dir[0] : 0.0, dir ? dir[1] : 0.0, dir ? dir[2] : 0.0, rq, pRef,
                       objd, objv);
}

// one forward-mode pass: gradient, then residual (and objective) tangents
// of the seed (colour col, component comp); col < 0: values only
void av_pass(Case& c, int col, int comp, double* Rv, double* Jv, int patch, const double* dir, double rq, double* objd) {
    const Mesh& m = *c.mesh;
    hipLaunchKernelGGL(k_av_grad, dim3(blocks_for(m.nC, 256)), dim3(256), 0, c.stream, c.meshDev(), c.f, c.d_bc.p, c.th,
                       bounds_of(c), m.d_wideColor.p, col, comp, c.avGrad.p);
    switch (c.st.flux_scheme) {
        case RDST_FLUX_AUSMPLUS: launch_av_resid<RDST_FLUX_AUSMPLUS>(c, col, comp, Rv, Jv, patch, dir, rq, objd); break;
        case RDST_FLUX_AUSMPLUSUP: launch_av_resid<RDST_FLUX_AUSMPLUSUP>(c, col, comp, Rv, Jv, patch, dir, rq, objd); break;
        case RDST_FLUX_HLL: launch_av_resid<RDST_FLUX_HLL>(c, col, comp, Rv, Jv, patch, dir, rq, objd); break;
        case RDST_FLUX_HLLC: launch_av_resid<RDST_FLUX_HLLC>(c, col, comp, Rv, Jv, patch, dir, rq, objd); break;
        default: launch_av_resid<RDST_FLUX_KURGANOV>(c, col, comp, Rv, Jv, patch, dir, rq, objd); break;
    }
    RDST_HIP(hipGetLastError());
}

void av_buffers(Case& c) {
    const size_t n = (size_t)c.mesh->nC;
    if (c.avGrad.n != 18 * n) c.avGrad.alloc(18 * n);
    if (c.avR.n != 5 * n) c.avR.alloc(5 * n);
    if (c.avJv.n != 5 * n) c.avJv.alloc(5 * n);
    if (c.avObj.n != n) c.avObj.alloc(n);
}

}  // namespace

// the second-neighbour pattern, its transpose map, the block-CSR slots in
// it, and the distance-4 colouring (greedy in cell order: cell j takes the
// smallest colour not held by a coloured cell of a row of its own row)
void build_wide(Mesh& m) {
    if (m.wideBuilt) return;
    const int n = m.nC;
    std::vector<int32_t> stamp(n, -1), row;
    m.wideRowPtr.assign(n + 1, 0);
    m.wideCols.clear();
    for (int r = 0; r < n; ++r) {
        row.clear();
        for (int s = m.rowPtr[r]; s < m.rowPtr[r + 1]; ++s) {
            const int a = m.cols[s];
            for (int t = m.rowPtr[a]; t < m.rowPtr[a + 1]; ++t) {
                const int b = m.cols[t];
                if (stamp[b] != r) { stamp[b] = r; row.push_back(b); }
            }
        }
        std::sort(row.begin(), row.end());
        m.wideCols.insert(m.wideCols.end(), row.begin(), row.end());
        m.wideRowPtr[r + 1] = (int32_t)m.wideCols.size();
    }
    auto find = [&](int r, int c) -> int32_t {
        const int32_t* b = m.wideCols.data() + m.wideRowPtr[r];
        const int32_t* e = m.wideCols.data() + m.wideRowPtr[r + 1];
        const int32_t* it = std::lower_bound(b, e, c);
        if (it == e || *it != c) throw std::runtime_error("second-neighbour pattern: missing block");
        return (int32_t)(it - m.wideCols.data());
    };
    const size_t nnz = m.wideCols.size();
    m.wideT.assign(nnz, 0);
    for (int r = 0; r < n; ++r)
        for (int e = m.wideRowPtr[r]; e < m.wideRowPtr[r + 1]; ++e) m.wideT[e] = find(m.wideCols[e], r);
    m.wideOfSlot.assign(m.nnzb, 0);
    for (int r = 0; r < n; ++r)
        for (int s = m.rowPtr[r]; s < m.rowPtr[r + 1]; ++s) m.wideOfSlot[s] = find(r, m.cols[s]);
    m.wideColor.assign(n, -1);
    std::vector<int32_t> forb;
    int ncol = 0;
    for (int j = 0; j < n; ++j) {
        for (int e = m.wideRowPtr[j]; e < m.wideRowPtr[j + 1]; ++e) {
            const int k = m.wideCols[e];
            for (int f = m.wideRowPtr[k]; f < m.wideRowPtr[k + 1]; ++f) {
                const int cl = m.wideColor[m.wideCols[f]];
                if (cl >= 0) forb[cl] = j;
            }
        }
        int cl = 0;
        while (cl < ncol && forb[cl] == j) ++cl;
        if (cl == ncol) { forb.push_back(-1); ++ncol; }
        m.wideColor[j] = cl;
    }
    m.wideColors = ncol;
    m.d_wideRowPtr.upload(m.wideRowPtr);
    m.d_wideCols.upload(m.wideCols);
    m.d_wideT.upload(m.wideT);
    m.d_wideOfSlot.upload(m.wideOfSlot);
    m.d_wideColor.upload(m.wideColor);
    m.wideBuilt = true;
}

namespace {
// J of the viscous residual on the second-neighbour pattern (adjW), J^T
// (adjWT) and the block-CSR part of J^T (adjJT, for the preconditioner)
void viscous_linearise(Case& c) {
    Mesh& m = *c.mesh;
    build_wide(m);
    av_buffers(c);
    const size_t nnz = m.wideCols.size();
    if (c.adjW.n != 25 * nnz) c.adjW.alloc(25 * nnz);
    if (c.adjWT.n != 25 * nnz) c.adjWT.alloc(25 * nnz);
    const int nb = blocks_for(m.nC, 256);
    for (int col = 0; col < m.wideColors; ++col)
        for (int comp = 0; comp < 5; ++comp) {
            av_pass(c, col, comp, nullptr, c.avJv.p, -1, nullptr, 0.0, nullptr);
            hipLaunchKernelGGL(k_av_scatter, dim3(nb), dim3(256), 0, c.stream, m.nC, m.d_wideRowPtr.p, m.d_wideCols.p,
                               m.d_wideColor.p, col, comp, c.avJv.p, c.adjW.p);
        }
    const int gb = std::min(4096, blocks_for((long)nnz, 256));
    hipLaunchKernelGGL(k_wide_transpose, dim3(gb), dim3(256), 0, c.stream, (long)nnz, m.d_wideT.p, c.adjW.p, c.adjWT.p);
    const size_t nn = 25 * (size_t)m.nnzb;
    if (c.adjJT.n != nn) c.adjJT.alloc(nn);
    hipLaunchKernelGGL(k_wide_to_slots, dim3(std::min(4096, blocks_for((long)m.nnzb, 256))), dim3(256), 0, c.stream,
                       (long)m.nnzb, m.d_wideOfSlot.p, c.adjWT.p, c.adjJT.p);
    RDST_HIP(hipGetLastError());
    RDST_HIP(hipStreamSynchronize(c.stream));
}
}  // namespace

void adjoint_check_scheme(const Case& c) {
    if (c.st.flux_scheme == RDST_FLUX_HLLCP)
        throw ArgError("adjoint: HLLCP's pressure sensor couples second neighbours; use AUSM+, AUSM+up, HLL, HLLC or "
                       "Kurganov", RDST_ERR_UNKNOWN_NAME);
    if (c.st.reconstruction != RDST_RECON_UPWIND)
        throw ArgError("adjoint: the exact linearisation covers first-order (upwind) reconstruction");
    if (c.sa.on && !c.st.viscous)
        throw ArgError("adjoint: a turbulence model needs the viscous residual (settings.viscous = 1)");
    if (c.st.viscous && grad_u_limited(c.st))
        throw ArgError("adjoint: cellLimited grad(U) (its min/max limiter) is not linearised; use Gauss linear");
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
    if (c.st.viscous) {
        adjoint_check_scheme(c);
        viscous_linearise(c);
        c.adjWide = true;
        c.hasAdjJ = true;
        return;
    }
    c.adjWide = false;
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
    if (c.st.viscous) {
        build_wide(*c.mesh);
        av_buffers(c);
        av_pass(c, -1, 0, R, nullptr, -1, nullptr, 0.0, nullptr);
        return;
    }
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
    if (c.st.viscous) {
        // the viscous force reads the corrected boundary gradient, which
        // depends on the wall cell's face neighbours: one pass per colour
        Mesh& m = *c.mesh;
        build_wide(m);
        av_buffers(c);
        for (int col = 0; col < m.wideColors; ++col)
            for (int comp = 0; comp < 5; ++comp) {
                av_pass(c, col, comp, nullptr, nullptr, patch, dir, 1.0 / q, c.avObj.p);
                hipLaunchKernelGGL(k_av_obj_gather, dim3(blocks_for(m.nC, 256)), dim3(256), 0, c.stream, c.meshDev(),
                                   m.d_wideColor.p, col, comp, c.avObj.p, g);
            }
        RDST_HIP(hipGetLastError());
        return;
    }
    const Mesh& m = *c.mesh;
    hipLaunchKernelGGL(k_adj_objective, dim3(blocks_for(m.nC, 256)), dim3(256), 0, c.stream, c.meshDev(), c.f, c.d_bc.p,
                       c.th, bounds_of(c), patch, dir[0], dir[1], dir[2], 1.0 / q, g);
    RDST_HIP(hipGetLastError());
}

// each cell's share of the force objective (values; sum = J): the pressure
// part of its faces on the patch and, viscous, their viscous force
void adjoint_objective_cells(Case& c, int patch, const double* dir, double q, double pRef, double* o) {
    adjoint_check_scheme(c);
    Mesh& m = *c.mesh;
    if (c.st.viscous) {
        build_wide(m);
        av_buffers(c);
        hipLaunchKernelGGL(k_av_grad, dim3(blocks_for(m.nC, 256)), dim3(256), 0, c.stream, c.meshDev(), c.f, c.d_bc.p,
                           c.th, bounds_of(c), m.d_wideColor.p, -1, 0, c.avGrad.p);
        switch (c.st.flux_scheme) {
            case RDST_FLUX_AUSMPLUS:
                launch_av_resid<RDST_FLUX_AUSMPLUS>(c, -1, 0, nullptr, nullptr, patch, dir, 1.0 / q, nullptr, pRef, o);
                break;
            case RDST_FLUX_AUSMPLUSUP:
                launch_av_resid<RDST_FLUX_AUSMPLUSUP>(c, -1, 0, nullptr, nullptr, patch, dir, 1.0 / q, nullptr, pRef, o);
                break;
            case RDST_FLUX_HLL:
                launch_av_resid<RDST_FLUX_HLL>(c, -1, 0, nullptr, nullptr, patch, dir, 1.0 / q, nullptr, pRef, o);
                break;
            case RDST_FLUX_HLLC:
                launch_av_resid<RDST_FLUX_HLLC>(c, -1, 0, nullptr, nullptr, patch, dir, 1.0 / q, nullptr, pRef, o);
                break;
            default:
                launch_av_resid<RDST_FLUX_KURGANOV>(c, -1, 0, nullptr, nullptr, patch, dir, 1.0 / q, nullptr, pRef, o);
        }
    } else {
        hipLaunchKernelGGL(k_adj_objective_cells, dim3(blocks_for(m.nC, 256)), dim3(256), 0, c.stream, c.meshDev(), c.f,
                           c.d_bc.p, c.th, bounds_of(c), patch, dir[0], dir[1], dir[2], 1.0 / q, pRef, o);
    }
    RDST_HIP(hipGetLastError());
}

}  // namespace rdst
