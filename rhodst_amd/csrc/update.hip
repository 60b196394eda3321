// update.hip -- state initialisation, primitive recovery with the boundary
// conditions, local time stepping and the CFL ramp.
//
// Replaces createFields.H:21-72 + timeDST::initialize (timeDST.C:67-208),
// #include "updateFields.H" (libraries/include/updateFields.H:40-94) with
// the BC updateCoeffs it triggers (farField{Velocity,Temperature,Pressure}DST,
// inletOutlet, symmetryPlane/slip, fixedValue, zeroGradient and the enthalpy
// BCs hPsiThermo derives from T's), and timeDST::setCourantDST /
// setSteadyStateDeltaT (timeDST.C:218-273, :316-443).
//
// The update runs as three kernels so that every boundary condition sees
// the field values the reference's evaluation order gives it: (1) U and the
// T-derived enthalpy coefficients, from the new rho/U and the not-yet-
// corrected p, T, psi; (2) the cell update; (3) h_b, T_b, psi_b, p_b (the
// far-field pressure reads the corrected cell p and psi), rho_b, rhoU_b, rhoE_b.
#include <cmath>
#include <vector>

#include "rdst_internal.hpp"
#include "fluxfun.hpp"
#include "reduce.hpp"

// No floating-point contraction: the reference evaluates every product
// rounded, so a*x - a*x cancels exactly (a tangential state on both sides of
// a boundary face gives phi = 0).  Fused into fma(a, x, -(a*x)) it leaves the
// rounding error of a*x instead, and the sign tests that read phi (far-field
// regimes mach <= -1, < 0, >= 0, >= 1; inletOutlet; upwind weights) flip on it.
#pragma clang fp contract(off)

namespace rdst {

namespace {

__device__ inline double dpos(double s) { return s > 0.0 ? 1.0 : 0.0; }
__device__ inline double dpos0(double s) { return s >= 0.0 ? 1.0 : 0.0; }
__device__ inline double dneg(double s) { return s < 0.0 ? 1.0 : 0.0; }
__device__ inline double dneg0(double s) { return s <= 0.0 ? 1.0 : 0.0; }
__device__ inline double fepos(double s) { return s >= 0.0 ? 1.0 : 0.0; }  // foam-extend pos()
__device__ inline double clampd(double v, double lo, double hi) { return fmin(fmax(v, lo), hi); }

__device__ inline double mu_of(const Thermo& th, double T) {
    return th.transport == 1 ? th.As * sqrt(T) / (1.0 + th.Ts / T) : th.mu;
}
__device__ inline double alpha_of(const Thermo& th, double T) {
    if (th.transport == 1) {
        double mu = mu_of(th, T);
        return mu * th.Cv * (1.32 + 1.77 * th.R / th.Cv) / th.Cp;
    }
    return th.mu / th.Pr;
}

// symmetryPlane/slip vector value: (transform(I - 2nn, U_I) + U_I)/2
__device__ inline void sym_value(double nx, double ny, double nz, double ux, double uy, double uz, double& ox,
                                 double& oy, double& oz) {
    double nU = nx * ux + ny * uy + nz * uz;
    double tx = ux - 2.0 * nx * nU, ty = uy - 2.0 * ny * nU, tz = uz - 2.0 * nz * nU;
    ox = (tx + ux) / 2.0;
    oy = (ty + uy) / 2.0;
    oz = (tz + uz) / 2.0;
}

// ---------------------------------------------------------------- set state
__global__ void k_init_cells(int n, Fields F, Thermo th, const double* p, const double* T, const double* U) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double pi = p[i], Ti = T[i], ux = U[3 * i], uy = U[3 * i + 1], uz = U[3 * i + 2];
    F.p[i] = pi; F.T[i] = Ti; F.Ux[i] = ux; F.Uy[i] = uy; F.Uz[i] = uz;
    double psi = 1.0 / (th.R * Ti), h = th.Cp * Ti;
    F.psi[i] = psi; F.h[i] = h;
    F.mu[i] = mu_of(th, Ti); F.alpha[i] = alpha_of(th, Ti);
    double rho = psi * pi;
    F.rho[i] = rho;
    F.rhoUx[i] = rho * ux; F.rhoUy[i] = rho * uy; F.rhoUz[i] = rho * uz;
    double rhoE = rho * (h * th.Cv / th.Cp + 0.5 * (ux * ux + uy * uy + uz * uz));
    F.rhoE[i] = rhoE;
    F.W[5 * i] = rho; F.W[5 * i + 1] = rho * ux; F.W[5 * i + 2] = rho * uy; F.W[5 * i + 3] = rho * uz;
    F.W[5 * i + 4] = rhoE;
    store_cell8(F, i, rho, rho * ux, rho * uy, rho * uz, rhoE, pi, h, psi, th.gamma);
    F.rhoOld[i] = rho;                       // thermo rho
    F.eOld[i] = h * th.Cv / th.Cp;
}

// initial boundary values: the BC "value" entries (fixedValue/inletOutlet
// value, far-field free stream) or the patch-internal value; valueFraction 0
__global__ void k_init_bnd(MeshDev m, Fields F, Thermo th, const BcPatchDev* bc) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.nB) return;
    const BcPatchDev& B = bc[m.bpatch[b]];
    int ci = m.bcell[b];
    const int k = m.nI + b;
    double ms = m.magSf[k];
    double nx = m.Sf[k] / ms, ny = m.Sf[m.nF + k] / ms, nz = m.Sf[2 * m.nF + k] / ms;
    double pb, Tb, ux, uy, uz;
    switch (B.p_type) {
        case RDST_BC_FIXED_VALUE: case RDST_BC_INLET_OUTLET: pb = B.p_value; break;
        case RDST_BC_FARFIELD: pb = B.p_inf; break;
        default: pb = F.p[ci];
    }
    switch (B.T_type) {
        case RDST_BC_FIXED_VALUE: case RDST_BC_INLET_OUTLET: Tb = B.T_value; break;
        case RDST_BC_FARFIELD: Tb = B.T_inf; break;
        default: Tb = F.T[ci];
    }
    switch (B.U_type) {
        case RDST_BC_FIXED_VALUE: case RDST_BC_INLET_OUTLET: ux = B.U_value[0]; uy = B.U_value[1]; uz = B.U_value[2]; break;
        case RDST_BC_FARFIELD: ux = B.U_inf[0]; uy = B.U_inf[1]; uz = B.U_inf[2]; break;
        case RDST_BC_SYMMETRY: sym_value(nx, ny, nz, F.Ux[ci], F.Uy[ci], F.Uz[ci], ux, uy, uz); break;
        default: ux = F.Ux[ci]; uy = F.Uy[ci]; uz = F.Uz[ci];
    }
    F.bp[b] = pb; F.bT[b] = Tb; F.bUx[b] = ux; F.bUy[b] = uy; F.bUz[b] = uz;
    F.vf_p[b] = 0.0; F.vf_T[b] = 0.0; F.vf_U[b] = 0.0; F.TrefH[b] = 0.0;
    F.refUx[b] = 0.0; F.refUy[b] = 0.0; F.refUz[b] = 0.0;
    double psi = 1.0 / (th.R * Tb), h = th.Cp * Tb;
    F.bpsi[b] = psi; F.bh[b] = h;
    F.bmu[b] = mu_of(th, Tb); F.balpha[b] = alpha_of(th, Tb);
    double rho = psi * pb;
    F.brho[b] = rho;
    F.brhoUx[b] = rho * ux; F.brhoUy[b] = rho * uy; F.brhoUz[b] = rho * uz;
    F.brhoE[b] = rho * (h * th.Cv / th.Cp + 0.5 * (ux * ux + uy * uy + uz * uz));
}

// phi = linearInterpolate(rho*U) & Sf (compressibleCreatePhi)
__global__ void k_init_phi(MeshDev m, Fields F) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m.nF) return;
    double sx = m.Sf[k], sy = m.Sf[m.nF + k], sz = m.Sf[2 * m.nF + k];
    if (k < m.nI) {
        int o = m.owner[k], nb = m.neigh[k];
        double w = m.w[k];
        double ox = F.rho[o] * F.Ux[o], oy = F.rho[o] * F.Uy[o], oz = F.rho[o] * F.Uz[o];
        double nx = F.rho[nb] * F.Ux[nb], ny = F.rho[nb] * F.Uy[nb], nz = F.rho[nb] * F.Uz[nb];
        double fx = w * (ox - nx) + nx, fy = w * (oy - ny) + ny, fz = w * (oz - nz) + nz;
        F.phi[k] = fx * sx + fy * sy + fz * sz;
    } else {
        int b = k - m.nI;
        double r = F.brho[b];
        F.phi[k] = (r * F.bUx[b]) * sx + (r * F.bUy[b]) * sy + (r * F.bUz[b]) * sz;
    }
}

// Blazek dS = 1/2 sum cmptMag(Sf) over internal and non-empty boundary faces
__global__ void k_init_dS(MeshDev m, Fields F) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const size_t E = (size_t)m.ellW * m.nC;
    double ax = 0, ay = 0, az = 0;
    for (int k = 0; k < m.ellW; ++k) {
        size_t e = (size_t)k * m.nC + c;
        if ((m.ellCode[e] >> 30) == ELL_NONE) break;
        ax += fabs(m.ellSf[e]); ay += fabs(m.ellSf[E + e]); az += fabs(m.ellSf[2 * E + e]);
    }
    F.dS[c] = ax * 0.5; F.dS[m.nC + c] = ay * 0.5; F.dS[2 * m.nC + c] = az * 0.5;
}

// ---------------------------------------------------------------- update
// kernel 1: per boundary face, U BC and the T-derived enthalpy coefficients
__global__ void k_update_bnd_pre(MeshDev m, Fields F, Thermo th, const BcPatchDev* bc, double rhoMin, double rhoMax) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.nB) return;
    const BcPatchDev& B = bc[m.bpatch[b]];
    const int ci = m.bcell[b];
    const int k = m.nI + b;
    const double ms = m.magSf[k];
    const double nx = m.Sf[k] / ms, ny = m.Sf[m.nF + k] / ms, nz = m.Sf[2 * m.nF + k] / ms;
    const double phib = F.phi[k];
    // new cell rho (bounded) and U from W; old p, T, psi still in place
    const double rhoI = clampd(F.W[5 * ci], rhoMin, rhoMax);
    const double uIx = F.W[5 * ci + 1] / rhoI, uIy = F.W[5 * ci + 2] / rhoI, uIz = F.W[5 * ci + 3] / rhoI;
    const double pOld = F.p[ci], TOld = F.T[ci], psiOld = F.psi[ci];
    const double gam = th.Cp / th.Cv;
    double ux = F.bUx[b], uy = F.bUy[b], uz = F.bUz[b];
    switch (B.U_type) {
        case RDST_BC_FIXED_VALUE: ux = B.U_value[0]; uy = B.U_value[1]; uz = B.U_value[2]; break;
        case RDST_BC_ZERO_GRADIENT: ux = uIx; uy = uIy; uz = uIz; break;
        case RDST_BC_SYMMETRY: sym_value(nx, ny, nz, uIx, uIy, uIz, ux, uy, uz); break;
        case RDST_BC_INLET_OUTLET: {
            double vf = 1.0 - fepos(phib);
            F.vf_U[b] = vf;
            F.refUx[b] = B.U_value[0]; F.refUy[b] = B.U_value[1]; F.refUz[b] = B.U_value[2];
            ux = vf * B.U_value[0] + (1.0 - vf) * uIx;
            uy = vf * B.U_value[1] + (1.0 - vf) * uIy;
            uz = vf * B.U_value[2] + (1.0 - vf) * uIz;
            break;
        }
        case RDST_BC_FARFIELD: {
            // farFieldVelocityDST::updateCoeffs (U.C:144-217)
            double cb = sqrt(gam / F.bpsi[b]);
            double cI = sqrt(gam / psiOld);
            double mach = phib / F.brho[b] / ms / cb;
            double supIn = dneg0(mach + 1), supOut = dpos0(mach - 1);
            double subIn = dneg(mach) - supIn, subOut = dpos0(mach) - supOut;
            double rho0c0 = rhoI * cI;
            double dUn = nx * (B.U_inf[0] - uIx) + ny * (B.U_inf[1] - uIy) + nz * (B.U_inf[2] - uIz);
            double pb = 0.5 * (B.p_inf + pOld - rho0c0 * dUn);
            double a = (B.p_inf - pb) / rho0c0, o = (pOld - pb) / rho0c0;
            double rx = supIn * B.U_inf[0] + subIn * (B.U_inf[0] - nx * a) + subOut * (nx * o);
            double ry = supIn * B.U_inf[1] + subIn * (B.U_inf[1] - ny * a) + subOut * (ny * o);
            double rz = supIn * B.U_inf[2] + subIn * (B.U_inf[2] - nz * a) + subOut * (nz * o);
            double vf = 1 - supOut - subOut;
            F.vf_U[b] = vf;
            F.refUx[b] = rx; F.refUy[b] = ry; F.refUz[b] = rz;
            ux = vf * rx + (1.0 - vf) * uIx;
            uy = vf * ry + (1.0 - vf) * uIy;
            uz = vf * rz + (1.0 - vf) * uIz;
            break;
        }
        default: break;
    }
    F.bUx[b] = ux; F.bUy[b] = uy; F.bUz[b] = uz;
    if (B.T_type == RDST_BC_INLET_OUTLET) {
        F.vf_T[b] = 1.0 - fepos(phib);
        F.TrefH[b] = B.T_value;
    } else if (B.T_type == RDST_BC_FARFIELD) {
        // farFieldTemperatureDST::updateCoeffs (T.C:144-221)
        double cb = sqrt(gam / F.bpsi[b]);
        double cI = sqrt(gam / psiOld);
        double mach = phib / F.brho[b] / ms / cb;
        double supIn = dneg0(mach + 1), supOut = dpos0(mach - 1);
        double subIn = dneg(mach) - supIn, subOut = dpos0(mach) - supOut;
        double rho0c0 = rhoI * cI;
        double dUn = nx * (B.U_inf[0] - uIx) + ny * (B.U_inf[1] - uIy) + nz * (B.U_inf[2] - uIz);
        double pb = 0.5 * (B.p_inf + pOld - rho0c0 * dUn);
        double RI = pOld / rhoI / TOld;
        double rhoa = B.p_inf / B.T_inf / RI;
        double cI2 = cI * cI;
        double subsIn = pb / (rhoa + (pb - B.p_inf) / cI2) / RI;
        double subsOutVF = pb / (pOld + RI * TOld * (pb - pOld) / cI2);
        F.TrefH[b] = supIn * B.T_inf + subIn * subsIn;
        F.vf_T[b] = 1 - supOut - subsOutVF * subOut;
    }
}

// kernel 2: cells -- retrieve from W, bound rho, U, h (bounded), thermo, p (bounded)
// update limit (an extension, before the rest of the update reads W):
// scale each cell's change of W so that rho and rhoE move by at most lim of
// their values before the update
__global__ void k_limit_update(int n, Fields F, double lim) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double o[5] = {F.rho[i], F.rhoUx[i], F.rhoUy[i], F.rhoUz[i], F.rhoE[i]};
    const double dr = fabs(F.W[5 * i] - o[0]), de = fabs(F.W[5 * i + 4] - o[4]);
    double a = 1.0;
    if (dr > lim * o[0]) a = fmin(a, lim * o[0] / dr);
    if (de > lim * o[4]) a = fmin(a, lim * o[4] / de);
    if (a < 1.0)
        for (int q = 0; q < 5; ++q) F.W[5 * i + q] = o[q] + a * (F.W[5 * i + q] - o[q]);
}

__global__ void k_update_cells(int n, Fields F, Thermo th, double rhoMin, double rhoMax, double hMin, double hMax,
                               double pMin, double pMax) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double rho = clampd(F.W[5 * i], rhoMin, rhoMax);
    double rux = F.W[5 * i + 1], ruy = F.W[5 * i + 2], ruz = F.W[5 * i + 3], rhoE = F.W[5 * i + 4];
    F.rho[i] = rho; F.rhoUx[i] = rux; F.rhoUy[i] = ruy; F.rhoUz[i] = ruz; F.rhoE[i] = rhoE;
    double ux = rux / rho, uy = ruy / rho, uz = ruz / rho;
    F.Ux[i] = ux; F.Uy[i] = uy; F.Uz[i] = uz;
    double h = (th.Cp / th.Cv) * (rhoE / rho - 0.5 * (ux * ux + uy * uy + uz * uz));
    h = clampd(h, hMin, hMax);
    F.h[i] = h;
    double T = h / th.Cp;
    double psi = 1.0 / (th.R * T);
    F.T[i] = T; F.psi[i] = psi;
    F.mu[i] = mu_of(th, T); F.alpha[i] = alpha_of(th, T);
    const double p = clampd(rho / psi, pMin, pMax);
    F.p[i] = p;
    store_cell8(F, i, rho, rux, ruy, ruz, rhoE, p, h, psi, th.gamma);
}

// kernel 3: boundary h, T, psi, p and the conserved boundary values
__global__ void k_update_bnd_post(MeshDev m, Fields F, Thermo th, const BcPatchDev* bc, double pMin, double pMax) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.nB) return;
    const BcPatchDev& B = bc[m.bpatch[b]];
    const int ci = m.bcell[b];
    const int k = m.nI + b;
    const double gam = th.Cp / th.Cv;
    // enthalpy BCs derived from T's
    double hb;
    switch (B.T_type) {
        case RDST_BC_FIXED_VALUE: hb = th.Cp * B.T_value; break;
        case RDST_BC_INLET_OUTLET: case RDST_BC_FARFIELD: {
            double vf = F.vf_T[b];
            hb = vf * (th.Cp * F.TrefH[b]) + (1.0 - vf) * F.h[ci];
            break;
        }
        default: hb = F.h[ci];
    }
    // thermo.correct() on the patch
    double Tb;
    if (B.T_type == RDST_BC_FIXED_VALUE) { Tb = B.T_value; hb = th.Cp * Tb; }
    else Tb = hb / th.Cp;
    double psib = 1.0 / (th.R * Tb);
    F.bh[b] = hb; F.bT[b] = Tb; F.bpsi[b] = psib;
    F.bmu[b] = mu_of(th, Tb); F.balpha[b] = alpha_of(th, Tb);
    // p.correctBoundaryConditions()
    const double phib = F.phi[k];
    const double pI = F.p[ci];
    double pb;
    switch (B.p_type) {
        case RDST_BC_FIXED_VALUE: pb = B.p_value; break;
        case RDST_BC_INLET_OUTLET: {
            double vf = 1.0 - fepos(phib);
            F.vf_p[b] = vf;
            pb = vf * B.p_value + (1.0 - vf) * pI;
            break;
        }
        case RDST_BC_FARFIELD: {
            // farFieldPressureDST::updateCoeffs (P.C:144-215)
            const double ms = m.magSf[k];
            const double nx = m.Sf[k] / ms, ny = m.Sf[m.nF + k] / ms, nz = m.Sf[2 * m.nF + k] / ms;
            double cb = sqrt(gam / psib);
            double cI = sqrt(gam / F.psi[ci]);
            double mach = phib / F.brho[b] / ms / cb;
            double supIn = dneg0(mach + 1), supOut = dpos0(mach - 1);
            double subIn = dneg(mach) - supIn, subOut = dpos0(mach) - supOut;
            double rho0c0 = F.rho[ci] * cI;
            double dUn = nx * (B.U_inf[0] - F.Ux[ci]) + ny * (B.U_inf[1] - F.Uy[ci]) + nz * (B.U_inf[2] - F.Uz[ci]);
            double pbf = B.p_inf - rho0c0 * dUn;
            double ref = supIn * B.p_inf + subIn * pbf + subOut * B.p_inf;
            double vf = 1 - supOut - 0.5 * subIn;
            F.vf_p[b] = vf;
            pb = vf * ref + (1.0 - vf) * pI;
            break;
        }
        default: pb = pI;
    }
    pb = clampd(pb, pMin, pMax);
    F.bp[b] = pb;
    double rb = pb * psib;
    double ux = F.bUx[b], uy = F.bUy[b], uz = F.bUz[b];
    F.brho[b] = rb;
    F.brhoUx[b] = rb * ux; F.brhoUy[b] = rb * uy; F.brhoUz[b] = rb * uz;
    F.brhoE[b] = rb * (hb + 0.5 * (ux * ux + uy * uy + uz * uz)) - pb;
}

// ---------------------------------------------------------------- LTS
// setSteadyStateDeltaT, neighbouring = true: max over the cell's faces of
// (|interp U| + interp sqrt(gamma/psi)) deltaCoeffs, wall faces add
// 0.5 deltaCoeffs c + |U| (timeDST.C:231-269), divided by courant
__global__ void k_lts_neighbouring(MeshDev m, Fields F, Thermo th) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const double gam = th.Cp / th.Cv;
    const double ux = F.Ux[c], uy = F.Uy[c], uz = F.Uz[c];
    const double cc = sqrt(gam / F.psi[c]);
    double dt = 0.0;
    for (int k = 0; k < m.ellW; ++k) {
        size_t e = (size_t)k * m.nC + c;
        uint32_t code = m.ellCode[e];
        uint32_t kind = code >> 30;
        if (kind == ELL_NONE) break;
        int idx = (int)(code & ELL_IDX_MASK);
        if (kind == ELL_BOUNDARY) {
            if (m.bwall[idx]) {
                double pL = 0.5 * m.ellDc[e] * cc + sqrt(ux * ux + uy * uy + uz * uz);
                dt = fmax(dt, pL);
            }
            continue;
        }
        double w = m.ellWt[e];
        double oUx, oUy, oUz, oc, nUx, nUy, nUz, nc;
        double ox = F.Ux[idx], oy = F.Uy[idx], oz = F.Uz[idx], occ = sqrt(gam / F.psi[idx]);
        if (kind == ELL_OWNER) { oUx = ux; oUy = uy; oUz = uz; oc = cc; nUx = ox; nUy = oy; nUz = oz; nc = occ; }
        else { oUx = ox; oUy = oy; oUz = oz; oc = occ; nUx = ux; nUy = uy; nUz = uz; nc = cc; }
        double fx = w * (oUx - nUx) + nUx, fy = w * (oUy - nUy) + nUy, fz = w * (oUz - nUz) + nUz;
        double fc = w * (oc - nc) + nc;
        double dtSf = (sqrt(fx * fx + fy * fy + fz * fz) + fc) * m.ellDc[e];
        dt = fmax(dt, dtSf);
    }
    F.rDeltaT[c] = dt / F.courant[c];
}

// neighbouring = false (Blazek): ((|u|+c)dSx + (|v|+c)dSy + (|w|+c)dSz)/(courant V)
__global__ void k_lts_blazek(MeshDev m, Fields F, Thermo th, double courant) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    double cc = sqrt(th.Cp / th.Cv / F.psi[c]);
    double lam = (fabs(F.Ux[c]) + cc) * F.dS[c] + (fabs(F.Uy[c]) + cc) * F.dS[m.nC + c] + (fabs(F.Uz[c]) + cc) * F.dS[2 * m.nC + c];
    F.rDeltaT[c] = lam / (courant * m.V[c]);
}

// setTransientDeltaT (timeDST.C:275-314): per-block minima of the physical
// time step, over the internal faces |Sf| / ((|U_f . Sf| + |Sf| c_f)
// deltaCoeffs) courant with linear interpolation (neighbouring), or over the
// cells courant V / ((|U| + c) . dS)
__global__ __launch_bounds__(256) void k_dt_faces(MeshDev m, Fields F, Thermo th, double courant, double* part) {
    __shared__ double sh[16];
    const double gam = th.Cp / th.Cv;
    double dt = 1e300;
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < m.nI; f += gridDim.x * blockDim.x) {
        const int o = m.owner[f], nb = m.neigh[f];
        const double w = m.w[f];
        const double fx = w * (F.Ux[o] - F.Ux[nb]) + F.Ux[nb];
        const double fy = w * (F.Uy[o] - F.Uy[nb]) + F.Uy[nb];
        const double fz = w * (F.Uz[o] - F.Uz[nb]) + F.Uz[nb];
        const double co = sqrt(gam / F.psi[o]), cn = sqrt(gam / F.psi[nb]);
        const double fc = w * (co - cn) + cn;
        const double sx = m.Sf[f], sy = m.Sf[m.nF + f], sz = m.Sf[2 * (size_t)m.nF + f];
        const double amax = fabs(fx * sx + fy * sy + fz * sz) + m.magSf[f] * fc;
        const double dtSf = m.magSf[f] / (amax * m.dc[f]);
        dt = fmin(dt, dtSf * courant);
    }
    const double t = block_min(dt, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}
__global__ __launch_bounds__(256) void k_dt_cells(MeshDev m, Fields F, Thermo th, double courant, double* part) {
    __shared__ double sh[16];
    double dt = 1e300;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < m.nC; c += gridDim.x * blockDim.x) {
        const double cc = sqrt(th.Cp / th.Cv / F.psi[c]);
        const double lam = (fabs(F.Ux[c]) + cc) * F.dS[c] + (fabs(F.Uy[c]) + cc) * F.dS[m.nC + c] +
                           (fabs(F.Uz[c]) + cc) * F.dS[2 * m.nC + c];
        dt = fmin(dt, courant * m.V[c] / lam);
    }
    const double t = block_min(dt, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// the conserved fields as [nC][5] (rho, rhoU, rhoE): an old time level of W
__global__ void k_conserved_aos(int n, Fields F, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[5 * (size_t)i] = F.rho[i];
    out[5 * (size_t)i + 1] = F.rhoUx[i];
    out[5 * (size_t)i + 2] = F.rhoUy[i];
    out[5 * (size_t)i + 3] = F.rhoUz[i];
    out[5 * (size_t)i + 4] = F.rhoE[i];
}

// ---------------------------------------------------------------- CFL ramp
// per-block partial sums of |R_c|^2 (mag then square, timeDST.C:328-331)
__global__ __launch_bounds__(256) void k_rnorm_partial(int n, const double* R, double* partial) {
    __shared__ double sh[256 / 64];
    double s = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double a = R[i], b = R[n + i], c = R[2 * n + i], d = R[3 * n + i], e = R[4 * n + i];
        double mg = sqrt(a * a + b * b + c * c + d * d + e * e);
        s += mg * mg;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) sh[wv] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += sh[k];
        partial[blockIdx.x] = t;
    }
}

// factor 0.5 where rho or e dropped, 0.75 next to such a cell; courant ramp and clamp
__global__ void k_courant(MeshDev m, Fields F, double ratio, double minCo, double maxCo) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    auto dropped = [&](int i) {
        double r = F.rho[i];
        double ux = F.rhoUx[i] / r, uy = F.rhoUy[i] / r, uz = F.rhoUz[i] / r;
        double eN = F.rhoE[i] / r - 0.5 * (ux * ux + uy * uy + uz * uz);
        return (r < 0.95 * F.rhoOld[i]) || (eN < 0.95 * F.eOld[i]) || (eN < 1.0e-15);
    };
    double factor = 1.0;
    bool any = false;
    for (int k = 0; k < m.ellW; ++k) {
        uint32_t code = m.ellCode[(size_t)k * m.nC + c];
        uint32_t kind = code >> 30;
        if (kind == ELL_NONE) break;
        if (kind == ELL_BOUNDARY) continue;
        any = true;
        if (dropped((int)(code & ELL_IDX_MASK))) factor = fmin(factor, 0.75);
    }
    if (any && dropped(c)) factor = fmin(factor, 0.5);
    double co = F.courant[c] * (ratio * factor);
    F.courant[c] = fmin(fmax(co, minCo), maxCo);
}

// rhoOld = rho, eOld = eNew (after every cell has read the old values)
__global__ void k_store_old(int n, Fields F) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double r = F.rho[i];
    double ux = F.rhoUx[i] / r, uy = F.rhoUy[i] / r, uz = F.rhoUz[i] / r;
    F.rhoOld[i] = r;
    F.eOld[i] = F.rhoE[i] / r - 0.5 * (ux * ux + uy * uy + uz * uz);
}

}  // namespace

void set_steady_state_deltaT(Case& c) {
    MeshDev md = c.meshDev();
    const int bs = 256;
    if (c.st.neighbouring)
        hipLaunchKernelGGL(k_lts_neighbouring, dim3(blocks_for(md.nC, bs)), dim3(bs), 0, c.stream, md, c.f, c.th);
    else
        hipLaunchKernelGGL(k_lts_blazek, dim3(blocks_for(md.nC, bs)), dim3(bs), 0, c.stream, md, c.f, c.th, c.courantScalar);
    RDST_HIP(hipGetLastError());
}

void set_state(Case& c, const double* p, const double* T, const double* U) {
    Mesh& m = *c.mesh;
    const int n = m.nC;
    DBuf<double> dp, dT, dU;
    dp.upload(p, n);
    dT.upload(T, n);
    dU.upload(U, 3 * (size_t)n);
    MeshDev md = c.meshDev();
    const int bs = 256;
    hipLaunchKernelGGL(k_init_cells, dim3(blocks_for(n, bs)), dim3(bs), 0, c.stream, n, c.f, c.th, dp.p, dT.p, dU.p);
    c.packDirty = false;
    if (m.nB) hipLaunchKernelGGL(k_init_bnd, dim3(blocks_for(m.nB, bs)), dim3(bs), 0, c.stream, md, c.f, c.th, c.d_bc.p);
    hipLaunchKernelGGL(k_init_phi, dim3(blocks_for(m.nF, bs)), dim3(bs), 0, c.stream, md, c.f);
    RDST_HIP(hipMemsetAsync(c.f.R, 0, sizeof(double) * 5 * (size_t)n, c.stream));
    if (!c.st.neighbouring) hipLaunchKernelGGL(k_init_dS, dim3(blocks_for(n, bs)), dim3(bs), 0, c.stream, md, c.f);
    c.residual = -1;
    c.residualRatio = 1;
    c.courantScalar = c.st.initial_co;
    std::vector<double> co(n, c.st.initial_co);
    RDST_HIP(hipMemcpyAsync(c.f.courant, co.data(), sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    c.time = 0;
    c.timeIndex = 0;
    if (c.st.time_integration == RDST_TIME_DUAL) {
        // timeDST::initialize (timeDST.C:193-207): the physical deltaT from the
        // initial courant; deltaT0 of the first step is the controlDict deltaT
        set_transient_deltaT(c);
        c.deltaT0 = c.deltaTSave = c.st.delta_t > 0 ? c.st.delta_t : c.deltaT;
    }
    set_steady_state_deltaT(c);
    RDST_HIP(hipGetLastError());
    RDST_HIP(hipStreamSynchronize(c.stream));
    c.hasState = true;
    c.hasMatrix = false;
}

void update_fields(Case& c) {
    Mesh& m = *c.mesh;
    MeshDev md = c.meshDev();
    const int bs = 256;
    const rdst_settings& s = c.st;
    if (c.sa.on) {
        // rho.oldTime() of the turbulence ddt: the density before this update
        RDST_HIP(hipMemcpyAsync(c.sa.rhoPrev, c.f.rho, sizeof(double) * m.nC, hipMemcpyDeviceToDevice, c.stream));
        c.sa.rhoPrevValid = true;
    }
    if (s.update_limit > 0.0)
        hipLaunchKernelGGL(k_limit_update, dim3(blocks_for(m.nC, bs)), dim3(bs), 0, c.stream, m.nC, c.f, s.update_limit);
    if (m.nB)
        hipLaunchKernelGGL(k_update_bnd_pre, dim3(blocks_for(m.nB, bs)), dim3(bs), 0, c.stream, md, c.f, c.th, c.d_bc.p,
                           s.rho_min, s.rho_max);
    hipLaunchKernelGGL(k_update_cells, dim3(blocks_for(m.nC, bs)), dim3(bs), 0, c.stream, m.nC, c.f, c.th, s.rho_min,
                       s.rho_max, c.th.Cp * s.T_min, c.th.Cp * s.T_max, s.p_min, s.p_max);
    c.packDirty = false;
    if (m.nB)
        hipLaunchKernelGGL(k_update_bnd_post, dim3(blocks_for(m.nB, bs)), dim3(bs), 0, c.stream, md, c.f, c.th, c.d_bc.p,
                           s.p_min, s.p_max);
    RDST_HIP(hipGetLastError());
}

void store_conserved(Case& c, double* out) {
    const int n = c.mesh->nC;
    hipLaunchKernelGGL(k_conserved_aos, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, n, c.f, out);
    RDST_HIP(hipGetLastError());
}

void set_transient_deltaT(Case& c) {
    Mesh& m = *c.mesh;
    MeshDev md = c.meshDev();
    const int bs = 256;
    const int nItems = c.st.neighbouring ? m.nI : m.nC;
    const int nb = std::max(1, std::min(1024, blocks_for(nItems, bs)));
    if (c.hostScratch.size() < (size_t)nb) c.hostScratch.resize(nb);
    double* dpart = c.f.dS + 3 * (size_t)m.nC;  // partial scratch reserved after dS
    if (c.st.neighbouring)
        hipLaunchKernelGGL(k_dt_faces, dim3(nb), dim3(bs), 0, c.stream, md, c.f, c.th, c.courantScalar, dpart);
    else
        hipLaunchKernelGGL(k_dt_cells, dim3(nb), dim3(bs), 0, c.stream, md, c.f, c.th, c.courantScalar, dpart);
    RDST_HIP(hipGetLastError());
    RDST_HIP(hipMemcpyAsync(c.hostScratch.data(), dpart, sizeof(double) * nb, hipMemcpyDeviceToHost, c.stream));
    RDST_HIP(hipStreamSynchronize(c.stream));
    double dt = 1e300;
    for (int i = 0; i < nb; ++i) dt = std::min(dt, c.hostScratch[i]);
    c.deltaT = dt;
}

double set_courant(Case& c) { return set_courant_dst(c, nullptr, true); }

// setCourantDST (timeDST.C:316-443)
double set_courant_dst(Case& c, const double* lin, bool pseudo) {
    Mesh& m = *c.mesh;
    MeshDev md = c.meshDev();
    const int bs = 256;
    const int nb = std::min(1024, blocks_for(m.nC, bs));
    if (c.hostScratch.size() < (size_t)nb) c.hostScratch.resize(nb);
    DBuf<double>& part = c.timeBuf;  // scratch after the time fields
    (void)part;
    double* dpart = c.f.dS + 3 * (size_t)m.nC;  // partial-sum scratch reserved after dS
    double residualR;
    if (lin) {
        // |solverPerf.initialResidual()| (timeDST.C:321-325)
        double s2 = 0.0;
        for (int q = 0; q < 5; ++q) s2 += lin[q] * lin[q];
        residualR = std::sqrt(s2);
    } else {
        hipLaunchKernelGGL(k_rnorm_partial, dim3(nb), dim3(bs), 0, c.stream, m.nC, c.f.R, dpart);
        RDST_HIP(hipGetLastError());
        RDST_HIP(hipMemcpyAsync(c.hostScratch.data(), dpart, sizeof(double) * nb, hipMemcpyDeviceToHost, c.stream));
        RDST_HIP(hipStreamSynchronize(c.stream));
        double ss = 0.0;
        for (int i = 0; i < nb; ++i) ss += c.hostScratch[i];
        residualR = std::sqrt(ss);
    }
    if (c.residual != -1) c.residualRatio = std::min(std::max(c.residual / residualR, 0.1), 2.0);
    c.residual = residualR;
    if (!pseudo) {
        // the physical branch: with adjustTimeStep the scalar courant follows
        // the residual ratio and sets the next physical deltaT
        if (c.st.adjust_time_step) {
            c.courantScalar = std::min(std::max(c.courantScalar * c.residualRatio, c.st.min_co), c.st.max_co);
            set_transient_deltaT(c);
        }
        return residualR;
    }
    hipLaunchKernelGGL(k_courant, dim3(blocks_for(m.nC, bs)), dim3(bs), 0, c.stream, md, c.f, c.residualRatio,
                       c.st.min_co, c.st.max_co);
    hipLaunchKernelGGL(k_store_old, dim3(blocks_for(m.nC, bs)), dim3(bs), 0, c.stream, m.nC, c.f);
    set_steady_state_deltaT(c);
    RDST_HIP(hipGetLastError());
    return residualR;
}

}  // namespace rdst
