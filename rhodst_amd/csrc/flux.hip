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
#include <mutex>

#include "rdst_internal.hpp"
#include "reduce.hpp"
#include "fluxfun.hpp"

// No floating-point contraction: the reference evaluates every product
// rounded, so a*x - a*x cancels exactly (a tangential state on both sides of
// a boundary face gives phi = 0).  Fused into fma(a, x, -(a*x)) it leaves the
// rounding error of a*x instead, and the sign tests that read phi (far-field
// regimes mach <= -1, < 0, >= 0, >= 1; inletOutlet; upwind weights) flip on it.
#pragma clang fp contract(off)

namespace rdst {

namespace {

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

// cellLimitedGrad's limitFace for one component: e is the change from the
// cell value to a face centre along the cell's gradient, [dmin, dmax] the
// admissible change; 1e-300 is VSMALL
__device__ inline void limit_face(double& l, double dmax, double dmin, double e) {
    if (e > dmax + 1.0e-300) l = fmin(l, dmax / e);
    else if (e < dmin - 1.0e-300) l = fmin(l, dmin / e);
}

// K2: fvc::grad(U), Gauss linear, one thread per cell.  LIM (cellLimited
// Gauss linear k): the Gauss linear gradient is kept in gradUc and, in the
// same pass over the cell's faces, the min/max of U over the face neighbours
// and boundary values is formed; a second pass over the faces gives the
// per-component limiter from the face centres, and gradU receives the
// limited gradient (column j scaled by l_j).
template <bool LIM>
__global__ __launch_bounds__(256) void k_grad_u(MeshDev m, Fields F, double k) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const size_t E = (size_t)m.ellW * m.nC;
    const double ux = F.Ux[c], uy = F.Uy[c], uz = F.Uz[c];
    double g[9];
    for (int q = 0; q < 9; ++q) g[q] = 0.0;
    double mx[3] = {ux, uy, uz}, mn[3] = {ux, uy, uz};
    for (int kk = 0; kk < m.ellW; ++kk) {
        const size_t e = (size_t)kk * m.nC + c;
        const uint32_t code = m.ellCode[e];
        const uint32_t kind = code >> 30;
        if (kind == ELL_NONE) break;
        const int idx = (int)(code & ELL_IDX_MASK);
        const double S[3] = {m.ellSf[e], m.ellSf[E + e], m.ellSf[2 * E + e]};
        double f[3];
        double sgn = 1.0;
        if (kind == ELL_BOUNDARY) {
            f[0] = F.bUx[idx]; f[1] = F.bUy[idx]; f[2] = F.bUz[idx];
            if (LIM)
                for (int j = 0; j < 3; ++j) { mx[j] = fmax(mx[j], f[j]); mn[j] = fmin(mn[j], f[j]); }
        } else {
            const double w = m.ellWt[e];
            const double ox = F.Ux[idx], oy = F.Uy[idx], oz = F.Uz[idx];
            const bool own = kind == ELL_OWNER;
            const double Px = own ? ux : ox, Py = own ? uy : oy, Pz = own ? uz : oz;
            const double Nx = own ? ox : ux, Ny = own ? oy : uy, Nz = own ? oz : uz;
            f[0] = w * (Px - Nx) + Nx; f[1] = w * (Py - Ny) + Ny; f[2] = w * (Pz - Nz) + Nz;
            if (!own) sgn = -1.0;
            if (LIM) {
                const double o[3] = {ox, oy, oz};
                for (int j = 0; j < 3; ++j) { mx[j] = fmax(mx[j], o[j]); mn[j] = fmin(mn[j], o[j]); }
            }
        }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const double t = S[i] * f[j];
                if (sgn > 0) g[3 * i + j] += t; else g[3 * i + j] -= t;
            }
    }
    const double V = m.V[c];
    for (int q = 0; q < 9; ++q) g[q] = g[q] / V;
    if (!LIM) {
        for (int q = 0; q < 9; ++q) F.gradU[(size_t)q * m.nC + c] = g[q];
        return;
    }
    for (int q = 0; q < 9; ++q) F.gradUc[(size_t)q * m.nC + c] = g[q];
    const double u[3] = {ux, uy, uz};
    for (int j = 0; j < 3; ++j) { mx[j] = mx[j] - u[j]; mn[j] = mn[j] - u[j]; }
    if (k < 1.0) {
        const double wd = 1.0 / k - 1.0;
        for (int j = 0; j < 3; ++j) {
            const double d = wd * (mx[j] - mn[j]);
            mx[j] = mx[j] + d; mn[j] = mn[j] - d;
        }
    }
    double l[3] = {1.0, 1.0, 1.0};
    const size_t nF = m.nF;
    const double cx = m.C[c], cy = m.C[m.nC + c], cz = m.C[2 * (size_t)m.nC + c];
    for (int kk = 0; kk < m.ellW; ++kk) {
        const size_t e = (size_t)kk * m.nC + c;
        const uint32_t kind = m.ellCode[e] >> 30;
        if (kind == ELL_NONE) break;
        const int fi = m.ellFace[e];
        const double dx = m.Cf[fi] - cx, dy = m.Cf[nF + fi] - cy, dz = m.Cf[2 * nF + fi] - cz;
        for (int j = 0; j < 3; ++j) limit_face(l[j], mx[j], mn[j], dx * g[j] + dy * g[3 + j] + dz * g[6 + j]);
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) F.gradU[(size_t)(3 * i + j) * m.nC + c] = l[j] * g[3 * i + j];
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

// a cell's flux side (make_side: 1/rho, U, rPsi, e, a*, c) from its packed
// record: four 16-byte loads of one 64-byte line.  For the schemes with a
// CELL8_SIDE record the side is read as stored (rhoU and c, which those
// fluxes do not read, are left NaN); otherwise it is formed from the state.
template <int S>
__device__ inline Side side_of_cell(const Fields& F, int i, double gamma) {
    const double2* q = reinterpret_cast<const double2*>(F.cell8 + 8 * (size_t)i);
    const double2 a = q[0], b = q[1], c = q[2], d = q[3];
    if constexpr (S == RDST_FLUX_AUSMPLUS || S == RDST_FLUX_AUSMPLUSUP || S == RDST_FLUX_HLL) {
        Side sd;
        sd.rho = a.x; sd.ux = a.y; sd.uy = b.x; sd.uz = b.y;
        sd.rr = c.x; sd.rPsi = c.y; sd.e = d.x; sd.aStar = d.y;
        sd.rux = sd.ruy = sd.ruz = sd.c = __builtin_nan("");
        return sd;
    } else {
        return make_side(a.x, a.y, b.x, b.y, c.x, c.y, d.x, d.y, gamma);
    }
}

// the packed record from the SoA fields, after the host overwrote a field
__global__ void k_pack_cells(int n, Fields F, double gamma) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    store_cell8(F, i, F.rho[i], F.rhoUx[i], F.rhoUy[i], F.rhoUz[i], F.rhoE[i], F.p[i], F.h[i], F.psi[i], gamma);
}

__device__ inline Side load_bnd(const Fields& F, int b, double gamma) {
    return make_side(F.brho[b], F.brhoUx[b], F.brhoUy[b], F.brhoUz[b], F.brhoE[b], F.bp[b], F.bh[b], F.bpsi[b], gamma);
}

// ---------------------------------------------------------------- limited reconstruction
// limitedScheme OSPRE (scalars) / OSPREV (rhoU): face value lambda (P - N) + N,
// lambda = lim w + (1 - lim) pos(faceFlux) with faceFlux = +1 on the owner side
// (own_) and -1 on the neighbour side (nei_); r = 2 (d . grad_upwind)/(phiN -
// phiP) - 1 (NVDTVD; vectors projected on phiN - phiP, NVDVTVDV), capped where
// |d . grad| >= 1000 |phiN - phiP|; OSPRE lim(r) = 1.5 r(r+1)/(r(r+1)+1).
constexpr int NREC = 8;  // rho, rhoUx, rhoUy, rhoUz, rPsi, e, a*, c
__device__ inline double ospre_lim(double r) {
    const double rrp1 = r * (r + 1.0);
    return 1.5 * rrp1 / (rrp1 + 1.0);
}
__device__ inline double r_factor(double gradcf, double gradf) {
    if (fabs(gradcf) >= 1000.0 * fabs(gradf)) return 2.0 * 1000.0 * fsign(gradcf) * fsign(gradf) - 1.0;
    return 2.0 * (gradcf / gradf) - 1.0;
}

// derived cell fields of the reconstruction: rPsi = 1/psi, e = h - p/rho,
// a* from H = max(rhoE/rho, SMALL) + max(p/rho, SMALL), c = sqrt(gamma rPsi)
__global__ void k_recon_prep(MeshDev m, Fields F, double gamma) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m.nC) {
        const Side s = make_side(F.rho[i], F.rhoUx[i], F.rhoUy[i], F.rhoUz[i], F.rhoE[i], F.p[i], F.h[i], F.psi[i], gamma);
        F.rec[i] = s.rPsi; F.rec[m.nC + i] = s.e; F.rec[2 * (size_t)m.nC + i] = s.aStar; F.rec[3 * (size_t)m.nC + i] = s.c;
    }
    if (i < m.nB) {
        const Side s = make_side(F.brho[i], F.brhoUx[i], F.brhoUy[i], F.brhoUz[i], F.brhoE[i], F.bp[i], F.bh[i], F.bpsi[i],
                                 gamma);
        F.brec[i] = s.rPsi; F.brec[m.nB + i] = s.e; F.brec[2 * (size_t)m.nB + i] = s.aStar; F.brec[3 * (size_t)m.nB + i] = s.c;
    }
}
__device__ inline void rec_cell(const Fields& F, int n, int i, double* v) {
    v[0] = F.rho[i]; v[1] = F.rhoUx[i]; v[2] = F.rhoUy[i]; v[3] = F.rhoUz[i];
    for (int q = 0; q < 4; ++q) v[4 + q] = F.rec[(size_t)q * n + i];
}
__device__ inline void rec_bnd(const Fields& F, int nB, int b, double* v) {
    v[0] = F.brho[b]; v[1] = F.brhoUx[b]; v[2] = F.brhoUy[b]; v[3] = F.brhoUz[b];
    for (int q = 0; q < 4; ++q) v[4 + q] = F.brec[(size_t)q * nB + b];
}
// Gauss linear gradients of the reconstructed fields, [24][nC]
__global__ __launch_bounds__(256) void k_recon_grad(MeshDev m, Fields F) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const size_t E = (size_t)m.ellW * m.nC;
    double vc[NREC], g[3 * NREC];
    rec_cell(F, m.nC, c, vc);
    for (int q = 0; q < 3 * NREC; ++q) g[q] = 0.0;
    for (int k = 0; k < m.ellW; ++k) {
        const size_t e = (size_t)k * m.nC + c;
        const uint32_t code = m.ellCode[e];
        const uint32_t kind = code >> 30;
        if (kind == ELL_NONE) break;
        const int idx = (int)(code & ELL_IDX_MASK);
        const double S[3] = {m.ellSf[e], m.ellSf[E + e], m.ellSf[2 * E + e]};
        double vf[NREC];
        if (kind == ELL_BOUNDARY) {
            rec_bnd(F, m.nB, idx, vf);
        } else {
            double vo[NREC];
            rec_cell(F, m.nC, idx, vo);
            const double w = m.ellWt[e];
            for (int q = 0; q < NREC; ++q)
                vf[q] = (kind == ELL_OWNER) ? w * (vc[q] - vo[q]) + vo[q] : w * (vo[q] - vc[q]) + vc[q];
        }
        for (int q = 0; q < NREC; ++q)
            for (int i = 0; i < 3; ++i) {
                const double t = S[i] * vf[q];
                if (kind == ELL_NEIGH) g[3 * q + i] -= t; else g[3 * q + i] += t;
            }
    }
    const double V = m.V[c];
    for (int q = 0; q < 3 * NREC; ++q) F.rgrad[(size_t)q * m.nC + c] = g[q] / V;
}
// owner- and neighbour-side sides of internal face (P, N) by OSPRE/OSPREV
__device__ inline void recon_sides(const MeshDev& m, const Fields& F, int P, int N, double w, Side& L, Side& R) {
    const int n = m.nC;
    double vP[NREC], vN[NREC];
    rec_cell(F, n, P, vP);
    rec_cell(F, n, N, vN);
    const double d0 = m.C[N] - m.C[P], d1 = m.C[n + N] - m.C[n + P], d2 = m.C[2 * (size_t)n + N] - m.C[2 * (size_t)n + P];
    double lo[NREC], ln[NREC];
    // scalars rho (q = 0) and rPsi, e, a*, c (q = 4..7)
    for (int q = 0; q < NREC; ++q) {
        if (q >= 1 && q <= 3) continue;
        const double gf = vN[q] - vP[q];
        const double* gPq = F.rgrad + (size_t)(3 * q) * n;
        const double gcfP = d0 * gPq[P] + d1 * gPq[n + P] + d2 * gPq[2 * (size_t)n + P];
        const double gcfN = d0 * gPq[N] + d1 * gPq[n + N] + d2 * gPq[2 * (size_t)n + N];
        lo[q] = ospre_lim(r_factor(gcfP, gf));
        ln[q] = ospre_lim(r_factor(gcfN, gf));
    }
    // rhoU (OSPREV): one limiter from the projection on phiN - phiP
    {
        const double gv0 = vN[1] - vP[1], gv1 = vN[2] - vP[2], gv2 = vN[3] - vP[3];
        const double gf = gv0 * gv0 + gv1 * gv1 + gv2 * gv2;
        double cP = 0.0, cN = 0.0;
        const double gv[3] = {gv0, gv1, gv2};
        for (int j = 0; j < 3; ++j) {
            const double* gj = F.rgrad + (size_t)(3 * (1 + j)) * n;  // d(rhoU_j)/dx_i at 3(1+j) + i
            const double dgP = d0 * gj[P] + d1 * gj[n + P] + d2 * gj[2 * (size_t)n + P];
            const double dgN = d0 * gj[N] + d1 * gj[n + N] + d2 * gj[2 * (size_t)n + N];
            cP += gv[j] * dgP;
            cN += gv[j] * dgN;
        }
        const double lP = ospre_lim(r_factor(cP, gf)), lN = ospre_lim(r_factor(cN, gf));
        for (int q = 1; q <= 3; ++q) { lo[q] = lP; ln[q] = lN; }
    }
    double fo[NREC], fn[NREC];
    for (int q = 0; q < NREC; ++q) {
        const double wo = lo[q] * w + (1.0 - lo[q]) * 1.0;
        const double wn = ln[q] * w + (1.0 - ln[q]) * 0.0;
        fo[q] = wo * (vP[q] - vN[q]) + vN[q];
        fn[q] = wn * (vP[q] - vN[q]) + vN[q];
    }
    const double* fs[2] = {fo, fn};
    Side* out[2] = {&L, &R};
    for (int k = 0; k < 2; ++k) {
        const double* v = fs[k];
        Side& sd = *out[k];
        sd.rho = v[0]; sd.rux = v[1]; sd.ruy = v[2]; sd.ruz = v[3];
        sd.rr = 1.0 / v[0];
        sd.ux = v[1] * sd.rr; sd.uy = v[2] * sd.rr; sd.uz = v[3] * sd.rr;
        sd.rPsi = v[4]; sd.e = v[5]; sd.aStar = v[6]; sd.c = v[7];
    }
}

// K1: face flux + residual, one thread per cell.
// VISC: 0 inviscid, 1 viscous, 2 viscous with a limited grad(U) (tauMC from
// gradU, the snGrad correction from the unlimited gradUc)
template <int S, int STORE, int VISC, bool RECON>
__global__ __launch_bounds__(256) void k_flux_residual(MeshDev m, Fields F, const BcPatchDev* bc, double gamma, int lowMach,
                                                       double Mainf, double lapPsi, double snPsi, int heat) {
    const int c = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const Side me = side_of_cell<S>(F, c, gamma);
    const double fHc = (S == RDST_FLUX_HLLCP) ? F.fH[c] : 0.0;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0;
    // viscous accumulators: laplacian(muEff, U), div(tauMC), div(sigmaDotU)
    double lx = 0.0, ly = 0.0, lz = 0.0, tx = 0.0, ty = 0.0, tz = 0.0, sd = 0.0, hq = 0.0;
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
        // the ELL stream is read once: non-temporal loads keep it from
        // evicting the cell records that neighbouring cells gather again
        const uint32_t code = __builtin_nontemporal_load(m.ellCode + e);
        const uint32_t kind = code >> 30;
        if (kind == ELL_NONE) break;
        const int idx = (int)(code & ELL_IDX_MASK);
        const double sx = __builtin_nontemporal_load(m.ellSf + e);
        const double sy = __builtin_nontemporal_load(m.ellSf + E + e);
        const double sz = __builtin_nontemporal_load(m.ellSf + 2 * E + e);
        // |Sf| formed here instead of read (8 of the 36 bytes per ELL entry):
        // the same rounded operations as the host's mag(Sf), IEEE sqrt
        const double ms = sqrt(sx * sx + sy * sy + sz * sz);
        Side L, R;
        double f = 1.0;
        if (kind == ELL_BOUNDARY) {
            L = load_bnd(F, idx, gamma);
            R = L;
        } else if (RECON) {
            const bool ownK = kind == ELL_OWNER;
            recon_sides(m, F, ownK ? c : idx, ownK ? idx : c, m.ellWt[e], L, R);
        } else {
            const Side other = side_of_cell<S>(F, idx, gamma);
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
            double lfx, lfy, lfz, tfx, tfy, tfz, sgx, sgy, sgz, qh = 0.0;
            if (kind == ELL_BOUNDARY) {
                const double rms = 1.0 / ms;
                const double nx = sx * rms, ny = sy * rms, nz = sz * rms;
                double snx, sny, snz;
                sngrad_u_bnd(bc[m.bpatch[idx]], F, idx, nx, ny, nz, m.ellDc[e], uCx, uCy, uCz, snx, sny, snz);
                const double bmuE = F.bmut[idx] + F.bmu[idx];
                const double mg = bmuE * ms;
                lfx = mg * snx; lfy = mg * sny; lfz = mg * snz;
                sgx = lfx; sgy = lfy; sgz = lfz;
                double gb[9], tb[9];
                for (int q = 0; q < 9; ++q) gb[q] = F.bgradU[(size_t)q * m.nB + idx];
                tau_mc(bmuE, gb, tb);
                tfx = sx * tb[0] + sy * tb[3] + sz * tb[6];
                tfy = sx * tb[1] + sy * tb[4] + sz * tb[7];
                tfz = sx * tb[2] + sy * tb[5] + sz * tb[8];
                if (heat)  // alphaEff_b |Sf| dc (h_b - h_P)
                    qh = (F.balphat[idx] + F.balpha[idx]) * ms * (m.ellDc[e] * (F.bh[idx] - F.h[c]));
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
                if constexpr (VISC == 2) {
                    const size_t iP = own ? c : idx, iN = own ? idx : c;
                    for (int q = 0; q < 9; ++q) {
                        const double a = F.gradUc[(size_t)q * m.nC + iP], b = F.gradUc[(size_t)q * m.nC + iN];
                        gf[q] = w * (a - b) + b;
                        tf[q] = w * (tP[q] - tN[q]) + tN[q];
                    }
                } else {
                    for (int q = 0; q < 9; ++q) {
                        gf[q] = w * (gP[q] - gN[q]) + gN[q];
                        tf[q] = w * (tP[q] - tN[q]) + tN[q];
                    }
                }
                const double cx = m.ellCorr[e], cy = m.ellCorr[E + e], cz = m.ellCorr[2 * E + e];
                const double dcf = m.ellDc[e];
                // orthogonal part and non-orthogonal correction of snGrad(U);
                // laplacianSchemes and snGradSchemes may limit the latter
                // differently (laplacian(muEff, U) and sigmaDotU)
                const double ux = dcf * (UNx - UPx), uy = dcf * (UNy - UPy), uz = dcf * (UNz - UPz);
                const double qx = cx * gf[0] + cy * gf[3] + cz * gf[6];
                const double qy = cx * gf[1] + cy * gf[4] + cz * gf[7];
                const double qz = cx * gf[2] + cy * gf[5] + cz * gf[8];
                double kL = 1.0, kS = 1.0;
                if (lapPsi >= 0.0 || snPsi >= 0.0) {
                    const double u2 = ux * ux + uy * uy + uz * uz, q2 = qx * qx + qy * qy + qz * qz;
                    kL = sngrad_limiter(lapPsi, u2, q2);
                    kS = sngrad_limiter(snPsi, u2, q2);
                }
                const double mg = muf * ms;
                lfx = mg * (ux + kL * qx); lfy = mg * (uy + kL * qy); lfz = mg * (uz + kL * qz);
                sgx = mg * (ux + kS * qx); sgy = mg * (uy + kS * qy); sgz = mg * (uz + kS * qz);
                tfx = sx * tf[0] + sy * tf[3] + sz * tf[6];
                tfy = sx * tf[1] + sy * tf[4] + sz * tf[7];
                tfz = sx * tf[2] + sy * tf[5] + sz * tf[8];
                if (heat) {  // interp(alphaEff) |Sf| dc (h_N - h_P)
                    const double alC = F.alphat[c] + F.alpha[c], alO = F.alphat[idx] + F.alpha[idx];
                    const double alP = own ? alC : alO, alN = own ? alO : alC;
                    const double hP = own ? F.h[c] : F.h[idx], hN = own ? F.h[idx] : F.h[c];
                    qh = (w * (alP - alN) + alN) * ms * (dcf * (hN - hP));
                }
            }
            const double sdf = (sgx + tfx) * o.upx + (sgy + tfy) * o.upy + (sgz + tfz) * o.upz;
            if (kind == ELL_NEIGH) {
                lx -= lfx; ly -= lfy; lz -= lfz; tx -= tfx; ty -= tfy; tz -= tfz; sd -= sdf; hq -= qh;
            } else {
                lx += lfx; ly += lfy; lz += lfz; tx += tfx; ty += tfy; tz += tfz; sd += sdf; hq += qh;
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
        if (heat) r4 -= hq / V;
    }
    __builtin_nontemporal_store(a0 / V, F.R + c);
    __builtin_nontemporal_store(r1, F.R + n + c);
    __builtin_nontemporal_store(r2, F.R + 2 * n + c);
    __builtin_nontemporal_store(r3, F.R + 3 * n + c);
    __builtin_nontemporal_store(r4, F.R + 4 * n + c);
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
// Force integral over the boundary faces [b0, b1) of one patch, first pass:
// block g sums its grid-stride share of the faces into part[q G + g]
// (q = pressure force x, y, z, viscous force x, y, z).
__global__ __launch_bounds__(RED_BS) void k_forces_part(MeshDev m, Fields F, int b0, int b1, double pRef, int visc,
                                                        double* part) {
    __shared__ double sh[16];
    double a[6] = {0, 0, 0, 0, 0, 0};
    const int G = (int)gridDim.x;
    for (int b = b0 + (int)(blockIdx.x * blockDim.x + threadIdx.x); b < b1; b += G * (int)blockDim.x) {
        const int k = m.nI + b;
        const double S[3] = {m.Sf[k], m.Sf[m.nF + k], m.Sf[2 * m.nF + k]};
        const double dp = F.bp[b] - pRef;
        for (int d = 0; d < 3; ++d) a[d] += S[d] * dp;
        if (visc) {
            double Gu[9];
            for (int q = 0; q < 9; ++q) Gu[q] = F.bgradU[(size_t)q * m.nB + b];
            const double tr = Gu[0] + Gu[4] + Gu[8];
            const double muE = F.bmut[b] + F.bmu[b];
            for (int j = 0; j < 3; ++j) {
                double v = 0.0;
                for (int i = 0; i < 3; ++i) {
                    const double ts = Gu[3 * i + j] + Gu[3 * j + i];
                    v += S[i] * (-muE * (ts - (i == j ? (2.0 / 3.0) * tr : 0.0)));
                }
                a[3 + j] += v;
            }
        }
    }
    for (int q = 0; q < 6; ++q) {
        const double t = block_sum(a[q], sh);
        if (threadIdx.x == 0) part[(size_t)q * G + blockIdx.x] = t;
    }
}

// second pass: one block sums the G partials of each component in block order
__global__ __launch_bounds__(RED_BS) void k_forces_final(const double* part, int G, double* out) {
    __shared__ double sh[16];
    for (int q = 0; q < 6; ++q) {
        const double t = reduce_partials(part + (size_t)q * G, G, sh);
        if (threadIdx.x == 0) out[q] = t;
    }
}

template <int S, int STORE, bool RECON>
void launch_flux_mode(Case& c) {
    MeshDev md = c.meshDev();
    const int bs = 256;
    const int nb = blocks_for(md.nC, bs);
    const BcPatchDev* bc = c.d_bc.p;
    const double lapPsi = sngrad_psi(c.st.laplacian_sngrad, c.st.laplacian_limit);
    const double snPsi = sngrad_psi(c.st.sngrad, c.st.sngrad_limit);
    if (c.st.viscous && grad_u_limited(c.st))
        hipLaunchKernelGGL((k_flux_residual<S, STORE, 2, RECON>), dim3(nb), dim3(bs), 0, c.stream, md, c.f, bc,
                           c.th.gamma, c.st.low_mach, c.st.Mainf, lapPsi, snPsi, c.st.heat_conduction);
    else if (c.st.viscous)
        hipLaunchKernelGGL((k_flux_residual<S, STORE, 1, RECON>), dim3(nb), dim3(bs), 0, c.stream, md, c.f, bc,
                           c.th.gamma, c.st.low_mach, c.st.Mainf, lapPsi, snPsi, c.st.heat_conduction);
    else
        hipLaunchKernelGGL((k_flux_residual<S, STORE, 0, RECON>), dim3(nb), dim3(bs), 0, c.stream, md, c.f, bc,
                           c.th.gamma, c.st.low_mach, c.st.Mainf, lapPsi, snPsi, 0);
}

template <int S, bool RECON>
void launch_flux_store(Case& c, int store) {
    switch (store) {
        case STORE_ALL: launch_flux_mode<S, STORE_ALL, RECON>(c); break;
        case STORE_PHI: launch_flux_mode<S, STORE_PHI, RECON>(c); break;
        default: launch_flux_mode<S, STORE_BPHI, RECON>(c);
    }
}

template <int S>
void launch_flux(Case& c, int store) {
    if (c.st.viscous) grad_u_launch(c);
    if (c.st.reconstruction == RDST_RECON_OSPRE) {
        MeshDev md = c.meshDev();
        const int nmax = std::max(md.nC, md.nB);
        hipLaunchKernelGGL(k_recon_prep, dim3(blocks_for(nmax, 256)), dim3(256), 0, c.stream, md, c.f, c.th.gamma);
        hipLaunchKernelGGL(k_recon_grad, dim3(blocks_for(md.nC, 256)), dim3(256), 0, c.stream, md, c.f);
        launch_flux_store<S, true>(c, store);
    } else {
        launch_flux_store<S, false>(c, store);
    }
}

}  // namespace

void grad_u_launch(Case& c) {
    MeshDev md = c.meshDev();
    const int bs = 256;
    if (grad_u_limited(c.st))
        hipLaunchKernelGGL(k_grad_u<true>, dim3(blocks_for(md.nC, bs)), dim3(bs), 0, c.stream, md, c.f,
                           c.st.grad_u_limit);
    else
        hipLaunchKernelGGL(k_grad_u<false>, dim3(blocks_for(md.nC, bs)), dim3(bs), 0, c.stream, md, c.f, 1.0);
    if (md.nB) hipLaunchKernelGGL(k_grad_u_bnd, dim3(blocks_for(md.nB, bs)), dim3(bs), 0, c.stream, md, c.f, c.d_bc.p);
}

void flux_launch(Case& c, int store) {
    MeshDev md = c.meshDev();
    if (c.packDirty) {
        hipLaunchKernelGGL(k_pack_cells, dim3(blocks_for(md.nC, 256)), dim3(256), 0, c.stream, md.nC, c.f, c.th.gamma);
        c.packDirty = false;
    }
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

// the face-flux kernel alone: the gradients it reads (viscous, limited
// reconstruction) depend only on the state and are formed once, before the
// timed launches
double flux_time(Case& c, int n) {
    if (!c.hasState) throw ArgError("rdst_time_flux: no state", RDST_ERR_STATE);
    flux_launch(c, STORE_BPHI);  // warm-up, and the gradients
    const bool recon = c.st.reconstruction == RDST_RECON_OSPRE;
    auto k1 = [&]() {
        switch (c.st.flux_scheme) {
            case RDST_FLUX_AUSMPLUS: recon ? launch_flux_store<RDST_FLUX_AUSMPLUS, true>(c, STORE_BPHI)
                                           : launch_flux_store<RDST_FLUX_AUSMPLUS, false>(c, STORE_BPHI); break;
            case RDST_FLUX_AUSMPLUSUP: recon ? launch_flux_store<RDST_FLUX_AUSMPLUSUP, true>(c, STORE_BPHI)
                                             : launch_flux_store<RDST_FLUX_AUSMPLUSUP, false>(c, STORE_BPHI); break;
            case RDST_FLUX_HLL: recon ? launch_flux_store<RDST_FLUX_HLL, true>(c, STORE_BPHI)
                                      : launch_flux_store<RDST_FLUX_HLL, false>(c, STORE_BPHI); break;
            case RDST_FLUX_HLLC: recon ? launch_flux_store<RDST_FLUX_HLLC, true>(c, STORE_BPHI)
                                       : launch_flux_store<RDST_FLUX_HLLC, false>(c, STORE_BPHI); break;
            case RDST_FLUX_HLLCP: recon ? launch_flux_store<RDST_FLUX_HLLCP, true>(c, STORE_BPHI)
                                        : launch_flux_store<RDST_FLUX_HLLCP, false>(c, STORE_BPHI); break;
            default: recon ? launch_flux_store<RDST_FLUX_KURGANOV, true>(c, STORE_BPHI)
                           : launch_flux_store<RDST_FLUX_KURGANOV, false>(c, STORE_BPHI);
        }
    };
    RDST_HIP(hipEventRecord(c.ev0, c.stream));
    for (int i = 0; i < n; ++i) k1();
    RDST_HIP(hipEventRecord(c.ev1, c.stream));
    RDST_HIP(hipStreamSynchronize(c.stream));
    float ms = 0.f;
    RDST_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    return (double)ms / n;
}

// Device scratch of the allocation-free force integral (rdst_forces_noalloc):
// the [6][RED_MAXG] block partials and the six sums.  Each GPU holds its own
// copy for the life of the process; host code reaches it only through the
// address hipGetSymbolAddress returns on that device (forces_scratch_setup).
constexpr size_t kForcesScratch = 6 * (size_t)RED_MAXG + 8;
static __device__ __align__(256) double scratch[kForcesScratch];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
double* g_scratchPtr[kMaxDevices] = {};
}  // namespace

void forces_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) throw ArgError("forces scratch: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (g_scratchPtr[device]) return;
    void* p = nullptr;
    RDST_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratchPtr[device] = static_cast<double*>(p);
}

double* forces_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) throw ArgError("forces scratch: device index out of range");
    {
        std::lock_guard<std::mutex> lock(g_scratchMutex);
        if (g_scratchPtr[device]) return g_scratchPtr[device];
    }
    forces_scratch_setup(device);
    return g_scratchPtr[device];
}

size_t forces_work_size() { return kForcesScratch; }

void forces(Case& c, int patch, const double* dragDir, const double* liftDir, double pRef, double rhoInf, double magUInf,
            double aRef, double* out, double* work) {
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
        if (!work) {
            d.alloc(kForcesScratch);
            work = d.p;
        }
        const int G = reduce_grid(b1 - b0);
        if (G > RED_MAXG) throw ArgError("forces: more reduction blocks than the scratch holds");
        double* res = work + 6 * (size_t)RED_MAXG;
        hipLaunchKernelGGL(k_forces_part, dim3(G), dim3(RED_BS), 0, c.stream, c.meshDev(), c.f, b0, b1, pRef,
                           c.st.viscous, work);
        RDST_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_forces_final, dim3(1), dim3(RED_BS), 0, c.stream, work, G, res);
        RDST_HIP(hipGetLastError());
        RDST_HIP(hipMemcpyAsync(F, res, sizeof(F), hipMemcpyDeviceToHost, c.stream));
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
