// jac.hip -- Roe 5x5 Jacobian assembly into the owner/neighbour-sorted
// block-CSR, with the local-time-step diagonal and the source.
//
// Replaces Roe<tensor5>::computeJacobianDST + computeBoundaryJacobians
// (jacobian/Roe/Roe.C:152-379), fvmDST::ddt with steadyStateDST::fvmDdt
// (fvmDST.C:77-143, steadyStateDST.C:173-198) and fluxDST::integrate
// (timeIntegration/fluxDST/fluxDST/fluxDST.C:73-127), summed as
// fvBlockMatrixDST::operator+ does (ddt first, then flux).
//
// One thread per cell.  For each internal face the cell builds the Euler
// Jacobian A(W_cell).n (Blazek A.7, Roe.C:265-271) and the face spectral
// radius Lambda_f, and writes the one off-diagonal block that depends on its
// own state: as owner the lower coefficient -beta (w A n |S| + Lambda I)
// (row neighbour, column owner), as neighbour the upper coefficient
// beta ((1-w) A n |S| - Lambda I) (row owner, column neighbour).  The
// diagonal is the negated sum of the same blocks in face order
// (negSumDiag), plus the boundary chain rule, plus rDeltaTDST V I.  With
// sellOf given, every block is also stored into the sliced-ELL copy the
// Krylov SpMV reads (entry q of slot s at sellOf[s] + 64 q), so the solve
// needs no separate block-CSR -> sliced-ELL pass.
#include "rdst_internal.hpp"

namespace rdst {

namespace {

// the physical-time part of the dual2ndDST ddt (dual2ndDST.C:395-411):
// a = 1/deltaT + 1/(deltaT + deltaT0), b = -(a + c), c = deltaT/(deltaT0
// (deltaT + deltaT0)) and the old time levels of W; on = 0: steadyStateDST
struct DualDdt {
    int on;
    double a, b, c;
    const double *Wo, *Woo;
};

template <bool F32>
__global__ __launch_bounds__(128) void k_assemble(MeshDev m, Fields F, Thermo th, const BcPatchDev* bc, double beta,
                                                  double* A64, float* A32, double* src, DualDdt dd2,
                                                  const int64_t* sellOf, double* S64, float* S32) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const double gam = th.Cp / th.Cv;
    const double rho = F.rho[c], rhoE = F.rhoE[c];
    const double ux = F.rhoUx[c] / rho, uy = F.rhoUy[c] / rho, uz = F.rhoUz[c] / rho;
    const double phiJ = 0.5 * (gam - 1) * (ux * ux + uy * uy + uz * uz);
    const double a1 = gam * rhoE / rho - phiJ;
    const double cc = sqrt(th.Cp / th.Cv / F.psi[c]);
    const double muC = F.mut[c] + F.mu[c], alC = F.alphat[c] + F.alpha[c], trC = F.psi[c] * F.p[c];
    double fd[25];
    for (int q = 0; q < 25; ++q) fd[q] = 0.0;
    const size_t E = (size_t)m.ellW * m.nC;
    for (int k = 0; k < m.ellW; ++k) {
        const size_t e = (size_t)k * m.nC + c;
        const uint32_t code = m.ellCode[e];
        const uint32_t kind = code >> 30;
        if (kind == ELL_NONE) break;
        const int idx = (int)(code & ELL_IDX_MASK);
        const double sx = m.ellSf[e], sy = m.ellSf[E + e], sz = m.ellSf[2 * E + e], ms = m.ellMagSf[e];
        if (kind == ELL_BOUNDARY) {
            // computeBoundaryJacobians (Roe.C:306-377): dR/dXb . dXb/dXi . dXi/dWi
            const BcPatchDev& B = bc[m.bpatch[idx]];
            const double rB = F.brho[idx];
            const double UBx = F.brhoUx[idx] / rB, UBy = F.brhoUy[idx] / rB, UBz = F.brhoUz[idx] / rB;
            const double pB = F.bp[idx], TB = F.bT[idx], CvB = th.Cv;
            const double EB = F.brhoE[idx] / rB;
            const double rRT = rB / pB;
            const double USf = UBx * sx + UBy * sy + UBz * sz;
            const double UB[3] = {UBx, UBy, UBz}, S[3] = {sx, sy, sz};
            double dR[25];
            dR[0] = USf * rRT;
            dR[1] = rB * sx; dR[2] = rB * sy; dR[3] = rB * sz;
            dR[4] = -rB / TB * USf;
            for (int i = 0; i < 3; ++i) {
                dR[5 * (1 + i)] = rRT * USf * UB[i] + S[i];
                for (int j = 0; j < 3; ++j) dR[5 * (1 + i) + 1 + j] = rB * UB[i] * S[j] + (i == j ? rB * USf : 0.0);
                dR[5 * (1 + i) + 4] = -rB / TB * USf * UB[i];
            }
            dR[20] = (EB * rRT + 1) * USf;
            for (int j = 0; j < 3; ++j) dR[21 + j] = rB * USf * UB[j] + S[j] * (rB * EB + pB);
            dR[24] = rB * CvB * USf;
            // dX/dW at the cell
            const double rRho = 1.0 / rho, Cv = th.Cv;
            const double rRhoCv = rRho / Cv;
            const double U2 = ux * ux + uy * uy + uz * uz;
            const double eI = rhoE * rRho - 0.5 * U2;
            const double RT = F.p[c] * rRho;
            const double RCv = RT / eI;
            const double UI[3] = {ux, uy, uz};
            double dX[25];
            for (int q = 0; q < 25; ++q) dX[q] = 0.0;
            dX[0] = RT - RCv * (eI - 0.5 * U2);
            dX[1] = -RCv * ux; dX[2] = -RCv * uy; dX[3] = -RCv * uz;
            dX[4] = RCv;
            for (int i = 0; i < 3; ++i) {
                dX[5 * (1 + i)] = -rRho * UI[i];
                dX[5 * (1 + i) + 1 + i] = rRho;
            }
            dX[20] = rRhoCv * (0.5 * U2 - eI);
            dX[21] = -rRhoCv * ux; dX[22] = -rRhoCv * uy; dX[23] = -rRhoCv * uz;
            dX[24] = rRhoCv;
            // dXb/dXi = diag(valueInternalCoeffs)
            double D[5];
            auto svic = [](int t, double vf) {
                return t == RDST_BC_FIXED_VALUE ? 0.0 : ((t == RDST_BC_INLET_OUTLET || t == RDST_BC_FARFIELD) ? 1.0 - vf : 1.0);
            };
            D[0] = svic(B.p_type, F.vf_p[idx]);
            D[4] = svic(B.T_type, F.vf_T[idx]);
            const double nx = sx / ms, ny = sy / ms, nz = sz / ms;
            const double nn[3] = {nx, ny, nz};
            for (int i = 0; i < 3; ++i) {
                double v;
                switch (B.U_type) {
                    case RDST_BC_FIXED_VALUE: v = 0.0; break;
                    case RDST_BC_ZERO_GRADIENT: v = 1.0; break;
                    case RDST_BC_SYMMETRY: v = 1.0 - fabs(nn[i]); break;
                    case RDST_BC_INLET_OUTLET: case RDST_BC_FARFIELD: v = 1.0 - F.vf_U[idx]; break;
                    default: v = 0.0;
                }
                D[1 + i] = v;
            }
            // (dR . D) . dX, then diag += beta (patchInternalCoeffs + 0)
            for (int i = 0; i < 5; ++i)
                for (int j = 0; j < 5; ++j) {
                    double s = 0.0;
                    for (int q = 0; q < 5; ++q) s += (dR[5 * i + q] * D[q]) * dX[5 * q + j];
                    fd[5 * i + j] += beta * (s + 0.0);
                }
            continue;
        }
        // the other cell of the face
        const double orho = F.rho[idx];
        const double oux = F.rhoUx[idx] / orho, ouy = F.rhoUy[idx] / orho, ouz = F.rhoUz[idx] / orho;
        const double occ = sqrt(th.Cp / th.Cv / F.psi[idx]);
        const double w = m.ellWt[e];
        const bool own = kind == ELL_OWNER;
        // owner (P) / neighbour (N) values for the linear interpolates w(P-N)+N
        const double Px = own ? ux : oux, Py = own ? uy : ouy, Pz = own ? uz : ouz, Nx = own ? oux : ux, Ny = own ? ouy : uy,
                     Nz = own ? ouz : uz;
        const double Pc = own ? cc : occ, Nc = own ? occ : cc;
        const double omu = F.mut[idx] + F.mu[idx], oal = F.alphat[idx] + F.alpha[idx];
        const double Pmu = own ? muC : omu, Nmu = own ? omu : muC;
        const double Pal = own ? alC : oal, Nal = own ? oal : alC;
        const double otr = F.psi[idx] * F.p[idx];
        const double Ptr = own ? trC : otr, Ntr = own ? otr : trC;
        const double Ufx = w * (Px - Nx) + Nx, Ufy = w * (Py - Ny) + Ny, Ufz = w * (Pz - Nz) + Nz;
        const double cf = w * (Pc - Nc) + Nc;
        const double muf = w * (Pmu - Nmu) + Nmu, alf = w * (Pal - Nal) + Nal, trf = w * (Ptr - Ntr) + Ntr;
        const double amaxSf = fabs(Ufx * sx + Ufy * sy + Ufz * sz) + ms * cf;
        const double lam = 0.5 * (amaxSf + ms * (muf + alf) / trf * m.ellDc[e]);
        // A(W_cell).n (Roe.C:265-271)
        const double nx = sx / ms, ny = sy / ms, nz = sz / ms;
        const double nv[3] = {nx, ny, nz}, uv[3] = {ux, uy, uz};
        const double Un = ux * nx + uy * ny + uz * nz;
        double J[25];
        J[0] = 0.0; J[1] = nx; J[2] = ny; J[3] = nz; J[4] = 0.0;
        for (int i = 0; i < 3; ++i) {
            J[5 * (1 + i)] = phiJ * nv[i] - uv[i] * Un;
            for (int j = 0; j < 3; ++j)
                J[5 * (1 + i) + 1 + j] = uv[i] * nv[j] + (i == j ? Un : 0.0) - (gam - 1) * (nv[i] * uv[j]);
            J[5 * (1 + i) + 4] = (gam - 1) * nv[i];
        }
        J[20] = Un * (phiJ - a1);
        for (int j = 0; j < 3; ++j) J[21 + j] = a1 * nv[j] - Un * (gam - 1) * uv[j];
        J[24] = gam * Un;
        const double wt = own ? w : (1 - w);
        const int slot = m.ellSlot[e];
        double* blk64 = A64 + 25 * (size_t)slot;
        float* blk32 = F32 ? A32 + 25 * (size_t)slot : nullptr;
        const size_t sb = sellOf ? (size_t)sellOf[slot] : 0;
        for (int q = 0; q < 25; ++q) {
            const double lq = (q % 6 == 0) ? lam : 0.0;
            const double jw = J[q] * wt;
            const double v = own ? -beta * (jw * ms + lq) : beta * (jw * ms - lq);
            if (F32) blk32[q] = (float)v;
            else blk64[q] = v;
            if (sellOf) {
                if (F32) S32[sb + 64 * q] = (float)v;
                else S64[sb + 64 * q] = v;
            }
            fd[q] -= v;
        }
    }
    // ddt diagonal (steadyStateDST: rDeltaTDST V on all five rows; dual2ndDST:
    // (rDeltaTDST + a) V) + flux diagonal
    const double V = m.V[c];
    const double diagS = dd2.on ? (F.rDeltaT[c] + dd2.a) * V : F.rDeltaT[c] * V;
    const int ds = m.slotDiag[c];
    const size_t sbd = sellOf ? (size_t)sellOf[ds] : 0;
    for (int q = 0; q < 25; ++q) {
        const double dd = (q % 6 == 0) ? diagS : 0.0;
        const double v = dd + fd[q];
        if (F32) A32[25 * (size_t)ds + q] = (float)v;
        else A64[25 * (size_t)ds + q] = v;
        if (sellOf) {
            if (F32) S32[sbd + 64 * q] = (float)v;
            else S64[sbd + 64 * q] = v;
        }
    }
    // source: (rDeltaT psi_old V - diag psi) + (-R V), psi_old = psi; dual2ndDST:
    // V (rDeltaT psi - b psi.oldTime() - c psi.oldTime().oldTime()) - diag psi
    const double x[5] = {rho, F.rhoUx[c], F.rhoUy[c], F.rhoUz[c], rhoE};
    for (int q = 0; q < 5; ++q) {
        const double sd = dd2.on ? V * (F.rDeltaT[c] * x[q] - dd2.b * dd2.Wo[5 * (size_t)c + q] -
                                        dd2.c * dd2.Woo[5 * (size_t)c + q]) -
                                       diagS * x[q]
                                 : F.rDeltaT[c] * x[q] * V - diagS * x[q];
        const double fs = 0.0 - F.R[(size_t)q * m.nC + c] * V;
        src[5 * (size_t)c + q] = sd + fs;
    }
}

}  // namespace

void jacobian_assemble(Case& c) {
    if (!c.hasState) throw ArgError("rdst_jacobian_assemble: no state", RDST_ERR_STATE);
    Mesh& m = *c.mesh;
    MeshDev md = c.meshDev();
    const bool mixed = c.st.precision == RDST_PREC_MIXED;
    DualDdt dd2{0, 0.0, 0.0, 0.0, nullptr, nullptr};
    if (c.st.time_integration == RDST_TIME_DUAL && c.timeIndex > 0) {
        const double rDeltaT = 1.0 / c.deltaT;
        dd2.on = 1;
        dd2.a = rDeltaT + 1.0 / (c.deltaT + c.deltaT0);
        dd2.c = c.deltaT / (c.deltaT0 * (c.deltaT + c.deltaT0));
        dd2.b = -(dd2.a + dd2.c);
        dd2.Wo = c.WOld.p;
        dd2.Woo = c.WOldOld.p;
    }
    if (c.st.jacobian == RDST_JACOBIAN_EXACT) {
        // the reference's Roe operator, kept as the preconditioner source
        const size_t nn = 25 * (size_t)m.nnzb;
        if (mixed) {
            if (c.P32.n != nn) c.P32.alloc(nn);
            hipLaunchKernelGGL((k_assemble<true>), dim3(blocks_for(m.nC, 128)), dim3(128), 0, c.stream, md, c.f, c.th,
                               c.d_bc.p, c.st.beta, (double*)nullptr, c.P32.p, c.src.p, dd2, (const int64_t*)nullptr,
                               (double*)nullptr, (float*)nullptr);
        } else {
            if (c.P64.n != nn) c.P64.alloc(nn);
            hipLaunchKernelGGL((k_assemble<false>), dim3(blocks_for(m.nC, 128)), dim3(128), 0, c.stream, md, c.f, c.th,
                               c.d_bc.p, c.st.beta, c.P64.p, (float*)nullptr, c.src.p, dd2, (const int64_t*)nullptr,
                               (double*)nullptr, (float*)nullptr);
        }
        newton_operator(c);
        RDST_HIP(hipGetLastError());
        c.hasMatrix = true;
        return;
    }
    // block-CSR for the ILU(0) factor and the read-back, sliced-ELL for the SpMV
    const size_t nn = 25 * (size_t)m.nnzb, nv = 25 * (size_t)m.sellBlocks;
    if (mixed) {
        if (c.A32.n != nn) c.A32.alloc(nn);
        if (c.S32.n != nv) {
            c.S32.alloc(nv);
            c.S32.zero(c.stream);  // padding blocks stay zero
        }
        hipLaunchKernelGGL((k_assemble<true>), dim3(blocks_for(m.nC, 128)), dim3(128), 0, c.stream, md, c.f, c.th, c.d_bc.p,
                           c.st.beta, (double*)nullptr, c.A32.p, c.src.p, dd2, m.d_sellOfSlot.p, (double*)nullptr, c.S32.p);
    } else {
        if (c.A64.n != nn) c.A64.alloc(nn);
        if (c.S64.n != nv) {
            c.S64.alloc(nv);
            c.S64.zero(c.stream);
        }
        hipLaunchKernelGGL((k_assemble<false>), dim3(blocks_for(m.nC, 128)), dim3(128), 0, c.stream, md, c.f, c.th, c.d_bc.p,
                           c.st.beta, c.A64.p, (float*)nullptr, c.src.p, dd2, m.d_sellOfSlot.p, c.S64.p, (float*)nullptr);
    }
    c.sellFresh = true;
    c.csrStale = false;
    RDST_HIP(hipGetLastError());
    c.hasMatrix = true;
}

}  // namespace rdst
