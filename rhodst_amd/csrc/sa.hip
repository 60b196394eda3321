// sa.hip -- Spalart-Allmaras turbulence: RASModel::correct() of SAneg
// (libraries/turbulence/SAneg/SAneg.C:490-605) and SARC2
// (libraries/turbulence/SARC2/SARC2.C:418-488), called after updateFields
// (Single.C:102).  mut and alphat = mut/Prt feed muEff = mut + mu and
// alphaEff = alphat + alpha of the face loop and the Roe spectral radius.
//
// Design (MI355X):
//  * one thread per cell assembles its whole row of the nuTilda matrix --
//    steadyStateDST ddt, upwind convection with the face mass flux, the
//    corrected laplacian of DnuTildaEff, the Cb2 gradient term, production
//    and implicit destruction, fvMatrix::relax and the boundary
//    coefficients -- into the block-CSR slots of the mesh (one scalar per
//    slot), walking the same coalesced ELL face map as the face loop;
//  * the smoothSolver's symmetric Gauss-Seidel runs block-Jacobi over the
//    mesh's RCB partitions (<= 1024 cells): one workgroup per partition
//    holds its part of nuTilda in LDS and sweeps it level by level (a level
//    schedule of the lower-triangular graph gives exactly the natural-order
//    sweep); couplings across partitions use the values at the start of the
//    sweep, as foam's processor interfaces do.  One partition is the serial
//    sweep of the reference;
//  * the residual norms are deterministic two-level reductions; the host
//    reads one scalar per sweep to take the convergence decision.
#include <algorithm>
#include <cmath>
#include <vector>

#include "rdst_internal.hpp"
#include "reduce.hpp"

namespace rdst {

// no floating-point contraction: the sign tests on phi and nuTilda (upwind
// weights, inletOutlet, pos(nuTilda)) must see the reference's rounding
#pragma clang fp contract(off)

namespace {

constexpr double kSMALL = 1.0e-15;

struct SaDev {
    const rdst_scalar_bc* bc;  // per patch
    double *nut, *bnut, *d, *rhoPrev, *fv1o, *gn, *B, *x2, *A, *Sold;
};

__device__ inline double sq(double s) { return s * s; }
__device__ inline double pow3(double s) { return s * sq(s); }
__device__ inline double pow6(double s) { return pow3(sq(s)); }
__device__ inline double fpos(double s) { return s >= 0.0 ? 1.0 : 0.0; }  // foam-extend pos()
__device__ inline double fneg(double s) { return s < 0.0 ? 1.0 : 0.0; }

__device__ inline double sa_fv1(const rdst_turbulence& t, double chi) {
    const double chi3 = pow3(chi);
    return chi3 / (chi3 + pow3(t.Cv1));
}
__device__ inline double sa_fv2(const rdst_turbulence& t, double chi) { return 1.0 / pow3(1.0 + chi / t.Cv2); }
__device__ inline double sa_fv3(const rdst_turbulence& t, double chi, double fv1) {
    const double cb = (1.0 / t.Cv2) * chi;
    return (1.0 + chi * fv1) * (1.0 / t.Cv2) * (3.0 * (1.0 + cb) + sq(cb)) / pow3(1.0 + cb);
}
__device__ inline double sa_fw(const rdst_turbulence& t, double Stilda, double nut, double d) {
    const double r = fmin(nut / (fmax(Stilda, kSMALL) * sq(t.kappa * d)), 10.0);
    const double g = r + t.Cw2 * (pow6(r) - r);
    return g * pow((1.0 + pow6(t.Cw3)) / (pow6(g) + pow6(t.Cw3)), 1.0 / 6.0);
}
__device__ inline double sa_gamma(const rdst_turbulence& t, double rho, double nut, double mu) {
    if (t.model == RDST_TURB_SANEG) {
        const double chi = rho * nut / mu;
        const double p3 = pow(chi, 3.0);
        const double fn = fpos(chi) + fneg(chi) * (t.Cn1 + p3) / (t.Cn1 - p3);
        return (rho * nut * fn + mu) / t.sigma_nut;
    }
    return rho * nut / t.sigma_nut + mu;
}
__device__ inline double mag_sqr_skew(const double* g) {
    double s = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) s += sq(0.5 * (g[3 * i + j] - g[3 * j + i]));
    return s;
}
__device__ inline double mag_sqr_symm(const double* g) {
    double s = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) s += sq(0.5 * (g[3 * i + j] + g[3 * j + i]));
    return s;
}
__device__ inline double sa_fr1(const rdst_turbulence& t, const double* g) {
    if (!t.rcm_correction) return 1.0;
    const double sqrS = 2.0 * mag_sqr_symm(g), sqrW = 2.0 * mag_sqr_skew(g);
    const double rStar = sqrt(sqrS / fmax(sqrW, kSMALL));
    const double rrStar = sqrt(sqrW / fmax(sqrS, kSMALL));
    const double rTilda = rrStar * (rrStar - 1.0);
    return (1.0 + t.Cr1) * 2.0 * rStar / (1.0 + rStar) * (1.0 - t.Cr3 * atan(t.Cr2 * rTilda)) - t.Cr1;
}

// SARC1::fr1 (SARC1.C:125-176): Spalart-Shur, with DSDt[9] = ddt(rho, S_ij) +
// div(phi, S_ij) per unit volume; sqrS = 2|S|^2, sqrW = 2|W|^2, sqrD =
// (sqrS + sqrW)/2, rTilda = 2 / (rho max(sqrD, SMALL)^2) sum_ij (sum_k W_ik
// S_jk) DSDt_ij
__device__ inline double sa_fr1_ss(const rdst_turbulence& t, const double* g, double rho, const double* DSDt) {
    if (!t.rcm_correction) return 1.0;
    double S[9], W[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            S[3 * i + j] = 0.5 * (g[3 * i + j] + g[3 * j + i]);
            W[3 * i + j] = 0.5 * (g[3 * i + j] - g[3 * j + i]);
        }
    double sqrS = 0.0, sqrW = 0.0;
    for (int q = 0; q < 9; ++q) { sqrS += S[q] * S[q]; sqrW += W[q] * W[q]; }
    sqrS *= 2.0;
    sqrW *= 2.0;
    const double sqrD = 0.5 * (sqrS + sqrW);
    const double rStar = sqrt(sqrS / fmax(sqrW, kSMALL));
    double sum = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double ws = 0.0;
            for (int k = 0; k < 3; ++k) ws += W[3 * i + k] * S[3 * j + k];
            sum += ws * DSDt[3 * i + j];
        }
    const double rTilda = 2.0 / (rho * sq(fmax(sqrD, kSMALL))) * sum;
    return (1.0 + t.Cr1) * 2.0 * rStar / (1.0 + rStar) * (1.0 - t.Cr3 * atan(t.Cr2 * rTilda)) - t.Cr1;
}
__device__ inline double symm_comp(const double* gU, size_t stride, size_t i, int q) {
    const int a = q / 3, b = q % 3;
    return 0.5 * (gU[(size_t)(3 * a + b) * stride + i] + gU[(size_t)(3 * b + a) * stride + i]);
}
// S.oldTime() for the next correct()
__global__ void k_sa_store_S(int n, Fields F, SaDev s) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    for (int q = 0; q < 9; ++q) s.Sold[(size_t)q * n + c] = symm_comp(F.gradU, n, c, q);
}

// ---------------------------------------------------------------- boundary values
// nuTilda boundary values (fixedValue, inletOutlet on phi, else zeroGradient)
__global__ void k_sa_nut_bnd(MeshDev m, Fields F, SaDev s, const double* x) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.nB) return;
    const rdst_scalar_bc bc = s.bc[m.bpatch[b]];
    const double vI = x[m.bcell[b]];
    double v;
    switch (bc.type) {
        case RDST_BC_FIXED_VALUE: v = bc.value; break;
        case RDST_BC_INLET_OUTLET: {
            const double vf = 1.0 - fpos(F.phi[m.nI + b]);
            v = vf * bc.value + (1.0 - vf) * vI;
            break;
        }
        default: v = vI;
    }
    s.bnut[b] = v;
}
// mut on the boundary: 0 on walls, else rho_b nuTilda_b |fv1(chi_b)| >= 0
__global__ void k_sa_mut_bnd(MeshDev m, Fields F, SaDev s, rdst_turbulence t) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.nB) return;
    double v = 0.0;
    if (!m.bwall[b]) {
        const double chi = F.brho[b] * s.bnut[b] / F.bmu[b];
        v = fmax(fabs(sa_fv1(t, chi)) * s.bnut[b] * F.brho[b], 0.0);
    }
    F.bmut[b] = v;
    F.balphat[b] = v / t.Prt;
}
__global__ void k_bnd_floor(int nB, double* v, double v0) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nB) v[b] = fmax(v[b], v0);
}

// ---------------------------------------------------------------- gradient
// fvc::grad(nuTilda), Gauss linear, [3][nC]
__global__ __launch_bounds__(256) void k_sa_grad(MeshDev m, SaDev s) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const size_t E = (size_t)m.ellW * m.nC;
    const double xc = s.nut[c];
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (int k = 0; k < m.ellW; ++k) {
        const size_t e = (size_t)k * m.nC + c;
        const uint32_t code = m.ellCode[e];
        const uint32_t kind = code >> 30;
        if (kind == ELL_NONE) break;
        const int idx = (int)(code & ELL_IDX_MASK);
        double vf;
        if (kind == ELL_BOUNDARY) {
            vf = s.bnut[idx];
        } else {
            const double w = m.ellWt[e], xo = s.nut[idx];
            vf = (kind == ELL_OWNER) ? w * (xc - xo) + xo : w * (xo - xc) + xc;
        }
        const double t0 = m.ellSf[e] * vf, t1 = m.ellSf[E + e] * vf, t2 = m.ellSf[2 * E + e] * vf;
        if (kind == ELL_NEIGH) { g0 -= t0; g1 -= t1; g2 -= t2; }
        else { g0 += t0; g1 += t1; g2 += t2; }
    }
    const double V = m.V[c];
    s.gn[c] = g0 / V;
    s.gn[m.nC + c] = g1 / V;
    s.gn[2 * (size_t)m.nC + c] = g2 / V;
}

// ---------------------------------------------------------------- assembly
// one thread per cell: the row of the relaxed nuTilda matrix in solve form
// (diagonal incl. internalCoeffs at slotDiag, off-diagonals at the face
// slots, source incl. boundaryCoeffs), and fv1(chi) of the current state
template <bool SS>
__global__ __launch_bounds__(256) void k_sa_assemble(MeshDev m, Fields F, SaDev s, rdst_turbulence t, int useRhoPrev,
                                                     int useSold, double lapPsi) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const size_t E = (size_t)m.ellW * m.nC;
    const int n = m.nC;
    const double V = m.V[c], rho = F.rho[c], mu = F.mu[c], nut = s.nut[c], d = s.d[c];
    const double rhoOld = useRhoPrev ? s.rhoPrev[c] : rho;
    const double Gc = sa_gamma(t, rho, nut, mu);
    const double gnc0 = s.gn[c], gnc1 = s.gn[n + c], gnc2 = s.gn[2 * (size_t)n + c];
    // fvm::ddt(rho, nuTilda) (steadyStateDST.C:230-255)
    double diag = F.rDeltaT[c] * rho * V;
    double src = F.rDeltaT[c] * rhoOld * nut * V;
    double sumOff = 0.0, icRel = 0.0, icRaw = 0.0, bcfSum = 0.0, srcBndRelax = 0.0;
    const double alpha = t.relax;
    // SARC1: div(phi, S_ij), Gauss upwind, and S of this cell
    double dS[9], Sc[9];
    if (SS)
        for (int q = 0; q < 9; ++q) { dS[q] = 0.0; Sc[q] = symm_comp(F.gradU, n, c, q); }
    for (int k = 0; k < m.ellW; ++k) {
        const size_t e = (size_t)k * m.nC + c;
        const uint32_t code = m.ellCode[e];
        const uint32_t kind = code >> 30;
        if (kind == ELL_NONE) break;
        const int idx = (int)(code & ELL_IDX_MASK);
        const double ms = m.ellMagSf[e];
        if (kind == ELL_BOUNDARY) {
            const rdst_scalar_bc bc = s.bc[m.bpatch[idx]];
            const double phib = F.phi[m.nI + idx];
            double vic, vbc, lw;
            switch (bc.type) {
                case RDST_BC_FIXED_VALUE: vic = 0.0; vbc = bc.value; lw = 1.0; break;
                case RDST_BC_INLET_OUTLET: {
                    const double vf = 1.0 - fpos(phib);
                    vic = 1.0 - vf; vbc = vf * bc.value; lw = vf;
                    break;
                }
                default: vic = 1.0; vbc = 0.0; lw = 0.0;
            }
            double ic = phib * vic;
            double bcf = -phib * vbc;
            if (lw != 0.0) {
                const double bG = sa_gamma(t, F.brho[idx], s.bnut[idx], F.bmu[idx]);
                const double gms = bG * ms;
                const double dcb = m.ellDc[e];
                ic += gms * lw * dcb;
                bcf += gms * lw * dcb * bc.value;
            }
            if (alpha > 0) {
                const double ica = fabs(ic);
                sumOff -= ica;
                const double icr = ica / alpha;
                srcBndRelax += (icr - ic) * nut;
                icRel += icr;
            }
            icRaw += ic;
            bcfSum += bcf;
            if (SS)
                for (int q = 0; q < 9; ++q) dS[q] += phib * symm_comp(F.bgradU, m.nB, idx, q);
            continue;
        }
        const bool own = kind == ELL_OWNER;
        const int fi = m.ellFace[e];
        const double phi = F.phi[fi];
        // upwind convection: lower = -w phi, upper = lower + phi, w = pos(phi)
        const double lo = -fpos(phi) * phi;
        const double up = lo + phi;
        if (SS) {
            const double wo = fpos(phi);  // upwind weight of the owner side
            for (int q = 0; q < 9; ++q) {
                const double So = symm_comp(F.gradU, n, idx, q);
                const double sf = own ? wo * Sc[q] + (1.0 - wo) * So : wo * So + (1.0 - wo) * Sc[q];
                if (own) dS[q] += phi * sf;
                else dS[q] -= phi * sf;
            }
        }
        // laplacian coefficient Gamma_f |Sf| deltaCoeffs and the explicit correction
        const double w = m.ellWt[e];
        const double Go = sa_gamma(t, F.rho[idx], s.nut[idx], F.mu[idx]);
        const double GP = own ? Gc : Go, GN = own ? Go : Gc;
        const double gf = w * (GP - GN) + GN;
        const double gms = gf * ms;
        const double coef = gms * m.ellDc[e];
        const double go0 = s.gn[idx], go1 = s.gn[n + idx], go2 = s.gn[2 * (size_t)n + idx];
        const double P0 = own ? gnc0 : go0, P1 = own ? gnc1 : go1, P2 = own ? gnc2 : go2;
        const double N0 = own ? go0 : gnc0, N1 = own ? go1 : gnc1, N2 = own ? go2 : gnc2;
        const double f0 = w * (P0 - N0) + N0, f1 = w * (P1 - N1) + N1, f2 = w * (P2 - N2) + N2;
        const double q = m.ellCorr[e] * f0 + m.ellCorr[E + e] * f1 + m.ellCorr[2 * E + e] * f2;
        double kL = 1.0;
        if (lapPsi >= 0.0) {  // limitedSnGrad / uncorrected on the nuTilda laplacian
            const double un = m.ellDc[e] * (own ? s.nut[idx] - nut : nut - s.nut[idx]);
            kL = sngrad_limiter(lapPsi, un * un, q * q);
        }
        const double corr = gms * (kL * q);
        double off;
        if (own) {
            diag -= lo;
            off = up;
            src += corr;
        } else {
            diag -= up;
            off = lo;
            src -= corr;
        }
        off -= coef;
        diag += coef;
        if (t.convection == 1) {
            // linearUpwind grad(nuTilda): phi (Cf - C_u).grad(nuTilda)_u of the
            // upwind cell u (the owner when phi > 0) leaves the source
            const int u = (phi > 0.0) == own ? c : idx;
            const double gu0 = u == c ? gnc0 : go0, gu1 = u == c ? gnc1 : go1, gu2 = u == c ? gnc2 : go2;
            const double lc = phi * ((m.Cf[fi] - m.C[u]) * gu0 + (m.Cf[m.nF + fi] - m.C[n + u]) * gu1 +
                                     (m.Cf[2 * (size_t)m.nF + fi] - m.C[2 * (size_t)n + u]) * gu2);
            if (own) src -= lc;
            else src += lc;
        }
        s.A[own ? m.slotUpper[fi] : m.slotLower[fi]] = off;  // this row's slot of the face
        sumOff += fabs(off);
    }
    // cell sources
    const double chi = rho * nut / mu;
    const double fv1 = sa_fv1(t, chi);
    const double fv2 = sa_fv2(t, chi);
    double g[9];
    for (int q = 0; q < 9; ++q) g[q] = F.gradU[(size_t)q * n + c];
    const double Omega = sqrt(2.0) * sqrt(mag_sqr_skew(g));
    const double Cw1 = t.Cb1 / sq(t.kappa) + (1.0 + t.Cb2) / t.sigma_nut;
    src += t.Cb2 / t.sigma_nut * rho * (gnc0 * gnc0 + gnc1 * gnc1 + gnc2 * gnc2) * V;
    if (t.model == RDST_TURB_SANEG) {
        double Stilda;
        if (t.stilda_modification) {
            const double Sbar = fv2 * nut / sq(t.kappa * d);
            const double a = t.Cn2 * Omega + Sbar;
            Stilda = Omega + fpos(a) * Sbar
                   + fneg(a) * (Omega * (sq(t.Cn2) * Omega + t.Cn3 * Sbar)) / fmax((t.Cn3 - 2.0 * t.Cn2) * Omega - Sbar, kSMALL);
        } else {
            Stilda = sa_fv3(t, chi, fv1) * sqrt(2.0) * sqrt(mag_sqr_skew(g)) + fv2 * nut / sq(t.kappa * d);
        }
        const double ft2 = t.negative_nu_tilda ? t.Ct3 * exp(-t.Ct4 * sq(chi)) : 0.0;
        const double fw = sa_fw(t, Stilda, nut, d);
        if (fpos(nut) > 0.0) {
            src += t.Cb1 * (1.0 - ft2) * rho * Stilda * nut * V;
            diag += (Cw1 * fw - t.Cb1 / sq(t.kappa) * ft2) * nut * rho / sq(d) * V;
        } else {
            src += t.Cb1 * (1.0 - t.Ct3) * rho * Omega * nut * V;
            diag -= Cw1 * nut * rho / sq(d) * V;
        }
    } else {
        const double Stilda = sa_fv3(t, chi, fv1) * sqrt(2.0) * sqrt(mag_sqr_skew(g)) + fv2 * nut / sq(t.kappa * d);
        const double fw = sa_fw(t, Stilda, nut, d);
        double fr1;
        if (SS) {
            // steadyStateDST::fvcDdt(rho, S): rDeltaT (rho S - rho.oldTime() S.oldTime())
            for (int q = 0; q < 9; ++q) {
                const double So = useSold ? s.Sold[(size_t)q * n + c] : Sc[q];
                dS[q] = F.rDeltaT[c] * (rho * Sc[q] - rhoOld * So) + dS[q] / V;
            }
            fr1 = sa_fr1_ss(t, g, rho, dS);
        } else {
            fr1 = sa_fr1(t, g);
        }
        src += t.Cb1 * rho * Stilda * nut * fr1 * V;
        diag += Cw1 * fw * nut * rho / sq(d) * V;
    }
    // fvMatrix::relax(alpha) and the boundary coefficients
    double icAdd = icRaw;
    if (alpha > 0) {
        const double D0 = diag;
        src += srcBndRelax;
        diag = fmax(diag, sumOff);
        diag /= alpha;
        src += (diag - D0) * nut;
        icAdd = icRel;
    }
    s.A[m.slotDiag[c]] = diag + icAdd;
    s.B[c] = src + bcfSum;
    s.fv1o[c] = fv1;
}

// ---------------------------------------------------------------- solve
__global__ __launch_bounds__(RED_BS) void k_sum_partial(long n, const double* x, double* part) {
    __shared__ double sh[16];
    double v = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) v += x[i];
    const double t = block_sum(v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}
__global__ __launch_bounds__(RED_BS) void k_sum_final(const double* part, int G, double* out) {
    __shared__ double sh[16];
    const double t = reduce_partials(part, G, sh);
    if (threadIdx.x == 0) *out = t;
}
// partials of sum|Ax - sumA xRef| + |b - sumA xRef| (normFactor) and of sum|b - Ax|
__global__ __launch_bounds__(RED_BS) void k_sa_norm_partial(int n, const int32_t* rowPtr, const int32_t* cols,
                                                            const double* A, const double* B, const double* x,
                                                            const double* xsum, double* part, int G) {
    __shared__ double sh[16];
    const double xRef = *xsum / n;
    double nf = 0.0, rs = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double ax = 0.0, sa = 0.0;
        for (int p = rowPtr[i]; p < rowPtr[i + 1]; ++p) {
            ax += A[p] * x[cols[p]];
            sa += A[p];
        }
        const double tmp = sa * xRef;
        nf += fabs(ax - tmp) + fabs(B[i] - tmp);
        rs += fabs(B[i] - ax);
    }
    double t = block_sum(nf, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
    t = block_sum(rs, sh);
    if (threadIdx.x == 0) part[G + blockIdx.x] = t;
}
__global__ __launch_bounds__(RED_BS) void k_sa_resid_partial(int n, const int32_t* rowPtr, const int32_t* cols,
                                                             const double* A, const double* B, const double* x,
                                                             double* part) {
    __shared__ double sh[16];
    double rs = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double ax = 0.0;
        for (int p = rowPtr[i]; p < rowPtr[i + 1]; ++p) ax += A[p] * x[cols[p]];
        rs += fabs(B[i] - ax);
    }
    const double t = block_sum(rs, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// one symmetric Gauss-Seidel sweep, block-Jacobi over the partitions:
// xin = values at the start of the sweep, xout = swept values
__global__ __launch_bounds__(256) void k_sa_sgs(const int32_t* partRowStart, const int32_t* partLevStart,
                                                const int32_t* partLevPtr, const int32_t* partRows,
                                                const int32_t* rowPtr, const int32_t* cols, const int32_t* slotDiag,
                                                const int32_t* localCol, const double* A, const double* B,
                                                const double* xin, double* xout) {
    extern __shared__ double lds[];
    const int pt = blockIdx.x;
    const int rs = partRowStart[pt], re = partRowStart[pt + 1];
    const int nl = re - rs;
    double* xl = lds;
    double* bl = lds + nl;
    for (int t = threadIdx.x; t < nl; t += blockDim.x) {
        const int i = partRows[rs + t];
        xl[t] = xin[i];
        double b = B[i];
        for (int p = rowPtr[i]; p < rowPtr[i + 1]; ++p)
            if (p != slotDiag[i] && localCol[p] < 0) b -= A[p] * xin[cols[p]];
        bl[t] = b;
    }
    __syncthreads();
    const int l0 = partLevStart[pt], l1 = partLevStart[pt + 1];
    for (int L = l0; L + 1 < l1; ++L) {
        for (int t = partLevPtr[L] + (int)threadIdx.x; t < partLevPtr[L + 1]; t += blockDim.x) {
            const int i = partRows[t];
            double v = bl[t - rs];
            for (int p = rowPtr[i]; p < rowPtr[i + 1]; ++p) {
                const int lc = localCol[p];
                if (p != slotDiag[i] && lc >= 0) v -= A[p] * xl[lc];
            }
            xl[t - rs] = v / A[slotDiag[i]];
        }
        __syncthreads();
    }
    for (int L = l1 - 2; L >= l0; --L) {
        for (int t = partLevPtr[L] + (int)threadIdx.x; t < partLevPtr[L + 1]; t += blockDim.x) {
            const int i = partRows[t];
            double v = bl[t - rs];
            for (int p = rowPtr[i]; p < rowPtr[i + 1]; ++p) {
                const int lc = localCol[p];
                if (p != slotDiag[i] && lc >= 0) v -= A[p] * xl[lc];
            }
            xl[t - rs] = v / A[slotDiag[i]];
        }
        __syncthreads();
    }
    for (int t = threadIdx.x; t < nl; t += blockDim.x) xout[partRows[rs + t]] = xl[t];
}

// ---------------------------------------------------------------- bound, mut
__global__ __launch_bounds__(RED_BS) void k_min_partial(int n, const double* x, int nB, const double* bx, double* part) {
    __shared__ double sh[16];
    double v = 1.0e300;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v = fmin(v, x[i]);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nB; i += gridDim.x * blockDim.x) v = fmin(v, bx[i]);
    const double t = block_min(v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}
__global__ __launch_bounds__(RED_BS) void k_min_final(const double* part, int G, double* out) {
    __shared__ double sh[16];
    double v = 1.0e300;
    for (int i = threadIdx.x; i < G; i += blockDim.x) v = fmin(v, part[i]);
    const double t = block_min(v, sh);
    if (threadIdx.x == 0) *out = t;
}
// foam bound(): v = max(max(v, average(max(v, v0)) pos(v0 - v)), v0), with
// average the |Sf|-weighted mean of the linear face values of max(v, v0)
__global__ __launch_bounds__(256) void k_bound_avg(MeshDev m, const double* v, const double* bv, double v0, double* out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const double mc = fmax(v[c], v0);
    double num = 0.0, den = 0.0;
    for (int k = 0; k < m.ellW; ++k) {
        const size_t e = (size_t)k * m.nC + c;
        const uint32_t code = m.ellCode[e];
        const uint32_t kind = code >> 30;
        if (kind == ELL_NONE) break;
        const int idx = (int)(code & ELL_IDX_MASK);
        const double ms = m.ellMagSf[e];
        double vf;
        if (kind == ELL_BOUNDARY) {
            vf = fmax(bv[idx], v0);
        } else {
            const double w = m.ellWt[e], mo = fmax(v[idx], v0);
            vf = (kind == ELL_OWNER) ? w * (mc - mo) + mo : w * (mo - mc) + mc;
        }
        num += ms * vf;
        den += ms;
    }
    const double avg = num / den;
    out[c] = fmax(fmax(v[c], avg * fpos(v0 - v[c])), v0);
}
__global__ void k_sa_mut(int n, Fields F, SaDev s, int absFv1) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double f = absFv1 ? fabs(s.fv1o[i]) : s.fv1o[i];
    F.mut[i] = f * s.nut[i] * F.rho[i];
}
__global__ void k_sa_alphat(int n, Fields F, double Prt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) F.alphat[i] = F.mut[i] / Prt;
}
__global__ void k_sa_mut_init(int n, Fields F, SaDev s, rdst_turbulence t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double chi = F.rho[i] * s.nut[i] / F.mu[i];
    const double f = sa_fv1(t, chi);
    const double v = fmax((t.model == RDST_TURB_SANEG ? fabs(f) : f) * s.nut[i] * F.rho[i], 0.0);
    F.mut[i] = v;
    F.alphat[i] = v / t.Prt;
}

}  // namespace

// ---------------------------------------------------------------- host
static SaDev sa_dev(Case& c) {
    SaState& s = c.sa;
    SaDev d;
    d.bc = s.d_bc.p;
    d.nut = s.nut; d.bnut = s.bnut; d.d = s.d; d.rhoPrev = s.rhoPrev; d.fv1o = s.fv1o; d.gn = s.gn;
    d.B = s.B; d.x2 = s.x2; d.A = s.A.p; d.Sold = s.Sold;
    return d;
}

static double read_scalar(Case& c, const double* dptr) {
    double v = 0.0;
    RDST_HIP(hipMemcpyAsync(&v, dptr, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    RDST_HIP(hipStreamSynchronize(c.stream));
    return v;
}

// min over cells and boundary values; returns true when below v0
static bool below(Case& c, const double* x, const double* bx, double v0) {
    const Mesh& m = *c.mesh;
    const int G = reduce_grid(std::max(m.nC, m.nB));
    hipLaunchKernelGGL(k_min_partial, dim3(G), dim3(RED_BS), 0, c.stream, m.nC, x, m.nB, bx, c.sa.part.p);
    hipLaunchKernelGGL(k_min_final, dim3(1), dim3(RED_BS), 0, c.stream, c.sa.part.p, G, c.sa.scal.p);
    RDST_HIP(hipGetLastError());
    return read_scalar(c, c.sa.scal.p) < v0;
}

static void nut_boundary(Case& c, const double* x) {
    const Mesh& m = *c.mesh;
    if (m.nB)
        hipLaunchKernelGGL(k_sa_nut_bnd, dim3(blocks_for(m.nB, 256)), dim3(256), 0, c.stream, c.meshDev(), c.f, sa_dev(c), x);
}

static void mut_boundary(Case& c) {
    const Mesh& m = *c.mesh;
    if (m.nB)
        hipLaunchKernelGGL(k_sa_mut_bnd, dim3(blocks_for(m.nB, 256)), dim3(256), 0, c.stream, c.meshDev(), c.f, sa_dev(c),
                           c.sa.t);
}

void sa_set(Case& c, const rdst_turbulence* t, const rdst_scalar_bc* bcs, const double* wallDist, const double* nuTilda) {
    Mesh& m = *c.mesh;
    SaState& s = c.sa;
    const size_t n = m.nC, nB = std::max(m.nB, 1);
    s.on = false;
    s.t = *t;
    // SpalartAllmaras: SARC2's equation with fr1 = 1 (SARC2.C:121-140 with
    // RCMCorrection off is foam-extend's stock model)
    if (s.t.model == RDST_TURB_SA) {
        s.t.model = RDST_TURB_SARC2;
        s.t.rcm_correction = 0;
    }
    s.nCorrect = 0;
    s.rhoPrevValid = false;
    // laminar: mut = alphat = 0
    c.turbBuf.zero();
    if (t->model == RDST_TURB_LAMINAR) {
        RDST_HIP(hipStreamSynchronize(c.stream));
        return;
    }
    for (int p = 0; p < m.nPatches; ++p) {
        const int ty = bcs[p].type;
        const bool empty = m.patches[p].type == RDST_PATCH_EMPTY;
        if (!empty && ty != RDST_BC_FIXED_VALUE && ty != RDST_BC_ZERO_GRADIENT && ty != RDST_BC_INLET_OUTLET &&
            ty != RDST_BC_SYMMETRY)
            throw ArgError("rdst_case_set_turbulence: unsupported nuTilda boundary condition type");
    }
    for (size_t i = 0; i < n; ++i)
        if (!(wallDist[i] > 0.0)) throw ArgError("rdst_case_set_turbulence: wall distance must be > 0 in every cell");
    s.bc.assign(bcs, bcs + m.nPatches);
    s.d_bc.upload(s.bc);
    s.buf.alloc(18 * n + nB);
    s.buf.zero();
    double* p = s.buf.p;
    s.nut = p; s.d = p + n; s.rhoPrev = p + 2 * n; s.fv1o = p + 3 * n; s.B = p + 4 * n; s.x2 = p + 5 * n;
    s.gn = p + 6 * n; s.bnut = p + 9 * n; s.Sold = p + 9 * n + nB;
    s.A.alloc(m.nnzb);
    s.A.zero();
    s.part.alloc(2 * RED_MAXG);
    s.scal.alloc(8);
    RDST_HIP(hipMemcpyAsync(s.nut, nuTilda, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    RDST_HIP(hipMemcpyAsync(s.d, wallDist, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    s.on = true;
    nut_boundary(c, s.nut);
    hipLaunchKernelGGL(k_sa_mut_init, dim3(blocks_for(m.nC, 256)), dim3(256), 0, c.stream, m.nC, c.f, sa_dev(c), s.t);
    mut_boundary(c);
    RDST_HIP(hipGetLastError());
    RDST_HIP(hipStreamSynchronize(c.stream));
}

void sa_put_nut(Case& c, const double* nuTilda) {
    if (!c.sa.on) throw ArgError("rdst_case_put: no turbulence model on the case", RDST_ERR_STATE);
    RDST_HIP(hipMemcpyAsync(c.sa.nut, nuTilda, sizeof(double) * c.mesh->nC, hipMemcpyHostToDevice, c.stream));
    nut_boundary(c, c.sa.nut);
    RDST_HIP(hipStreamSynchronize(c.stream));
}

void sa_correct(Case& c, rdst_solver_perf* perf) {
    Mesh& m = *c.mesh;
    SaState& s = c.sa;
    if (!s.on) throw ArgError("rdst_turbulence_correct: no turbulence model on the case", RDST_ERR_STATE);
    const rdst_turbulence& t = s.t;
    const int n = m.nC;
    const bool isNeg = t.model == RDST_TURB_SANEG;
    MeshDev md = c.meshDev();
    const int bs = 256;
    const int nb = blocks_for(n, bs);
    // grad(U) of the updated state, grad(nuTilda), the matrix
    grad_u_launch(c);
    hipLaunchKernelGGL(k_sa_grad, dim3(nb), dim3(bs), 0, c.stream, md, sa_dev(c));
    const int useRhoPrev = (s.nCorrect > 0 && s.rhoPrevValid) ? 1 : 0;
    const double lapPsi = sngrad_psi(c.st.laplacian_sngrad, c.st.laplacian_limit);
    if (t.model == RDST_TURB_SARC1) {
        hipLaunchKernelGGL((k_sa_assemble<true>), dim3(nb), dim3(bs), 0, c.stream, md, c.f, sa_dev(c), t, useRhoPrev,
                           s.nCorrect > 0 ? 1 : 0, lapPsi);
        hipLaunchKernelGGL(k_sa_store_S, dim3(nb), dim3(bs), 0, c.stream, n, c.f, sa_dev(c));
    } else {
        hipLaunchKernelGGL((k_sa_assemble<false>), dim3(nb), dim3(bs), 0, c.stream, md, c.f, sa_dev(c), t, useRhoPrev,
                           0, lapPsi);
    }
    RDST_HIP(hipGetLastError());
    // smoothSolver: normFactor, initial residual
    const int G = reduce_grid(n);
    double* part = s.part.p;
    double* sc = s.scal.p;
    hipLaunchKernelGGL(k_sum_partial, dim3(G), dim3(RED_BS), 0, c.stream, (long)n, s.nut, part);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(RED_BS), 0, c.stream, part, G, sc);
    hipLaunchKernelGGL(k_sa_norm_partial, dim3(G), dim3(RED_BS), 0, c.stream, n, m.d_rowPtr.p, m.d_cols.p, s.A.p, s.B,
                       s.nut, sc, part, G);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(RED_BS), 0, c.stream, part, G, sc + 1);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(RED_BS), 0, c.stream, part + G, G, sc + 2);
    RDST_HIP(hipGetLastError());
    double h[3];
    RDST_HIP(hipMemcpyAsync(h, sc, sizeof(h), hipMemcpyDeviceToHost, c.stream));
    RDST_HIP(hipStreamSynchronize(c.stream));
    const double nf = h[1] + 1.0e-20;
    const double init = h[2] / nf;
    double fin = init;
    int nIter = 0;
    auto conv = [&]() { return fin < t.tolerance || (t.rel_tol > 0 && fin < t.rel_tol * init); };
    if (!conv() || t.min_iter > 0) {
        const size_t shmem = 2 * sizeof(double) * (size_t)m.sa.maxRows;
        double* xa = s.nut;
        double* xb = s.x2;
        do {
            hipLaunchKernelGGL(k_sa_sgs, dim3(m.sa.nParts), dim3(256), shmem, c.stream, m.sa.d_rowStart.p,
                               m.sa.d_levStart.p, m.sa.d_levPtr.p, m.sa.d_rows.p, m.d_rowPtr.p, m.d_cols.p,
                               m.d_slotDiag.p, m.sa.d_localCol.p, s.A.p, s.B, xa, xb);
            std::swap(xa, xb);
            hipLaunchKernelGGL(k_sa_resid_partial, dim3(G), dim3(RED_BS), 0, c.stream, n, m.d_rowPtr.p, m.d_cols.p,
                               s.A.p, s.B, xa, part);
            hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(RED_BS), 0, c.stream, part, G, sc + 3);
            RDST_HIP(hipGetLastError());
            fin = read_scalar(c, sc + 3) / nf;
            ++nIter;
        } while ((nIter < t.max_iter && !conv()) || nIter < t.min_iter);
        if (xa != s.nut) RDST_HIP(hipMemcpyAsync(s.nut, xa, sizeof(double) * n, hipMemcpyDeviceToDevice, c.stream));
    }
    // bound nuTilda (SARC2; SAneg without negativeNuTilda), boundary values
    nut_boundary(c, s.nut);
    if (!isNeg || !t.negative_nu_tilda) {
        if (below(c, s.nut, s.bnut, 0.0)) {
            hipLaunchKernelGGL(k_bound_avg, dim3(nb), dim3(bs), 0, c.stream, md, s.nut, s.bnut, 0.0, s.x2);
            RDST_HIP(hipMemcpyAsync(s.nut, s.x2, sizeof(double) * n, hipMemcpyDeviceToDevice, c.stream));
            nut_boundary(c, s.nut);
            if (m.nB) hipLaunchKernelGGL(k_bnd_floor, dim3(blocks_for(m.nB, 256)), dim3(256), 0, c.stream, m.nB, s.bnut, 0.0);
        }
    }
    // mut = |fv1(chi_old)| nuTilda rho (SAneg), fv1 nuTilda rho (SARC2)
    hipLaunchKernelGGL(k_sa_mut, dim3(nb), dim3(bs), 0, c.stream, n, c.f, sa_dev(c), isNeg ? 1 : 0);
    mut_boundary(c);
    if (isNeg && t.negative_nu_tilda) {
        if (below(c, c.f.mut, c.f.bmut, 0.0)) {
            hipLaunchKernelGGL(k_bound_avg, dim3(nb), dim3(bs), 0, c.stream, md, c.f.mut, c.f.bmut, 0.0, s.x2);
            RDST_HIP(hipMemcpyAsync(c.f.mut, s.x2, sizeof(double) * n, hipMemcpyDeviceToDevice, c.stream));
            mut_boundary(c);
        }
    }
    hipLaunchKernelGGL(k_sa_alphat, dim3(nb), dim3(bs), 0, c.stream, n, c.f, t.Prt);
    RDST_HIP(hipGetLastError());
    s.nCorrect++;
    if (perf) {
        *perf = rdst_solver_perf{};
        perf->initial_residual[0] = init;
        perf->final_residual[0] = fin;
        perf->n_iterations = nIter;
        perf->converged = conv() ? 1 : 0;
    }
}

}  // namespace rdst
