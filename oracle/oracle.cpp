/*
 * oracle.cpp -- CPU FP64 restatement of rhoDST's block-coupled implicit step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it is built
 * into oracle/_build/liboracle.so and is loaded only by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as
 * the checker.  The product path (rhodst_amd/) never links or calls it.
 *
 * It follows the reference field-at-a-time, the way the foam-extend
 * operators evaluate (one whole-field expression after another, branch
 * masks multiplied in rather than branched on), citing each reference
 * file:line it restates.  Paths are relative to the rhoDST tree.
 *
 * Pinning (see tests/test_oracle_*.py): flux consistency F(W,W)=F_phys(W),
 * supersonic upwinding, mirror antisymmetry, discrete conservation,
 * Euler-Jacobian homogeneity A(W)W = F(W).n and central finite
 * differences, closed-form geometry on uniform grids, exact block LU on
 * 1-D meshes against scipy, GMRES/BiCGStab against numpy.linalg.solve,
 * far-field regime closed forms, LTS closed forms.
 *
 * Parity unpinned (no reference run is possible: foam-extend 4.1 is absent):
 *  - thermo T(h) is taken as h/Cp exactly (hConstThermo Newton not restated);
 *  - boundMinMax is restated as a plain clamp;
 *  - the foam-extend BlockGMRES/BlockBiCGStab/ILUC0 internals (normFactor,
 *    convergence test on |residual vector5|, modified Gram-Schmidt,
 *    left-preconditioned GMRES, right-preconditioned BiCGStab) are restated
 *    from the published foam-extend algorithms, not pinned by a run.
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <vector>
#include <string>
#include <algorithm>
#include <stdexcept>

#include "../include/rhodst.h"

namespace orc {

static const double SMALL = 1.0e-15;   // foam doubleScalar SMALL
static const double VSMALL = 1.0e-300; // foam doubleScalar VSMALL
static const double ROOTVSMALL = 1.0e-150;
static const double BLOCK_SMALL = 1.0e-20; // BlockLduMatrix::small_
static const double BLOCK_GREAT = 1.0e20;  // BlockLduMatrix::great_

static std::string g_err;

// ------------------------------------------------------------------ helpers
struct V3 { double x, y, z; };
static inline V3 v3(double x, double y, double z) { V3 r = {x, y, z}; return r; }
static inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 operator*(double s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
static inline V3 operator*(V3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline V3 operator/(V3 a, double s) { return v3(a.x / s, a.y / s, a.z / s); }
static inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double magSqr(V3 a) { return dot(a, a); }
static inline double mag(V3 a) { return std::sqrt(magSqr(a)); }
static inline double cmp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// foam scalar functions
static inline double fsign(double s) { return (s >= 0) ? 1.0 : -1.0; } // foam sign: sign(0)=+1
static inline double sqr(double s) { return s * s; }
// DST::pos/pos0/neg/neg0 (DST.C:323-341)
static inline double pos(double s) { return (s > 0) ? 1.0 : 0.0; }
static inline double pos0(double s) { return (s >= 0) ? 1.0 : 0.0; }
static inline double neg(double s) { return (s < 0) ? 1.0 : 0.0; }
static inline double neg0(double s) { return (s <= 0) ? 1.0 : 0.0; }
// foam-extend's own pos() used by inletOutlet (s >= 0)
static inline double fe_pos(double s) { return (s >= 0) ? 1.0 : 0.0; }

typedef double T25[25];

// ------------------------------------------------------------------ mesh
struct Mesh {
    int nPoints = 0, nFaces = 0, nInternal = 0, nCells = 0;
    std::vector<V3> points;
    std::vector<int> faceOff, facePts, owner, neighbour;
    std::vector<rdst_patch> patches;
    // polyMesh-face geometry
    std::vector<V3> Sf, Cf;
    std::vector<double> magSf;
    std::vector<V3> C;
    std::vector<double> V;
    // fvMesh surface fields: internal faces, then compacted boundary faces
    int nB = 0;
    std::vector<int> bFace;   // polyMesh face of each boundary face
    std::vector<int> bPatch;  // patch of each boundary face
    std::vector<int> bCell;   // faceCells
    std::vector<double> w;    // weights, size nInternal + nB
    std::vector<double> dc;   // deltaCoeffs
    std::vector<V3> corr;     // non-orthogonal correction vectors, internal faces
    std::vector<std::vector<int>> cellCells;
    // faces of each cell in the order the serial face loops reach it
    // (internal faces ascending, then the boundary faces): entry k + 1 where
    // the cell is the owner of surface-field face k (every boundary face),
    // -(k + 1) where it is the neighbour.  A cell-parallel gather over this
    // list adds the same terms in the same order as the serial scatter, so
    // the OpenMP build of the face loop gives bit-identical sums.
    std::vector<int> cfStart, cfIdx;
};

// primitiveMesh::makeFaceCentresAndAreas (foam-extend primitiveMeshFaceCentresAndAreas.C)
static void faceCentresAndAreas(Mesh& m) {
    m.Sf.assign(m.nFaces, v3(0, 0, 0));
    m.Cf.assign(m.nFaces, v3(0, 0, 0));
    m.magSf.assign(m.nFaces, 0.0);
    for (int f = 0; f < m.nFaces; ++f) {
        const int s = m.faceOff[f], n = m.faceOff[f + 1] - s;
        const int* fp = &m.facePts[s];
        if (n == 3) {
            V3 p0 = m.points[fp[0]], p1 = m.points[fp[1]], p2 = m.points[fp[2]];
            m.Cf[f] = (1.0 / 3.0) * (p0 + p1 + p2);
            m.Sf[f] = 0.5 * cross(p1 - p0, p2 - p0);
        } else {
            V3 sumN = v3(0, 0, 0), sumAc = v3(0, 0, 0);
            double sumA = 0.0;
            V3 fCentre = m.points[fp[0]];
            for (int pi = 1; pi < n; ++pi) fCentre = fCentre + m.points[fp[pi]];
            fCentre = fCentre / double(n);
            for (int pi = 0; pi < n; ++pi) {
                V3 p = m.points[fp[pi]];
                V3 nextPoint = m.points[fp[(pi + 1) % n]];
                V3 c = p + nextPoint + fCentre;
                V3 nv = cross(nextPoint - p, fCentre - p);
                double a = mag(nv);
                sumN = sumN + nv;
                sumA += a;
                sumAc = sumAc + a * c;
            }
            if (sumA < ROOTVSMALL) {
                m.Cf[f] = fCentre;
                m.Sf[f] = v3(0, 0, 0);
            } else {
                m.Cf[f] = (1.0 / 3.0) * sumAc / sumA;
                m.Sf[f] = 0.5 * sumN;
            }
        }
        m.magSf[f] = mag(m.Sf[f]);
    }
}

// primitiveMesh::makeCellCentresAndVols (foam-extend primitiveMeshCellCentresAndVols.C)
static void cellCentresAndVols(Mesh& m) {
    std::vector<V3> cEst(m.nCells, v3(0, 0, 0));
    std::vector<int> nCellFaces(m.nCells, 0);
    for (int f = 0; f < m.nFaces; ++f) {
        cEst[m.owner[f]] = cEst[m.owner[f]] + m.Cf[f];
        nCellFaces[m.owner[f]]++;
    }
    for (int f = 0; f < m.nInternal; ++f) {
        cEst[m.neighbour[f]] = cEst[m.neighbour[f]] + m.Cf[f];
        nCellFaces[m.neighbour[f]]++;
    }
    for (int c = 0; c < m.nCells; ++c) cEst[c] = cEst[c] / double(nCellFaces[c]);
    m.C.assign(m.nCells, v3(0, 0, 0));
    m.V.assign(m.nCells, 0.0);
    for (int f = 0; f < m.nFaces; ++f) {
        int o = m.owner[f];
        double pyr3Vol = std::max(dot(m.Sf[f], m.Cf[f] - cEst[o]), VSMALL);
        V3 pc = (3.0 / 4.0) * m.Cf[f] + (1.0 / 4.0) * cEst[o];
        m.C[o] = m.C[o] + pyr3Vol * pc;
        m.V[o] += pyr3Vol;
    }
    for (int f = 0; f < m.nInternal; ++f) {
        int nb = m.neighbour[f];
        double pyr3Vol = std::max(dot(m.Sf[f], cEst[nb] - m.Cf[f]), VSMALL);
        V3 pc = (3.0 / 4.0) * m.Cf[f] + (1.0 / 4.0) * cEst[nb];
        m.C[nb] = m.C[nb] + pyr3Vol * pc;
        m.V[nb] += pyr3Vol;
    }
    for (int c = 0; c < m.nCells; ++c) {
        m.C[c] = m.C[c] / m.V[c];
        m.V[c] *= (1.0 / 3.0);
    }
}

static void buildMesh(Mesh& m, const rdst_mesh_desc* d) {
    m.nPoints = d->n_points; m.nFaces = d->n_faces; m.nInternal = d->n_internal_faces;
    m.nCells = d->n_cells;
    m.points.resize(m.nPoints);
    for (int i = 0; i < m.nPoints; ++i)
        m.points[i] = v3(d->points[3 * i], d->points[3 * i + 1], d->points[3 * i + 2]);
    m.faceOff.assign(d->face_offsets, d->face_offsets + m.nFaces + 1);
    m.facePts.assign(d->face_points, d->face_points + m.faceOff[m.nFaces]);
    m.owner.assign(d->owner, d->owner + m.nFaces);
    m.neighbour.assign(d->neighbour, d->neighbour + m.nInternal);
    m.patches.assign(d->patches, d->patches + d->n_patches);
    faceCentresAndAreas(m);
    cellCentresAndVols(m);
    // compacted boundary faces: fvPatch of an empty patch has size 0
    m.bFace.clear(); m.bPatch.clear(); m.bCell.clear();
    for (int p = 0; p < (int)m.patches.size(); ++p) {
        if (m.patches[p].type == RDST_PATCH_EMPTY) continue;
        for (int i = 0; i < m.patches[p].size; ++i) {
            int f = m.patches[p].start + i;
            m.bFace.push_back(f); m.bPatch.push_back(p); m.bCell.push_back(m.owner[f]);
        }
    }
    m.nB = (int)m.bFace.size();
    // surfaceInterpolation::makeWeights / makeDeltaCoeffs (foam-extend 4.1)
    m.w.assign(m.nInternal + m.nB, 1.0);
    m.dc.assign(m.nInternal + m.nB, 0.0);
    for (int f = 0; f < m.nInternal; ++f) {
        double dOwn = std::fabs(dot(m.Sf[f], m.Cf[f] - m.C[m.owner[f]]));
        double dNei = std::fabs(dot(m.Sf[f], m.C[m.neighbour[f]] - m.Cf[f]));
        m.w[f] = dNei / (dOwn + dNei);
        V3 delta = m.C[m.neighbour[f]] - m.C[m.owner[f]];
        V3 unitArea = m.Sf[f] / m.magSf[f];
        // stabilised form, comment at timeDST.C:237
        m.dc[f] = 1.0 / std::max(dot(unitArea, delta), 0.05 * mag(delta));
    }
    // correction vectors n - d deltaCoeffs (boundary: 0 on non-coupled patches)
    m.corr.assign(m.nInternal, v3(0, 0, 0));
    for (int f = 0; f < m.nInternal; ++f) {
        V3 delta = m.C[m.neighbour[f]] - m.C[m.owner[f]];
        V3 unitArea = m.Sf[f] / m.magSf[f];
        m.corr[f] = unitArea - delta * m.dc[f];
    }
    for (int b = 0; b < m.nB; ++b) {
        int f = m.bFace[b];
        V3 delta = m.Cf[f] - m.C[m.bCell[b]];
        m.dc[m.nInternal + b] = 1.0 / mag(delta);
    }
    m.cfStart.assign(m.nCells + 1, 0);
    for (int f = 0; f < m.nInternal; ++f) { m.cfStart[m.owner[f] + 1]++; m.cfStart[m.neighbour[f] + 1]++; }
    for (int b = 0; b < m.nB; ++b) m.cfStart[m.bCell[b] + 1]++;
    for (int i = 0; i < m.nCells; ++i) m.cfStart[i + 1] += m.cfStart[i];
    m.cfIdx.assign(m.cfStart[m.nCells], 0);
    {
        std::vector<int> fill(m.cfStart.begin(), m.cfStart.end() - 1);
        for (int f = 0; f < m.nInternal; ++f) {
            m.cfIdx[fill[m.owner[f]]++] = f + 1;
            m.cfIdx[fill[m.neighbour[f]]++] = -(f + 1);
        }
        for (int b = 0; b < m.nB; ++b) m.cfIdx[fill[m.bCell[b]]++] = m.nInternal + b + 1;
    }
    // mesh.cellCells(): neighbours across internal faces
    m.cellCells.assign(m.nCells, std::vector<int>());
    for (int f = 0; f < m.nInternal; ++f) {
        m.cellCells[m.owner[f]].push_back(m.neighbour[f]);
        m.cellCells[m.neighbour[f]].push_back(m.owner[f]);
    }
}

// surface-field accessors: face k < nInternal internal, else boundary k-nInternal
static inline V3 SfOf(const Mesh& m, int k) { return k < m.nInternal ? m.Sf[k] : m.Sf[m.bFace[k - m.nInternal]]; }
static inline double magSfOf(const Mesh& m, int k) { return k < m.nInternal ? m.magSf[k] : m.magSf[m.bFace[k - m.nInternal]]; }

// ------------------------------------------------------------------ case
struct BCState {
    // per boundary face mixed-BC data (valueFraction, U refValue) kept between steps
    std::vector<double> vf_p, vf_T, vf_U;
    std::vector<V3> refU;
};

struct T9 { double a[9]; };

// Spalart-Allmaras state (SAneg / SARC2 members nuTilda_, mut_, alphat_, d_)
struct Turb {
    rdst_turbulence t{};
    bool on = false;
    std::vector<rdst_scalar_bc> bc;          // 0/nuTilda per patch
    std::vector<double> d;                   // wall distance, cells
    std::vector<double> nut, bnut;           // nuTilda: cells, boundary faces
    std::vector<double> mut, bmut, alphat, balphat;
    std::vector<double> rhoPrev;             // rho before this step's updateFields
    bool rhoPrevValid = false;
    std::vector<T9> Sold;                    // SARC1: S = symm(grad U) of the previous correct()
    int nCorrect = 0;
    // nuTilda matrix after relax(), solve form: diagonal incl. internalCoeffs,
    // source incl. boundaryCoeffs
    std::vector<double> Ad, Au, Al, As;
};

struct Case {
    Mesh m;
    rdst_thermo th;
    std::vector<rdst_bc> bc;
    rdst_settings st;
    double Cv = 0, gamma = 0;

    // volume fields (cells)
    std::vector<double> rho, rhoE, p, T, h, psi, mu, alpha;
    std::vector<V3> rhoU, U;
    // boundary values, per compacted boundary face
    std::vector<double> brho, brhoE, bp, bT, bh, bpsi, bmu, balpha;
    std::vector<V3> brhoU, bU;
    BCState bcs;
    // surface fields, nInternal + nB
    std::vector<double> phi, phiEp;
    std::vector<V3> phiUp, Up;
    // residual and solution vector
    std::vector<double> R;  // [n][5]
    std::vector<double> W;  // [n][5]
    // Jacobian
    std::vector<double> ownC, neiC;  // [nInternal][25]
    std::vector<double> lambda;      // [nInternal] spectral radius
    std::vector<double> patchIC;     // [nB][25]
    // assembled LDU
    std::vector<double> diag, upper, lower;  // [..][25]
    std::vector<double> source;              // [n][5]
    // time control
    std::vector<double> rDeltaT, courant, rhoOld, eOld, dS;  // dS [n][3]
    double residual = -1, residualRatio = 1, courantScalar = 1;
    // dual time stepping: old time levels of W [n][5], physical time control
    std::vector<double> WOld, WOldOld;
    double time = 0, deltaT = 0, deltaT0 = 0, deltaTSave = 0;
    int timeIndex = 0;
    bool hasState = false, hasMatrix = false;
    // cell partition for the block-Jacobi ILU(0) (empty: one partition)
    std::vector<int> part;
    // cell partition of the nuTilda Gauss-Seidel sweep (empty: the serial sweep)
    std::vector<int> saPart;
    Turb tb;
};

// muEff = mut + mu, alphaEff = alphat + alpha (RASModel::muEff/alphaEff);
// laminar: mut = alphat = 0
static void muAlphaEff(const Case& c, std::vector<double>& mu, std::vector<double>& bmu, std::vector<double>* al,
                       std::vector<double>* bal) {
    mu = c.mu; bmu = c.bmu;
    if (al) *al = c.alpha;
    if (bal) *bal = c.balpha;
    if (!c.tb.on) return;
    for (size_t i = 0; i < mu.size(); ++i) mu[i] = c.tb.mut[i] + c.mu[i];
    for (size_t b = 0; b < bmu.size(); ++b) bmu[b] = c.tb.bmut[b] + c.bmu[b];
    if (al) for (size_t i = 0; i < al->size(); ++i) (*al)[i] = c.tb.alphat[i] + c.alpha[i];
    if (bal) for (size_t b = 0; b < bal->size(); ++b) (*bal)[b] = c.tb.balphat[b] + c.balpha[b];
}

// ------------------------------------------------------------------ thermo
static inline double muOf(const Case& c, double T) {
    if (c.th.transport == 1) return c.th.As * std::sqrt(T) / (1.0 + c.th.Ts / T);
    return c.th.mu;
}
static inline double alphaOf(const Case& c, double T) {
    // constTransport alpha reduces to mu/Pr for hConstThermo; sutherland uses
    // the modified Eucken conductivity kappa = mu*Cv*(1.32 + 1.77*R/Cv), alpha = kappa/Cp
    if (c.th.transport == 1) {
        double mu = muOf(c, T);
        return mu * c.Cv * (1.32 + 1.77 * c.th.R / c.Cv) / c.th.Cp;
    }
    return c.th.mu / c.th.Pr;
}

// boundary value of a vector field on a symmetryPlane/slip patch:
// (transform(I - 2 n n, U_I) + U_I)/2   (basicSymmetryFvPatchField::evaluate)
static inline V3 symmetricValue(V3 n, V3 UI) {
    double nU = dot(n, UI);
    V3 t = v3(UI.x - 2.0 * n.x * nU, UI.y - 2.0 * n.y * nU, UI.z - 2.0 * n.z * nU);
    return (t + UI) / 2.0;
}

// valueInternalCoeffs of each BC (Roe.C:363-374 reads them)
static inline double scalarVIC(int type, double vf) {
    switch (type) {
        case RDST_BC_FIXED_VALUE: return 0.0;
        case RDST_BC_ZERO_GRADIENT: return 1.0;
        case RDST_BC_SYMMETRY: return 1.0;
        case RDST_BC_INLET_OUTLET:
        case RDST_BC_FARFIELD: return 1.0 - vf;
        default: return 0.0;
    }
}
static inline V3 vectorVIC(int type, double vf, V3 n) {
    switch (type) {
        case RDST_BC_FIXED_VALUE: return v3(0, 0, 0);
        case RDST_BC_ZERO_GRADIENT: return v3(1, 1, 1);
        case RDST_BC_SYMMETRY: return v3(1.0 - std::fabs(n.x), 1.0 - std::fabs(n.y), 1.0 - std::fabs(n.z));
        case RDST_BC_INLET_OUTLET:
        case RDST_BC_FARFIELD: return v3(1.0 - vf, 1.0 - vf, 1.0 - vf);
        default: return v3(0, 0, 0);
    }
}

// ------------------------------------------------------------------ set state
// createFields.H:21-72 and compressibleCreatePhi; timeDST::initialize (timeDST.C:67-208)
static void setSteadyStateDeltaT(Case& c);
static void setTransientDeltaT(Case& c);

static void initialBoundary(Case& c) {
    const Mesh& m = c.m;
    int nB = m.nB;
    c.bp.assign(nB, 0); c.bT.assign(nB, 0); c.bU.assign(nB, v3(0, 0, 0));
    c.bcs.vf_p.assign(nB, 0.0); c.bcs.vf_T.assign(nB, 0.0); c.bcs.vf_U.assign(nB, 0.0);
    c.bcs.refU.assign(nB, v3(0, 0, 0));
    for (int b = 0; b < nB; ++b) {
        const rdst_bc& bc = c.bc[m.bPatch[b]];
        int ci = m.bCell[b];
        V3 n = m.Sf[m.bFace[b]] / m.magSf[m.bFace[b]];
        switch (bc.p_type) {
            case RDST_BC_FIXED_VALUE: case RDST_BC_INLET_OUTLET: c.bp[b] = bc.p_value; break;
            case RDST_BC_FARFIELD: c.bp[b] = bc.p_inf; break;
            default: c.bp[b] = c.p[ci];
        }
        switch (bc.T_type) {
            case RDST_BC_FIXED_VALUE: case RDST_BC_INLET_OUTLET: c.bT[b] = bc.T_value; break;
            case RDST_BC_FARFIELD: c.bT[b] = bc.T_inf; break;
            default: c.bT[b] = c.T[ci];
        }
        switch (bc.U_type) {
            case RDST_BC_FIXED_VALUE: case RDST_BC_INLET_OUTLET:
                c.bU[b] = v3(bc.U_value[0], bc.U_value[1], bc.U_value[2]); break;
            case RDST_BC_FARFIELD: c.bU[b] = v3(bc.U_inf[0], bc.U_inf[1], bc.U_inf[2]); break;
            case RDST_BC_SYMMETRY: c.bU[b] = symmetricValue(n, c.U[ci]); break;
            default: c.bU[b] = c.U[ci];
        }
    }
}

static void setState(Case& c, const double* p, const double* T, const double* U) {
    const Mesh& m = c.m;
    int n = m.nCells, nB = m.nB;
    c.p.assign(p, p + n); c.T.assign(T, T + n);
    c.U.resize(n);
    for (int i = 0; i < n; ++i) c.U[i] = v3(U[3 * i], U[3 * i + 1], U[3 * i + 2]);
    c.psi.resize(n); c.h.resize(n); c.mu.resize(n); c.alpha.resize(n);
    for (int i = 0; i < n; ++i) {
        c.psi[i] = 1.0 / (c.th.R * c.T[i]);
        c.h[i] = c.th.Cp * c.T[i];
        c.mu[i] = muOf(c, c.T[i]);
        c.alpha[i] = alphaOf(c, c.T[i]);
    }
    initialBoundary(c);
    c.bpsi.resize(nB); c.bh.resize(nB); c.bmu.resize(nB); c.balpha.resize(nB);
    for (int b = 0; b < nB; ++b) {
        c.bpsi[b] = 1.0 / (c.th.R * c.bT[b]);
        c.bh[b] = c.th.Cp * c.bT[b];
        c.bmu[b] = muOf(c, c.bT[b]);
        c.balpha[b] = alphaOf(c, c.bT[b]);
    }
    // rho = thermo.rho() = psi*p; rhoU = rho*U; rhoE = rho*(h*Cv/Cp + 0.5|U|^2)
    c.rho.resize(n); c.rhoU.resize(n); c.rhoE.resize(n);
    for (int i = 0; i < n; ++i) {
        c.rho[i] = c.psi[i] * c.p[i];
        c.rhoU[i] = c.rho[i] * c.U[i];
        c.rhoE[i] = c.rho[i] * (c.h[i] * c.Cv / c.th.Cp + 0.5 * magSqr(c.U[i]));
    }
    c.brho.resize(nB); c.brhoU.resize(nB); c.brhoE.resize(nB);
    for (int b = 0; b < nB; ++b) {
        c.brho[b] = c.bpsi[b] * c.bp[b];
        c.brhoU[b] = c.brho[b] * c.bU[b];
        c.brhoE[b] = c.brho[b] * (c.bh[b] * c.Cv / c.th.Cp + 0.5 * magSqr(c.bU[b]));
    }
    // phi = linearInterpolate(rho*U) & Sf
    c.phi.assign(m.nInternal + nB, 0.0);
    for (int f = 0; f < m.nInternal; ++f) {
        int o = m.owner[f], nb = m.neighbour[f];
        double w = m.w[f];
        V3 rUo = c.rho[o] * c.U[o], rUn = c.rho[nb] * c.U[nb];
        V3 rUf = v3(w * (rUo.x - rUn.x) + rUn.x, w * (rUo.y - rUn.y) + rUn.y, w * (rUo.z - rUn.z) + rUn.z);
        c.phi[f] = dot(rUf, m.Sf[f]);
    }
    for (int b = 0; b < nB; ++b) c.phi[m.nInternal + b] = dot(c.brho[b] * c.bU[b], m.Sf[m.bFace[b]]);
    c.phiUp.assign(m.nInternal + nB, v3(0, 0, 0));
    c.phiEp.assign(m.nInternal + nB, 0.0);
    c.Up.assign(m.nInternal + nB, v3(0, 0, 0));
    c.R.assign(5 * n, 0.0);
    // W.set(rho, rhoU, rhoE)
    c.W.resize(5 * n);
    for (int i = 0; i < n; ++i) {
        c.W[5 * i] = c.rho[i];
        c.W[5 * i + 1] = c.rhoU[i].x; c.W[5 * i + 2] = c.rhoU[i].y; c.W[5 * i + 3] = c.rhoU[i].z;
        c.W[5 * i + 4] = c.rhoE[i];
    }
    // timeDST::initialize
    c.rhoOld = c.rho;   // thermo rho
    c.eOld.resize(n);
    for (int i = 0; i < n; ++i) c.eOld[i] = c.h[i] * c.Cv / c.th.Cp;
    if (!c.st.neighbouring) {
        c.dS.assign(3 * n, 0.0);
        for (int f = 0; f < m.nInternal; ++f) {
            V3 a = v3(std::fabs(m.Sf[f].x), std::fabs(m.Sf[f].y), std::fabs(m.Sf[f].z));
            int o = m.owner[f], nb = m.neighbour[f];
            c.dS[3 * o] += a.x; c.dS[3 * o + 1] += a.y; c.dS[3 * o + 2] += a.z;
            c.dS[3 * nb] += a.x; c.dS[3 * nb + 1] += a.y; c.dS[3 * nb + 2] += a.z;
        }
        for (int b = 0; b < nB; ++b) {
            V3 sf = m.Sf[m.bFace[b]];
            int ci = m.bCell[b];
            c.dS[3 * ci] += std::fabs(sf.x); c.dS[3 * ci + 1] += std::fabs(sf.y); c.dS[3 * ci + 2] += std::fabs(sf.z);
        }
        for (auto& v : c.dS) v *= 0.5;
    }
    c.residual = -1; c.residualRatio = 1;
    c.courantScalar = c.st.initial_co;
    c.courant.assign(n, c.st.initial_co);
    c.rDeltaT.assign(n, 0.0);
    c.time = 0; c.timeIndex = 0;
    c.WOld.clear(); c.WOldOld.clear();
    if (c.st.time_integration == RDST_TIME_DUAL) {
        // timeDST::initialize (timeDST.C:193-207): the physical deltaT from the
        // initial courant, then the pseudo time step
        setTransientDeltaT(c);
        c.deltaT0 = c.deltaTSave = c.st.delta_t > 0 ? c.st.delta_t : c.deltaT;
    }
    setSteadyStateDeltaT(c);
    c.hasState = true; c.hasMatrix = false;
}

// ------------------------------------------------------------------ flux
// Face states used by every scheme: upwind reconstruction with own_=+1 and
// nei_=-1 (baseFluxDST.C:144-171) gives the owner cell value on the owner
// side and the neighbour cell value on the neighbour side; boundary faces
// take the patch value on both sides.
struct FaceStates {
    int nf;
    std::vector<double> rho_o, rho_n, rPsi_o, rPsi_n, e_o, e_n, a_o, a_n, gam_f, p_cellf;
    std::vector<V3> rhoU_o, rhoU_n, Sf;
    std::vector<double> magSf;
};

// cell-centred auxiliary fields with boundary values
struct CellAux {
    std::vector<double> rPsi, e, aStar, c, gam;  // cells
    std::vector<double> brPsi, be, baStar, bc, bgam;
};

static void makeAux(const Case& c, CellAux& a) {
    const Mesh& m = c.m;
    int n = m.nCells, nB = m.nB;
    double Cp = c.th.Cp, Cv = c.Cv;
    a.rPsi.resize(n); a.e.resize(n); a.aStar.resize(n); a.c.resize(n); a.gam.resize(n);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        a.rPsi[i] = 1.0 / c.psi[i];
        a.e[i] = c.h[i] - c.p[i] / c.rho[i];                       // AUSMPLDST.C:145
        double H = std::max(c.rhoE[i] / c.rho[i], SMALL) + std::max(c.p[i] / c.rho[i], SMALL);  // :156-160
        a.gam[i] = Cp / Cv;
        a.aStar[i] = std::sqrt(2.0 * (a.gam[i] - 1.0) / (a.gam[i] + 1.0) * H);  // :161
        a.c[i] = std::sqrt(a.gam[i] * a.rPsi[i]);                   // KurganovDST.C: sqrt(Cp/Cv*rPsi)
    }
    a.brPsi.resize(nB); a.be.resize(nB); a.baStar.resize(nB); a.bc.resize(nB); a.bgam.resize(nB);
    for (int b = 0; b < nB; ++b) {
        a.brPsi[b] = 1.0 / c.bpsi[b];
        a.be[b] = c.bh[b] - c.bp[b] / c.brho[b];
        double H = std::max(c.brhoE[b] / c.brho[b], SMALL) + std::max(c.bp[b] / c.brho[b], SMALL);
        a.bgam[b] = Cp / Cv;
        a.baStar[b] = std::sqrt(2.0 * (a.bgam[b] - 1.0) / (a.bgam[b] + 1.0) * H);
        a.bc[b] = std::sqrt(a.bgam[b] * a.brPsi[b]);
    }
}

// upwind-interpolated face value of a cell field (own side / nei side)
template <class T>
static void reconstruct(const Mesh& m, const std::vector<T>& cell, const std::vector<T>& bnd,
                        std::vector<T>& own, std::vector<T>& nei) {
    int nf = m.nInternal + m.nB;
    own.resize(nf); nei.resize(nf);
#pragma omp parallel for schedule(static)
    for (int f = 0; f < m.nInternal; ++f) { own[f] = cell[m.owner[f]]; nei[f] = cell[m.neighbour[f]]; }
    for (int b = 0; b < m.nB; ++b) { own[m.nInternal + b] = bnd[b]; nei[m.nInternal + b] = bnd[b]; }
}
// limitedScheme OSPRE / OSPREV (foam LimitedScheme + NVDTVD / NVDVTVDV r-factor):
// face value lambda (P - N) + N with lambda = lim w + (1 - lim) pos(faceFlux),
// faceFlux = +1 on the owner side (own_) and -1 on the neighbour side (nei_);
// r = 2 (d & grad_upwind)/(phiN - phiP) - 1 (scalars; for vectors both sides
// projected on phiN - phiP), |d.grad| >= 1000 |phiN - phiP| capped, d = C_N - C_P;
// OSPRE lim(r) = 1.5 r(r + 1)/(r(r + 1) + 1).  Boundary faces take the patch value.
static inline double ospre(double r) {
    double rrp1 = r * (r + 1.0);
    return 1.5 * rrp1 / (rrp1 + 1.0);
}
static inline double rFactor(double gradcf, double gradf) {
    if (std::fabs(gradcf) >= 1000.0 * std::fabs(gradf)) return 2.0 * 1000.0 * fsign(gradcf) * fsign(gradf) - 1.0;
    return 2.0 * (gradcf / gradf) - 1.0;
}
static void gradScalarB(const Mesh& m, const std::vector<double>& v, const std::vector<double>& bv,
                        std::vector<V3>& g);
static void reconstructOSPRE(const Mesh& m, const std::vector<double>& cell, const std::vector<double>& bnd,
                             std::vector<double>& own, std::vector<double>& nei) {
    std::vector<V3> g;
    gradScalarB(m, cell, bnd, g);
    int nf = m.nInternal + m.nB;
    own.resize(nf); nei.resize(nf);
    for (int f = 0; f < m.nInternal; ++f) {
        int o = m.owner[f], nb = m.neighbour[f];
        double P = cell[o], N = cell[nb], w = m.w[f];
        V3 d = m.C[nb] - m.C[o];
        double gradf = N - P;
        double lo = ospre(rFactor(dot(d, g[o]), gradf));
        double ln = ospre(rFactor(dot(d, g[nb]), gradf));
        double wo = lo * w + (1.0 - lo) * 1.0;
        double wn = ln * w + (1.0 - ln) * 0.0;
        own[f] = wo * (P - N) + N;
        nei[f] = wn * (P - N) + N;
    }
    for (int b = 0; b < m.nB; ++b) { own[m.nInternal + b] = bnd[b]; nei[m.nInternal + b] = bnd[b]; }
}
static void reconstructOSPREV(const Mesh& m, const std::vector<V3>& cell, const std::vector<V3>& bnd,
                              std::vector<V3>& own, std::vector<V3>& nei) {
    // Gauss linear gradient of a vector field, G[3i+j] = d phi_j / dx_i
    int n = m.nCells;
    std::vector<double> gc(9 * (size_t)n, 0.0);
    for (int f = 0; f < m.nInternal; ++f) {
        int o = m.owner[f], nb = m.neighbour[f];
        double w = m.w[f];
        V3 P = cell[o], N = cell[nb];
        V3 vf = v3(w * (P.x - N.x) + N.x, w * (P.y - N.y) + N.y, w * (P.z - N.z) + N.z);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double t = cmp(m.Sf[f], i) * cmp(vf, j);
                gc[9 * (size_t)o + 3 * i + j] += t;
                gc[9 * (size_t)nb + 3 * i + j] -= t;
            }
    }
    for (int b = 0; b < m.nB; ++b) {
        int ci = m.bCell[b];
        V3 sf = m.Sf[m.bFace[b]];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) gc[9 * (size_t)ci + 3 * i + j] += cmp(sf, i) * cmp(bnd[b], j);
    }
    for (int c = 0; c < n; ++c)
        for (int k = 0; k < 9; ++k) gc[9 * (size_t)c + k] /= m.V[c];
    auto dG = [&](V3 d, int c) {  // d & grad
        const double* G = &gc[9 * (size_t)c];
        return v3(d.x * G[0] + d.y * G[3] + d.z * G[6], d.x * G[1] + d.y * G[4] + d.z * G[7],
                  d.x * G[2] + d.y * G[5] + d.z * G[8]);
    };
    int nf = m.nInternal + m.nB;
    own.resize(nf); nei.resize(nf);
    for (int f = 0; f < m.nInternal; ++f) {
        int o = m.owner[f], nb = m.neighbour[f];
        V3 P = cell[o], N = cell[nb];
        double w = m.w[f];
        V3 d = m.C[nb] - m.C[o];
        V3 gradfV = N - P;
        double gradf = dot(gradfV, gradfV);
        double lo = ospre(rFactor(dot(gradfV, dG(d, o)), gradf));
        double ln = ospre(rFactor(dot(gradfV, dG(d, nb)), gradf));
        double wo = lo * w + (1.0 - lo) * 1.0;
        double wn = ln * w + (1.0 - ln) * 0.0;
        own[f] = v3(wo * (P.x - N.x) + N.x, wo * (P.y - N.y) + N.y, wo * (P.z - N.z) + N.z);
        nei[f] = v3(wn * (P.x - N.x) + N.x, wn * (P.y - N.y) + N.y, wn * (P.z - N.z) + N.z);
    }
    for (int b = 0; b < m.nB; ++b) { own[m.nInternal + b] = bnd[b]; nei[m.nInternal + b] = bnd[b]; }
}

// linear interpolation w*(P-N)+N; boundary value on boundary faces
static void linearInterp(const Mesh& m, const std::vector<double>& cell, const std::vector<double>& bnd,
                         std::vector<double>& out) {
    int nf = m.nInternal + m.nB;
    out.resize(nf);
#pragma omp parallel for schedule(static)
    for (int f = 0; f < m.nInternal; ++f) {
        double P = cell[m.owner[f]], N = cell[m.neighbour[f]];
        out[f] = m.w[f] * (P - N) + N;
    }
    for (int b = 0; b < m.nB; ++b) out[m.nInternal + b] = bnd[b];
}
static void linearInterpV(const Mesh& m, const std::vector<V3>& cell, const std::vector<V3>& bnd,
                          std::vector<V3>& out) {
    int nf = m.nInternal + m.nB;
    out.resize(nf);
    for (int f = 0; f < m.nInternal; ++f) {
        V3 P = cell[m.owner[f]], N = cell[m.neighbour[f]];
        double w = m.w[f];
        out[f] = v3(w * (P.x - N.x) + N.x, w * (P.y - N.y) + N.y, w * (P.z - N.z) + N.z);
    }
    for (int b = 0; b < m.nB; ++b) out[m.nInternal + b] = bnd[b];
}

// The six schemes, evaluated face by face from the reconstructed states.
// Each writes phi, phiUp, phiEp, Up for face k.
struct FaceIn {
    double rho_o, rho_n, rPsi_o, rPsi_n, e_o, e_n, a_o, a_n, c_o, c_n, gam_f, fHLLCP;
    V3 rhoU_o, rhoU_n, Sf;
    double magSf;
};
struct FaceOut { double phi, phiE; V3 phiU, Up; };

// AUSMPLDST::computeFluxDST (AUSMPLDST.C:127-217), alpha=3/16, beta=1/8 (AUSMPLDST.H:78-79)
static FaceOut fluxAUSMplus(const FaceIn& in) {
    const double alpha_ = 3.0 / 16.0, beta_ = 0.125;
    V3 U_own = in.rhoU_o / in.rho_o, U_nei = in.rhoU_n / in.rho_n;
    V3 nrm = in.Sf / in.magSf;
    double Uv_own = dot(U_own, nrm), Uv_nei = dot(U_nei, nrm);
    double rhoE_own = in.rho_o * (in.e_o + 0.5 * magSqr(U_own));
    double rhoE_nei = in.rho_n * (in.e_n + 0.5 * magSqr(U_nei));
    double p_own = in.rho_o * in.rPsi_o, p_nei = in.rho_n * in.rPsi_n;
    double a12 = 0.5 * (in.a_o + in.a_n);
    double Ma_own = Uv_own / a12, Ma_nei = Uv_nei / a12;
    double s1o = fsign(1.0 - std::fabs(Ma_own)), s2o = fsign(-1.0 + std::fabs(Ma_own));
    double s1n = fsign(1.0 - std::fabs(Ma_nei)), s2n = fsign(-1.0 + std::fabs(Ma_nei));
    double m1o = 0.5 * (s1o + std::fabs(s1o)), m2o = 0.5 * (s2o + std::fabs(s2o));
    double m1n = 0.5 * (s1n + std::fabs(s1n)), m2n = 0.5 * (s2n + std::fabs(s2n));
    double Ma4_own = m1o * (0.25 * sqr(Ma_own + 1.0)) * (1.0 - 16.0 * beta_ * (-0.25) * sqr(Ma_own - 1.0))
                   + m2o * std::max(Ma_own, 0.0);
    double P5_own = m1o * 0.25 * sqr(Ma_own + 1.0) * (2.0 - Ma_own - 16.0 * alpha_ * Ma_own * (-0.25) * sqr(Ma_own - 1.0))
                  + m2o * pos(Ma_own);
    double Ma4_nei = m1n * (-0.25) * sqr(Ma_nei - 1.0) * (1.0 + 16.0 * beta_ * 0.25 * sqr(Ma_nei + 1.0))
                   + m2n * std::min(Ma_nei, 0.0);
    double P5_nei = m1n * (-0.25) * sqr(Ma_nei - 1.0) * (-2.0 - Ma_nei + 16.0 * alpha_ * Ma_nei * 0.25 * sqr(Ma_nei + 1.0))
                  + m2n * neg(Ma_nei);
    double Ma12 = Ma4_own + Ma4_nei;
    double P12 = P5_own * p_own + P5_nei * p_nei;
    FaceOut o;
    double ps = pos(Ma12), ng = neg0(Ma12);
    o.Up = ps * U_own + ng * U_nei;
    o.phi = Ma12 * a12 * (ps * in.rho_o + ng * in.rho_n) * in.magSf;
    o.phiU = o.phi * o.Up + P12 * in.Sf;
    double H_own = (rhoE_own + p_own) / in.rho_o, H_nei = (rhoE_nei + p_nei) / in.rho_n;
    o.phiE = o.phi * (ps * H_own + ng * H_nei);
    return o;
}

// AUSMPLUPDST::computeFluxDST (AUSMPLUPDST.C:209-320); Kp .25, Ku .75, sigma 1 (AUSMPLUPDST.H:84-87)
static FaceOut fluxAUSMplusUp(const FaceIn& in, bool lowMach, double Mainf) {
    const double beta_ = 0.125, Kp_ = 0.25, Ku_ = 0.75, sigma_ = 1.0;
    V3 U_own = in.rhoU_o / in.rho_o, U_nei = in.rhoU_n / in.rho_n;
    V3 nrm = in.Sf / in.magSf;
    double Uv_own = dot(U_own, nrm), Uv_nei = dot(U_nei, nrm);
    double rhoE_own = in.rho_o * (in.e_o + 0.5 * magSqr(U_own));
    double rhoE_nei = in.rho_n * (in.e_n + 0.5 * magSqr(U_nei));
    double p_own = in.rho_o * in.rPsi_o, p_nei = in.rho_n * in.rPsi_n;
    double a_own = sqr(in.a_o) / std::max(in.a_o, Uv_own);      // :249
    double a_nei = sqr(in.a_n) / std::max(in.a_n, -Uv_nei);     // :250
    double a12 = std::min(a_own, a_nei);                         // :252
    double Ma_own = Uv_own / a12, Ma_nei = Uv_nei / a12;
    double alpha = 0.0, Mmean2 = 0.0, fa = 0.0;                  // alpha stays 0 without lowMach (:118-131)
    if (lowMach) {
        Mmean2 = 0.5 * (sqr(Uv_own) + sqr(Uv_nei)) / sqr(a12);
        double M0 = std::sqrt(std::min(1.0, std::max(Mmean2, sqr(Mainf))));
        fa = M0 * (2.0 - M0);
        alpha = 3.0 / 16.0 * (-4.0 + 5.0 * sqr(fa));
    }
    double s1o = fsign(1.0 - std::fabs(Ma_own)), s2o = fsign(-1.0 + std::fabs(Ma_own));
    double s1n = fsign(1.0 - std::fabs(Ma_nei)), s2n = fsign(-1.0 + std::fabs(Ma_nei));
    double m1o = 0.5 * (s1o + std::fabs(s1o)), m2o = 0.5 * (s2o + std::fabs(s2o));
    double m1n = 0.5 * (s1n + std::fabs(s1n)), m2n = 0.5 * (s2n + std::fabs(s2n));
    double Ma4_own = m1o * (0.25 * sqr(Ma_own + 1.0)) * (1.0 - 16.0 * beta_ * (-0.25) * sqr(Ma_own - 1.0))
                   + m2o * std::max(Ma_own, 0.0);
    double P5_own = m1o * 0.25 * sqr(Ma_own + 1.0) * (2.0 - Ma_own - 16.0 * alpha * Ma_own * (-0.25) * sqr(Ma_own - 1.0))
                  + m2o * pos(Ma_own);
    double Ma4_nei = m1n * (-0.25) * sqr(Ma_nei - 1.0) * (1.0 + 16.0 * beta_ * 0.25 * sqr(Ma_nei + 1.0))
                   + m2n * std::min(Ma_nei, 0.0);
    double P5_nei = m1n * (-0.25) * sqr(Ma_nei - 1.0) * (-2.0 - Ma_nei + 16.0 * alpha * Ma_nei * 0.25 * sqr(Ma_nei + 1.0))
                  + m2n * neg(Ma_nei);
    double Ma12 = Ma4_own + Ma4_nei;
    double P12 = P5_own * p_own + P5_nei * p_nei;
    if (lowMach) {
        double Mp = -Kp_ / fa * std::max(1.0 - sigma_ * Mmean2, 0.0) * (p_nei - p_own) / (0.5 * (in.rho_o + in.rho_n)) / sqr(a12);
        double pu = -Ku_ * P5_own * P5_nei * (in.rho_o + in.rho_n) * a12 * fa * (Uv_own - Uv_nei);
        Ma12 += Mp;
        P12 += pu;
    }
    FaceOut o;
    double ps = pos(Ma12), ng = neg0(Ma12);
    o.Up = ps * U_own + ng * U_nei;
    o.phi = Ma12 * a12 * (ps * in.rho_o + ng * in.rho_n) * in.magSf;
    o.phiU = o.phi * o.Up + P12 * in.Sf;
    double H_own = (rhoE_own + p_own) / in.rho_o, H_nei = (rhoE_nei + p_nei) / in.rho_n;
    o.phiE = o.phi * (ps * H_own + ng * H_nei);
    return o;
}

// HLLDST::computeFluxDST (HLLDST.C:91-204)
static FaceOut fluxHLL(const FaceIn& in) {
    V3 U_own = in.rhoU_o / in.rho_o, U_nei = in.rhoU_n / in.rho_n;
    double E_own = in.e_o + 0.5 * magSqr(U_own), E_nei = in.e_n + 0.5 * magSqr(U_nei);
    double p_own = in.rho_o * in.rPsi_o, p_nei = in.rho_n * in.rPsi_n;
    double H_own = E_own + p_own / in.rho_o, H_nei = E_nei + p_nei / in.rho_n;
    double coefR = std::sqrt(std::max(VSMALL, in.rho_n) / std::max(VSMALL, in.rho_o));
    double HTilde = (coefR * H_nei + H_own) / (coefR + 1.0);
    V3 UTilde = (coefR * U_nei + U_own) / (coefR + 1.0);
    double aTilde = std::sqrt((in.gam_f - 1.0) * (HTilde - 0.5 * magSqr(UTilde)));
    V3 nrm = in.Sf / in.magSf;
    double Uv_own = dot(U_own, nrm), Uv_nei = dot(U_nei, nrm), UvTilde = dot(UTilde, nrm);
    double S_own = std::min(Uv_own - in.a_o, UvTilde - aTilde);
    double S_nei = std::max(Uv_nei + in.a_n, UvTilde + aTilde);
    double cf_own = pos0(S_own), cf_m = neg(S_own) * pos0(S_nei), cf_nei = neg(S_nei) * neg(S_own);
    V3 rhoUPhi_own = in.rho_o * U_own * Uv_own + p_own * nrm;
    V3 rhoUPhi_nei = in.rho_n * U_nei * Uv_nei + p_nei * nrm;
    double dS = S_nei - S_own;
    FaceOut o;
    o.phi = (cf_own * in.rho_o * Uv_own
           + cf_m * (S_nei * in.rho_o * Uv_own - S_own * in.rho_n * Uv_nei + S_own * S_nei * (in.rho_n - in.rho_o)) / dS
           + cf_nei * in.rho_n * Uv_nei) * in.magSf;
    V3 mid = (S_nei * rhoUPhi_own - S_own * rhoUPhi_nei + S_own * S_nei * (in.rho_n * U_nei - in.rho_o * U_own)) / dS;
    o.phiU = (cf_own * rhoUPhi_own + cf_m * mid + cf_nei * rhoUPhi_nei) * in.magSf;
    o.phiE = (cf_own * in.rho_o * Uv_own * H_own
            + cf_m * (S_nei * in.rho_o * Uv_own * H_own - S_own * in.rho_n * Uv_nei * H_nei + S_own * S_nei * (in.rho_n * E_nei - in.rho_o * E_own)) / dS
            + cf_nei * in.rho_n * Uv_nei * H_nei) * in.magSf;
    V3 upm = (S_nei * U_own - S_own * U_nei + Uv_own * U_own - Uv_nei * U_nei) / dS;
    o.Up = cf_own * U_own + cf_m * upm + cf_nei * U_nei;
    return o;
}

// HLLCDST::computeFluxDST (HLLCDST.C:135-272), directWaveSpeed = true
// HLLCPDST::computeFluxDST (HLLCPDST.C:166-317) when hllcp
static FaceOut fluxHLLC(const FaceIn& in, bool hllcp) {
    V3 U_own = in.rhoU_o / in.rho_o, U_nei = in.rhoU_n / in.rho_n;
    double E_own = in.e_o + 0.5 * magSqr(U_own), E_nei = in.e_n + 0.5 * magSqr(U_nei);
    double p_own = in.rho_o * in.rPsi_o, p_nei = in.rho_n * in.rPsi_n;
    double H_own = E_own + p_own / in.rho_o, H_nei = E_nei + p_nei / in.rho_n;
    V3 normal = in.Sf / in.magSf;
    double Uv_own = dot(U_own, normal), Uv_nei = dot(U_nei, normal);
    double coefR = std::sqrt(std::max(VSMALL, in.rho_n) / std::max(VSMALL, in.rho_o));
    double HTilde = (coefR * H_nei + H_own) / (coefR + 1.0);
    V3 UTilde = (coefR * U_nei + U_own) / (coefR + 1.0);
    double aTilde = std::sqrt((in.gam_f - 1.0) * (HTilde - 0.5 * magSqr(UTilde)));
    double UvTilde = dot(UTilde, normal);
    double S_own = std::min(Uv_own - in.a_o, UvTilde - aTilde);
    double S_nei = std::max(Uv_nei + in.a_n, UvTilde + aTilde);
    double S_star = (p_nei - p_own + in.rho_o * Uv_own * (S_own - Uv_own) - in.rho_n * Uv_nei * (S_nei - Uv_nei))
                  / (in.rho_o * (S_own - Uv_own) - in.rho_n * (S_nei - Uv_nei));
    double cf_own = pos(S_own);
    double cf_m1 = neg0(S_own) * pos(S_star);
    double cf_m2 = neg0(S_own) * neg0(S_star) * pos0(S_nei);
    double cf_nei = neg0(S_own) * neg0(S_star) * neg(S_nei);
    double rhoE_own = in.rho_o * E_own, rhoE_nei = in.rho_n * E_nei;
    V3 rhoUPhi_own = in.rhoU_o * Uv_own + p_own * normal;
    V3 rhoUPhi_nei = in.rhoU_n * Uv_nei + p_nei * normal;
    double rhoEPhi_own = (rhoE_own + p_own) * Uv_own, rhoEPhi_nei = (rhoE_nei + p_nei) * Uv_nei;
    FaceOut o;
    if (!hllcp) {
        double pstar_own = p_own + in.rho_o * (S_own - Uv_own) * (S_star - Uv_own);
        double pstar_nei = p_nei + in.rho_n * (S_nei - Uv_nei) * (S_star - Uv_nei);
        o.phi = (cf_own * in.rho_o * Uv_own
               + cf_m1 * S_star * in.rho_o * (S_own - Uv_own) / (S_own - S_star)
               + cf_m2 * S_star * in.rho_n * (S_nei - Uv_nei) / (S_nei - S_star)
               + cf_nei * in.rho_n * Uv_nei) * in.magSf;
        V3 m1 = (S_star * (S_own * in.rhoU_o - rhoUPhi_own) + S_own * pstar_own * normal) / (S_own - S_star);
        V3 m2 = (S_star * (S_nei * in.rhoU_n - rhoUPhi_nei) + S_nei * pstar_nei * normal) / (S_nei - S_star);
        o.phiU = (cf_own * rhoUPhi_own + cf_m1 * m1 + cf_m2 * m2 + cf_nei * rhoUPhi_nei) * in.magSf;
        o.phiE = (cf_own * rhoEPhi_own
                + cf_m1 * S_star * (S_own * rhoE_own - rhoEPhi_own + S_own * pstar_own) / (S_own - S_star)
                + cf_m2 * S_star * (S_nei * rhoE_nei - rhoEPhi_nei + S_nei * pstar_nei) / (S_nei - S_star)
                + cf_nei * rhoEPhi_nei) * in.magSf;
    } else {
        double Ma_own = Uv_own / in.a_o, Ma_nei = Uv_nei / in.a_n;
        double MaTilde = (coefR * Ma_nei + Ma_own) / (coefR + 1.0);
        double c_own = in.rho_o * (S_own - Uv_own), c_nei = in.rho_n * (S_nei - Uv_nei);
        double p_star = (c_nei * p_own - c_own * p_nei - c_own * c_nei * (Uv_own - Uv_nei)) / (c_nei - c_own);
        double theta = std::min(std::max(std::fabs(Ma_own), std::fabs(Ma_nei)), 1.0);
        double p_av = 0.5 * (p_own + p_nei);
        double p_ss = p_star * theta + (1.0 - theta) * p_av;
        double f = in.fHLLCP;
        double p_sss = f * p_ss + (1.0 - f) * p_star;
        double phip = (f - 1.0) * S_own * S_nei / (S_nei - S_own) / (1.0 + std::fabs(MaTilde)) * (p_nei - p_own) / sqr(aTilde);
        o.phi = (cf_own * in.rho_o * Uv_own
               + cf_m1 * (S_star * in.rho_o * (S_own - Uv_own) / (S_own - S_star) + phip)
               + cf_m2 * (S_star * in.rho_n * (S_nei - Uv_nei) / (S_nei - S_star) + phip)
               + cf_nei * in.rho_n * Uv_nei) * in.magSf;
        V3 m1 = (S_star * (S_own * in.rhoU_o - rhoUPhi_own) + S_own * p_sss * normal) / (S_own - S_star) + phip * UTilde;
        V3 m2 = (S_star * (S_nei * in.rhoU_n - rhoUPhi_nei) + S_nei * p_sss * normal) / (S_nei - S_star) + phip * UTilde;
        o.phiU = (cf_own * rhoUPhi_own + cf_m1 * m1 + cf_m2 * m2 + cf_nei * rhoUPhi_nei) * in.magSf;
        double UT2 = magSqr(UTilde);
        o.phiE = (cf_own * rhoEPhi_own
                + cf_m1 * (S_star * (S_own * rhoE_own - rhoEPhi_own + S_own * p_star) / (S_own - S_star) + 0.5 * phip * UT2)
                + cf_m2 * (S_star * (S_nei * rhoE_nei - rhoEPhi_nei + S_nei * p_star) / (S_nei - S_star) + 0.5 * phip * UT2)
                + cf_nei * rhoEPhi_nei) * in.magSf;
    }
    o.Up = cf_own * U_own + cf_m1 * (S_own - Uv_own) * U_own / (S_own - S_star)
         + cf_m2 * (S_nei - Uv_nei) * U_nei / (S_nei - S_star) + cf_nei * U_nei;
    return o;
}

// KurganovDST::computeFluxDST (KurganovDST.C:93-167)
static FaceOut fluxKurganov(const FaceIn& in) {
    V3 U_own = in.rhoU_o / in.rho_o, U_nei = in.rhoU_n / in.rho_n;
    double rhoE_own = in.rho_o * (in.e_o + 0.5 * magSqr(U_own));
    double rhoE_nei = in.rho_n * (in.e_n + 0.5 * magSqr(U_nei));
    double p_own = in.rho_o * in.rPsi_o, p_nei = in.rho_n * in.rPsi_n;
    double phiv_own = dot(U_own, in.Sf), phiv_nei = dot(U_nei, in.Sf);
    double cSf_own = in.c_o * in.magSf, cSf_nei = in.c_n * in.magSf;
    double ap = std::max(std::max(phiv_own + cSf_own, phiv_nei + cSf_nei), 0.0);
    double am = std::min(std::min(phiv_own - cSf_own, phiv_nei - cSf_nei), 0.0);
    double a_own = ap / (ap - am);
    double aSf = am * a_own;
    double a_nei = 1.0 - a_own;
    phiv_own *= a_own;
    phiv_nei *= a_nei;
    double aphiv_own = phiv_own - aSf, aphiv_nei = phiv_nei + aSf;
    FaceOut o;
    o.phi = aphiv_own * in.rho_o + aphiv_nei * in.rho_n;
    o.phiU = (aphiv_own * in.rhoU_o + aphiv_nei * in.rhoU_n) + (a_own * p_own + a_nei * p_nei) * in.Sf;
    o.phiE = aphiv_own * (rhoE_own + p_own) + aphiv_nei * (rhoE_nei + p_nei) + aSf * p_own - aSf * p_nei;
    o.Up = a_own * U_own + a_nei * U_nei;
    return o;
}

static FaceOut faceFlux(int scheme, const FaceIn& in, const rdst_settings& st) {
    switch (scheme) {
        case RDST_FLUX_AUSMPLUS: return fluxAUSMplus(in);
        case RDST_FLUX_AUSMPLUSUP: return fluxAUSMplusUp(in, st.low_mach != 0, st.Mainf);
        case RDST_FLUX_HLL: return fluxHLL(in);
        case RDST_FLUX_HLLC: return fluxHLLC(in, false);
        case RDST_FLUX_HLLCP: return fluxHLLC(in, true);
        case RDST_FLUX_KURGANOV: return fluxKurganov(in);
    }
    throw std::runtime_error("unknown flux scheme");
}

// HLLCPDST::computef (HLLCPDST.C:131-164): f = interpolate(pow3(fCells)), fCells boundary = 1
static void computeHLLCPf(const Case& c, std::vector<double>& f) {
    const Mesh& m = c.m;
    std::vector<double> fc(m.nCells, 1.0);
    for (int i = 0; i < m.nCells; ++i)
        for (int j : m.cellCells[i])
            fc[i] = std::min(fc[i], std::min(c.p[i] / c.p[j], c.p[j] / c.p[i]));
    std::vector<double> f3(m.nCells), b1(m.nB, 1.0);
    for (int i = 0; i < m.nCells; ++i) f3[i] = fc[i] * fc[i] * fc[i];
    linearInterp(m, f3, b1, f);
}


// ------------------------------------------------------------------ viscous terms of calcFields
// rank-2 tensor, T[3*i + j]

// snGrad() of the U boundary condition on boundary face b (fvPatchField
// snGrad: fixedValue dc (U_b - U_I); zeroGradient 0; basicSymmetry
// (transform(I - 2nn, U_I) - U_I) dc/2; mixed vf (ref - U_I) dc + (1-vf) 0)
static V3 snGradU_b(const Case& c, int b) {
    const Mesh& m = c.m;
    const rdst_bc& bc = c.bc[m.bPatch[b]];
    int f = m.bFace[b], ci = m.bCell[b];
    double dcb = m.dc[m.nInternal + b];
    V3 UI = c.U[ci];
    switch (bc.U_type) {
        case RDST_BC_FIXED_VALUE: return dcb * (c.bU[b] - UI);
        case RDST_BC_SYMMETRY: {
            V3 n = m.Sf[f] / m.magSf[f];
            double nU = dot(n, UI);
            V3 t = v3(UI.x - 2.0 * n.x * nU, UI.y - 2.0 * n.y * nU, UI.z - 2.0 * n.z * nU);
            return (t - UI) * (dcb / 2.0);
        }
        case RDST_BC_INLET_OUTLET: case RDST_BC_FARFIELD: {
            double vf = c.bcs.vf_U[b];
            return vf * (c.bcs.refU[b] - UI) * dcb + (1.0 - vf) * v3(0, 0, 0);
        }
        default: return v3(0, 0, 0);
    }
}

// cellLimitedGrad (foam-extend 4.1 gradSchemes, not part of the reference
// tree; tutorials/RAE2822/system/fvSchemes:27 selects "cellLimited Gauss
// linear 1" for grad(U)), restated from its published algorithm: the max and
// min of each U component over a cell, its face neighbours (a face loop,
// owner and neighbour sides) and its boundary-face values, less the cell
// value; for k < 1 both widened by (1/k - 1)(max - min); a limiter l = 1 per
// component, lowered face by face (internal faces for both sides, then
// boundary faces) by limitFace on the change (Cf - C) & grad towards the
// face centre; grad_ij *= l_j.  limitFace: where the change exceeds
// max + VSMALL, l = min(l, max/change); else where it is below min - VSMALL,
// l = min(l, min/change); VSMALL = 1e-300.
static void limitFaceCmp(double& l, double dmax, double dmin, double e) {
    if (e > dmax + 1.0e-300) l = std::min(l, dmax / e);
    else if (e < dmin - 1.0e-300) l = std::min(l, dmin / e);
}
static void limitFace(V3& l, const V3& dmax, const V3& dmin, const V3& e) {
    limitFaceCmp(l.x, dmax.x, dmin.x, e.x);
    limitFaceCmp(l.y, dmax.y, dmin.y, e.y);
    limitFaceCmp(l.z, dmax.z, dmin.z, e.z);
}
static V3 maxV(const V3& a, const V3& b) { return v3(std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)); }
static V3 minV(const V3& a, const V3& b) { return v3(std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)); }
static inline V3 dotVT(V3 v, const T9& t);
static void cellLimitGradU(const Case& c, double k, std::vector<T9>& g) {
    const Mesh& m = c.m;
    const int n = m.nCells;
    std::vector<V3> mx(c.U.begin(), c.U.begin() + n), mn(c.U.begin(), c.U.begin() + n);
    for (int f = 0; f < m.nInternal; ++f) {
        const int o = m.owner[f], nb = m.neighbour[f];
        mx[o] = maxV(mx[o], c.U[nb]); mn[o] = minV(mn[o], c.U[nb]);
        mx[nb] = maxV(mx[nb], c.U[o]); mn[nb] = minV(mn[nb], c.U[o]);
    }
    for (int b = 0; b < m.nB; ++b) {
        const int o = m.bCell[b];
        mx[o] = maxV(mx[o], c.bU[b]); mn[o] = minV(mn[o], c.bU[b]);
    }
    for (int i = 0; i < n; ++i) { mx[i] = mx[i] - c.U[i]; mn[i] = mn[i] - c.U[i]; }
    if (k < 1.0) {
        for (int i = 0; i < n; ++i) {
            const V3 d = (1.0 / k - 1.0) * (mx[i] - mn[i]);
            mx[i] = mx[i] + d; mn[i] = mn[i] - d;
        }
    }
    std::vector<V3> lim(n, v3(1.0, 1.0, 1.0));
    for (int f = 0; f < m.nInternal; ++f) {
        const int o = m.owner[f], nb = m.neighbour[f];
        limitFace(lim[o], mx[o], mn[o], dotVT(m.Cf[f] - m.C[o], g[o]));
        limitFace(lim[nb], mx[nb], mn[nb], dotVT(m.Cf[f] - m.C[nb], g[nb]));
    }
    for (int b = 0; b < m.nB; ++b) {
        const int o = m.bCell[b];
        limitFace(lim[o], mx[o], mn[o], dotVT(m.Cf[m.bFace[b]] - m.C[o], g[o]));
    }
    for (int i = 0; i < n; ++i)
        for (int r = 0; r < 3; ++r)
            for (int j = 0; j < 3; ++j) g[i].a[3 * r + j] = cmp(lim[i], j) * g[i].a[3 * r + j];
}

// fvc::grad(U) Gauss linear (cells), cellLimited when the settings select it,
// and the corrected boundary gradient (gaussGrad::correctBoundaryConditions:
// g_b = g_I + n (snGrad_b - n.g_I)).  graw, when given, receives the
// unlimited Gauss linear cell gradient (the snGrad correction's).
static void gradU(const Case& c, std::vector<T9>& g, std::vector<T9>& gb, std::vector<T9>* graw = nullptr) {
    const Mesh& m = c.m;
    int n = m.nCells;
    g.assign(n, T9{});
    // face values of U (linear interpolation; the boundary value on boundary faces)
    std::vector<V3> Uface(m.nInternal + m.nB);
#pragma omp parallel for schedule(static)
    for (int f = 0; f < m.nInternal; ++f) {
        double w = m.w[f];
        V3 P = c.U[m.owner[f]], N = c.U[m.neighbour[f]];
        Uface[f] = v3(w * (P.x - N.x) + N.x, w * (P.y - N.y) + N.y, w * (P.z - N.z) + N.z);
    }
    for (int b = 0; b < m.nB; ++b) Uface[m.nInternal + b] = c.bU[b];
    // Gauss sum in face order (owner +, neighbour -, boundary +), then / V
#pragma omp parallel for schedule(static)
    for (int ci = 0; ci < n; ++ci) {
        for (int q = m.cfStart[ci]; q < m.cfStart[ci + 1]; ++q) {
            const int e = m.cfIdx[q];
            const int k = e > 0 ? e - 1 : -e - 1;
            const V3 sf = SfOf(m, k), Uf = Uface[k];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double t = cmp(sf, i) * cmp(Uf, j);
                    if (e > 0) g[ci].a[3 * i + j] += t;
                    else g[ci].a[3 * i + j] -= t;
                }
        }
        for (int k = 0; k < 9; ++k) g[ci].a[k] /= m.V[ci];
    }
    if (graw) *graw = g;
    if (c.st.grad_u == RDST_GRAD_CELL_LIMITED && c.st.grad_u_limit >= 1.0e-15) cellLimitGradU(c, c.st.grad_u_limit, g);
    gb.assign(m.nB, T9{});
    for (int b = 0; b < m.nB; ++b) {
        int f = m.bFace[b], ci = m.bCell[b];
        V3 nrm = m.Sf[f] / m.magSf[f];
        const T9& gi = g[ci];
        V3 nG = v3(nrm.x * gi.a[0] + nrm.y * gi.a[3] + nrm.z * gi.a[6],
                   nrm.x * gi.a[1] + nrm.y * gi.a[4] + nrm.z * gi.a[7],
                   nrm.x * gi.a[2] + nrm.y * gi.a[5] + nrm.z * gi.a[8]);
        V3 d = snGradU_b(c, b) - nG;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) gb[b].a[3 * i + j] = gi.a[3 * i + j] + cmp(nrm, i) * cmp(d, j);
    }
}

// tauMC = muEff dev2(T(grad U)), dev2(A) = A - (2/3) tr(A) I
static T9 tauMC(double mu, const T9& g) {
    T9 t;
    double tr = g.a[0] + g.a[4] + g.a[8];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double a = g.a[3 * j + i] - (i == j ? (2.0 / 3.0) * tr : 0.0);
            t.a[3 * i + j] = mu * a;
        }
    return t;
}
static inline V3 dotVT(V3 v, const T9& t) {  // (v & T)_j = v_i T_ij
    return v3(v.x * t.a[0] + v.y * t.a[3] + v.z * t.a[6], v.x * t.a[1] + v.y * t.a[4] + v.z * t.a[7],
              v.x * t.a[2] + v.y * t.a[5] + v.z * t.a[8]);
}

// Surface-normal gradient schemes (foam-extend 4.1 snGradSchemes, not part of
// the reference tree; the reference selects them in system/fvSchemes, e.g.
// tutorials/oneraM6Turb/system/fvSchemes:44 "Gauss linear limited 0.33").
// corrected: dc (x_N - x_P) + corr & interp(grad x); uncorrected: the first
// term; limited psi (limitedSnGrad::correction): the correction times
// min(psi |orthogonal part| / ((1 - psi) |correction| + SMALL), 1), with |.|
// the magnitude of the face value (vector or scalar) and SMALL = 1e-15.
// Encoded as psi = -1 (corrected), 0 (uncorrected) or the limited psi.
static double snGradPsi(int kind, double limit) {
    return kind == RDST_SNGRAD_CORRECTED ? -1.0 : (kind == RDST_SNGRAD_UNCORRECTED ? 0.0 : limit);
}
static double snGradLimiter(double psi, double u2, double c2) {
    if (psi < 0.0) return 1.0;
    return std::fmin(psi * std::sqrt(u2) / ((1.0 - psi) * std::sqrt(c2) + 1.0e-15), 1.0);
}

// baseFluxDST.C:229-250: R1 -= laplacian(muEff, U) + div(tauMC);
// R4 -= div((interp(muEff)|S| snGrad(U) + Sf & interp(tauMC)) & Up).
// Gauss linear gradient, "corrected" snGrad with non-orthogonal correction
// corr & interp(grad U); no heat-conduction term (none in the reference).
static void viscousTerms(Case& c) {
    const Mesh& m = c.m;
    const double psiL = snGradPsi(c.st.laplacian_sngrad, c.st.laplacian_limit);
    const double psiS = snGradPsi(c.st.sngrad, c.st.sngrad_limit);
    int n = m.nCells;
    std::vector<T9> g, gb, gr;
    gradU(c, g, gb, &gr);
    std::vector<double> muEff, bmuEff;
    muAlphaEff(c, muEff, bmuEff, nullptr, nullptr);
    std::vector<T9> tau(n), btau(m.nB);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) tau[i] = tauMC(muEff[i], g[i]);
    for (int b = 0; b < m.nB; ++b) btau[b] = tauMC(bmuEff[b], gb[b]);
    std::vector<V3> lap(n, v3(0, 0, 0)), dtau(n, v3(0, 0, 0));
    std::vector<double> sdu(n, 0.0);
    // face terms, then summed per cell in face order
    const int nf = m.nInternal + m.nB;
    std::vector<V3> lapFace(nf), tauFace(nf);
    std::vector<double> sdFace(nf);
#pragma omp parallel for schedule(static)
    for (int f = 0; f < m.nInternal; ++f) {
        int o = m.owner[f], nb = m.neighbour[f];
        double w = m.w[f];
        double muf = w * (muEff[o] - muEff[nb]) + muEff[nb];
        T9 gf, tf;
        // the correction interpolates the unlimited Gauss linear gradient
        // (foam-extend 4.1 correctedSnGrad: gradScheme "snGradCorr(U)")
        for (int k = 0; k < 9; ++k) {
            gf.a[k] = w * (gr[o].a[k] - gr[nb].a[k]) + gr[nb].a[k];
            tf.a[k] = w * (tau[o].a[k] - tau[nb].a[k]) + tau[nb].a[k];
        }
        // orthogonal part and non-orthogonal correction of snGrad(U): the
        // laplacian's scheme (laplacianSchemes) and fvc::snGrad's
        // (snGradSchemes) limit the correction separately
        V3 u = m.dc[f] * (c.U[nb] - c.U[o]);
        V3 q = dotVT(m.corr[f], gf);
        double kL = snGradLimiter(psiL, dot(u, u), dot(q, q));
        double kS = snGradLimiter(psiS, dot(u, u), dot(q, q));
        double mg = muf * m.magSf[f];
        V3 lapF = mg * (u + kL * q);
        V3 sgF = mg * (u + kS * q);
        V3 tauF = dotVT(m.Sf[f], tf);
        lapFace[f] = lapF;
        tauFace[f] = tauF;
        sdFace[f] = dot(sgF + tauF, c.Up[f]);
    }
    for (int b = 0; b < m.nB; ++b) {
        int f = m.bFace[b], k = m.nInternal + b;
        V3 lapF = (bmuEff[b] * m.magSf[f]) * snGradU_b(c, b);
        V3 tauF = dotVT(m.Sf[f], btau[b]);
        lapFace[k] = lapF;
        tauFace[k] = tauF;
        sdFace[k] = dot(lapF + tauF, c.Up[k]);
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        for (int q = m.cfStart[i]; q < m.cfStart[i + 1]; ++q) {
            const int e = m.cfIdx[q];
            if (e > 0) {
                const int k = e - 1;
                lap[i] = lap[i] + lapFace[k]; dtau[i] = dtau[i] + tauFace[k]; sdu[i] += sdFace[k];
            } else {
                const int k = -e - 1;
                lap[i] = lap[i] - lapFace[k]; dtau[i] = dtau[i] - tauFace[k]; sdu[i] -= sdFace[k];
            }
        }
    }
    // heat conduction (an extension, off in the reference's equations):
    // R4 -= laplacian(alphaEff, h), orthogonal part, summed in face order
    std::vector<double> hq(n, 0.0);
    if (c.st.heat_conduction) {
        std::vector<double> al, bal;
        std::vector<double> mu_, bmu_;
        muAlphaEff(c, mu_, bmu_, &al, &bal);
        std::vector<double> qFace(nf);
        for (int f = 0; f < m.nInternal; ++f) {
            int o = m.owner[f], nb = m.neighbour[f];
            double alf = m.w[f] * (al[o] - al[nb]) + al[nb];
            qFace[f] = alf * m.magSf[f] * (m.dc[f] * (c.h[nb] - c.h[o]));
        }
        for (int b = 0; b < m.nB; ++b) {
            int f = m.bFace[b], k = m.nInternal + b;
            qFace[k] = bal[b] * m.magSf[f] * (m.dc[k] * (c.bh[b] - c.h[m.bCell[b]]));
        }
        for (int i = 0; i < n; ++i)
            for (int q = m.cfStart[i]; q < m.cfStart[i + 1]; ++q) {
                const int e = m.cfIdx[q];
                if (e > 0) hq[i] += qFace[e - 1];
                else hq[i] -= qFace[-e - 1];
            }
    }
    for (int i = 0; i < n; ++i) {
        double V = m.V[i];
        V3 L = lap[i] / V, D = dtau[i] / V;
        V3 s = L + D;
        c.R[5 * i + 1] -= s.x; c.R[5 * i + 2] -= s.y; c.R[5 * i + 3] -= s.z;
        c.R[5 * i + 4] -= sdu[i] / V;
        if (c.st.heat_conduction) c.R[5 * i + 4] -= hq[i] / V;
    }
}

// computeFluxDST + calcFields (baseFluxDST.C:223-255)
// the face loop's arrays, kept from call to call (every entry is rewritten
// by each call; first-touch page faults of fresh arrays otherwise dominate a
// threaded call)
struct FluxWork {
    CellAux a;
    std::vector<double> rho_o, rho_n, rPsi_o, rPsi_n, e_o, e_n, a_o, a_n, c_o, c_n, gam_f;
    std::vector<V3> rhoU_o, rhoU_n;
    std::vector<double> d0, d4;
    std::vector<V3> d1;
};

static void computeFlux(Case& c) {
    const Mesh& m = c.m;
    static FluxWork w;
    CellAux& a = w.a;
    makeAux(c, a);
    std::vector<double>&rho_o = w.rho_o, &rho_n = w.rho_n, &rPsi_o = w.rPsi_o, &rPsi_n = w.rPsi_n, &e_o = w.e_o,
                       &e_n = w.e_n, &a_o = w.a_o, &a_n = w.a_n, &c_o = w.c_o, &c_n = w.c_n, &gam_f = w.gam_f;
    std::vector<double> fH;
    std::vector<V3>&rhoU_o = w.rhoU_o, &rhoU_n = w.rhoU_n;
    if (c.st.reconstruction == RDST_RECON_OSPRE) {
        reconstructOSPRE(m, c.rho, c.brho, rho_o, rho_n);           // reconstruct(rho) OSPRE
        reconstructOSPREV(m, c.rhoU, c.brhoU, rhoU_o, rhoU_n);      // reconstruct(U) OSPREV
        reconstructOSPRE(m, a.rPsi, a.brPsi, rPsi_o, rPsi_n);       // reconstruct(T) OSPRE
        reconstructOSPRE(m, a.e, a.be, e_o, e_n);
        reconstructOSPRE(m, a.aStar, a.baStar, a_o, a_n);
        reconstructOSPRE(m, a.c, a.bc, c_o, c_n);
    } else {
        reconstruct(m, c.rho, c.brho, rho_o, rho_n);
        reconstruct(m, c.rhoU, c.brhoU, rhoU_o, rhoU_n);
        reconstruct(m, a.rPsi, a.brPsi, rPsi_o, rPsi_n);
        reconstruct(m, a.e, a.be, e_o, e_n);
        reconstruct(m, a.aStar, a.baStar, a_o, a_n);
        reconstruct(m, a.c, a.bc, c_o, c_n);
    }
    linearInterp(m, a.gam, a.bgam, gam_f);
    if (c.st.flux_scheme == RDST_FLUX_HLLCP) computeHLLCPf(c, fH);
    int nf = m.nInternal + m.nB;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nf; ++k) {
        FaceIn in;
        in.rho_o = rho_o[k]; in.rho_n = rho_n[k]; in.rPsi_o = rPsi_o[k]; in.rPsi_n = rPsi_n[k];
        in.e_o = e_o[k]; in.e_n = e_n[k]; in.a_o = a_o[k]; in.a_n = a_n[k];
        in.c_o = c_o[k]; in.c_n = c_n[k]; in.gam_f = gam_f[k];
        in.fHLLCP = fH.empty() ? 1.0 : fH[k];
        in.rhoU_o = rhoU_o[k]; in.rhoU_n = rhoU_n[k];
        in.Sf = SfOf(m, k); in.magSf = magSfOf(m, k);
        FaceOut o = faceFlux(c.st.flux_scheme, in, c.st);
        c.phi[k] = o.phi; c.phiUp[k] = o.phiU; c.phiEp[k] = o.phiE; c.Up[k] = o.Up;
    }
    // calcFields: fvc::div = surfaceIntegrate (sum owner +, neighbour -, boundary +) / V
    int n = m.nCells;
    std::vector<double>&d0 = w.d0, &d4 = w.d4;
    std::vector<V3>& d1 = w.d1;
    d0.resize(n); d4.resize(n); d1.resize(n);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double s0 = 0.0, s4 = 0.0;
        V3 s1 = v3(0, 0, 0);
        for (int q = m.cfStart[i]; q < m.cfStart[i + 1]; ++q) {
            const int e = m.cfIdx[q];
            if (e > 0) {
                const int k = e - 1;
                s0 += c.phi[k]; s1 = s1 + c.phiUp[k]; s4 += c.phiEp[k];
            } else {
                const int k = -e - 1;
                s0 -= c.phi[k]; s1 = s1 - c.phiUp[k]; s4 -= c.phiEp[k];
            }
        }
        d0[i] = s0; d1[i] = s1; d4[i] = s4;
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        c.R[5 * i] = d0[i] / m.V[i];
        c.R[5 * i + 1] = d1[i].x / m.V[i]; c.R[5 * i + 2] = d1[i].y / m.V[i]; c.R[5 * i + 3] = d1[i].z / m.V[i];
        c.R[5 * i + 4] = d4[i] / m.V[i];
    }
    if (c.st.viscous) viscousTerms(c);
}

// ------------------------------------------------------------------ Jacobian
// Euler flux Jacobian A(W).n in the Blazek A.7 form, Roe::calcJ (Roe.C:247-272)
static void eulerJacobian(V3 U, double gamma, double rhoE, double rho, V3 n, double* J) {
    double phiJ = 0.5 * (gamma - 1) * magSqr(U);
    double a1 = gamma * rhoE / rho - phiJ;
    for (int i = 0; i < 25; ++i) J[i] = 0.0;
    double Un = dot(U, n);
    // insertScalarEqn(nSf, 0, 1): row 0, cols 1..3
    J[1] = n.x; J[2] = n.y; J[3] = n.z;
    // insertVectorEqn(phiJ*n - (U*U)&n, 1, 0): column 0, rows 1..3
    for (int i = 0; i < 3; ++i) J[5 * (1 + i)] = phiJ * cmp(n, i) - cmp(U, i) * Un;
    // insertVectorEqn(U*n + (U&n) I - (gamma-1)(n*U), 1, 1)
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            J[5 * (1 + i) + 1 + j] = cmp(U, i) * cmp(n, j) + (i == j ? Un : 0.0) - (gamma - 1) * (cmp(n, i) * cmp(U, j));
    // insertVectorEqn((gamma-1) n, 1, 4): column 4
    for (int i = 0; i < 3; ++i) J[5 * (1 + i) + 4] = (gamma - 1) * cmp(n, i);
    // insertScalarEqn((U&n)(phiJ - a1), 4, 0)
    J[20] = Un * (phiJ - a1);
    // insertScalarEqn(a1 n - (U&n)(gamma-1) U, 4, 1)
    for (int j = 0; j < 3; ++j) J[21 + j] = a1 * cmp(n, j) - Un * (gamma - 1) * cmp(U, j);
    // insertScalarEqn(gamma (U&n), 4, 4)
    J[24] = gamma * Un;
}

static void matmul5(const double* A, const double* B, double* C) {
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) {
            double s = 0.0;
            for (int k = 0; k < 5; ++k) s += A[5 * i + k] * B[5 * k + j];
            C[5 * i + j] = s;
        }
}

// Roe::computeJacobianDST (Roe.C:152-214) + computeBoundaryJacobians (:274-379)
static void computeJacobian(Case& c) {
    const Mesh& m = c.m;
    int n = m.nCells;
    std::vector<V3> Uc(n);
    std::vector<double> gam(n);
    for (int i = 0; i < n; ++i) { Uc[i] = c.rhoU[i] / c.rho[i]; gam[i] = c.th.Cp / c.Cv; }
    c.ownC.assign(25 * m.nInternal, 0.0);
    c.neiC.assign(25 * m.nInternal, 0.0);
    for (int f = 0; f < m.nInternal; ++f) {
        V3 nSf = m.Sf[f] / m.magSf[f];
        int o = m.owner[f], nb = m.neighbour[f];
        eulerJacobian(Uc[o], gam[o], c.rhoE[o], c.rho[o], nSf, &c.ownC[25 * f]);
        eulerJacobian(Uc[nb], gam[nb], c.rhoE[nb], c.rho[nb], nSf, &c.neiC[25 * f]);
        for (int k = 0; k < 25; ++k) {
            c.ownC[25 * f + k] *= m.w[f];
            c.neiC[25 * f + k] *= (1 - m.w[f]);
        }
    }
    // spectral radius (Roe.C:181-201)
    std::vector<V3> bUc(m.nB);
    for (int b = 0; b < m.nB; ++b) bUc[b] = c.brhoU[b] / c.brho[b];
    // interpolate(U_) with U_ = rhoU/rho (boundary values rhoU_b/rho_b)
    std::vector<V3> Uf;
    linearInterpV(m, Uc, bUc, Uf);
    std::vector<double> cc(n), bcc(m.nB), cf;
    for (int i = 0; i < n; ++i) cc[i] = std::sqrt(c.th.Cp / c.Cv / c.psi[i]);
    for (int b = 0; b < m.nB; ++b) bcc[b] = std::sqrt(c.th.Cp / c.Cv / c.bpsi[b]);
    linearInterp(m, cc, bcc, cf);
    std::vector<double> muEffSf, alphaEffSf, rhoSf, trho(n), btrho(m.nB);
    std::vector<double> muE, bmuE, alE, balE;
    muAlphaEff(c, muE, bmuE, &alE, &balE);
    linearInterp(m, muE, bmuE, muEffSf);
    linearInterp(m, alE, balE, alphaEffSf);
    for (int i = 0; i < n; ++i) trho[i] = c.psi[i] * c.p[i];       // thermo.rho() = psi*p
    for (int b = 0; b < m.nB; ++b) btrho[b] = c.bpsi[b] * c.bp[b];
    linearInterp(m, trho, btrho, rhoSf);
    c.lambda.assign(m.nInternal, 0.0);
    for (int f = 0; f < m.nInternal; ++f) {
        double amaxSf = std::fabs(dot(Uf[f], m.Sf[f])) + m.magSf[f] * cf[f];
        c.lambda[f] = 0.5 * (amaxSf + m.magSf[f] * (muEffSf[f] + alphaEffSf[f]) / rhoSf[f] * m.dc[f]);
    }
    // boundary Jacobians on non-coupled patches (Roe.C:306-377)
    c.patchIC.assign(25 * m.nB, 0.0);
    for (int b = 0; b < m.nB; ++b) {
        int f = m.bFace[b], ci = m.bCell[b];
        const rdst_bc& bc = c.bc[m.bPatch[b]];
        V3 pSf = m.Sf[f];
        V3 UB = c.brhoU[b] / c.brho[b];
        double pB = c.bp[b], TB = c.bT[b], rhoB = c.brho[b], CvB = c.Cv;
        double EB = c.brhoE[b] / c.brho[b];
        double rRT = rhoB / pB;
        double USf = dot(UB, pSf);
        double dR[25] = {0};
        dR[0] = USf * rRT;
        dR[1] = rhoB * pSf.x; dR[2] = rhoB * pSf.y; dR[3] = rhoB * pSf.z;
        dR[4] = -rhoB / TB * USf;
        for (int i = 0; i < 3; ++i) {
            dR[5 * (1 + i)] = rRT * USf * cmp(UB, i) + cmp(pSf, i);
            for (int j = 0; j < 3; ++j)
                dR[5 * (1 + i) + 1 + j] = rhoB * cmp(UB, i) * cmp(pSf, j) + (i == j ? rhoB * USf : 0.0);
            dR[5 * (1 + i) + 4] = -rhoB / TB * USf * cmp(UB, i);
        }
        dR[20] = (EB * rRT + 1) * USf;
        for (int j = 0; j < 3; ++j) dR[21 + j] = rhoB * USf * cmp(UB, j) + cmp(pSf, j) * (rhoB * EB + pB);
        dR[24] = rhoB * CvB * USf;
        // dX/dW at the patch-internal cell
        V3 UI = Uc[ci];
        double rhoEI = c.rhoE[ci], pI = c.p[ci], rRho = 1.0 / c.rho[ci], Cv = c.Cv;
        double rRhoCv = rRho / Cv;
        double e = rhoEI * rRho - 0.5 * magSqr(UI);
        double RT = pI * rRho;
        double RCv = RT / e;
        double dX[25] = {0};
        dX[0] = RT - RCv * (e - 0.5 * magSqr(UI));
        dX[1] = -RCv * UI.x; dX[2] = -RCv * UI.y; dX[3] = -RCv * UI.z;
        dX[4] = RCv;
        for (int i = 0; i < 3; ++i) {
            dX[5 * (1 + i)] = -rRho * cmp(UI, i);
            dX[5 * (1 + i) + 1 + i] = rRho;
        }
        dX[20] = rRhoCv * (0.5 * magSqr(UI) - e);
        dX[21] = -rRhoCv * UI.x; dX[22] = -rRhoCv * UI.y; dX[23] = -rRhoCv * UI.z;
        dX[24] = rRhoCv;
        // dXb/dXi = diag(valueInternalCoeffs)
        V3 nf = pSf / m.magSf[f];
        V3 uvic = vectorVIC(bc.U_type, c.bcs.vf_U[b], nf);
        double D[25] = {0};
        D[0] = scalarVIC(bc.p_type, c.bcs.vf_p[b]);
        D[6] = uvic.x; D[12] = uvic.y; D[18] = uvic.z;
        D[24] = scalarVIC(bc.T_type, c.bcs.vf_T[b]);
        double tmp[25];
        matmul5(dR, D, tmp);
        matmul5(tmp, dX, &c.patchIC[25 * b]);
    }
}

// ------------------------------------------------------------------ assembly
// fvmDST::ddt (fvmDST.C:77-143) with steadyStateDST::fvmDdt (steadyStateDST.C:173-198)
// + fluxDST::integrate (fluxDST.C:73-127), summed by operator+ (fvBlockMatrixDST.C:201-213)
// dual2ndDST coefficients (dual2ndDST.C:395-401): a = 1/deltaT + 1/(deltaT +
// deltaT0), c = deltaT/(deltaT0 (deltaT + deltaT0)), b = -(a + c)
static void dualCoeffs(const Case& c, double& a, double& b, double& cc) {
    double rDeltaT = 1.0 / c.deltaT;
    double deltaT = c.deltaT, deltaT0 = c.deltaT0;
    a = rDeltaT + 1.0 / (deltaT + deltaT0);
    cc = deltaT / (deltaT0 * (deltaT + deltaT0));
    b = -(a + cc);
}

static void assemble(Case& c) {
    const Mesh& m = c.m;
    int n = m.nCells, F = m.nInternal;
    double beta = c.st.beta;
    std::vector<double> dd(25 * n, 0.0), sd(5 * n, 0.0);
    const bool dual = c.st.time_integration == RDST_TIME_DUAL;
    double da = 0, db = 0, dcf = 0;
    if (dual) dualCoeffs(c, da, db, dcf);
    for (int i = 0; i < n; ++i) {
        double x[5] = {c.rho[i], c.rhoU[i].x, c.rhoU[i].y, c.rhoU[i].z, c.rhoE[i]};
        if (dual) {
            // dual2ndDST::fvmDdt (dual2ndDST.C:377-411): diag = (rDeltaTau + a) V,
            // source = V (rDeltaTau psi - b psi.oldTime() - c psi.oldTime().oldTime());
            // the block keeps source - diag*psi (fvmDST.C:101-133)
            double diagS = (c.rDeltaT[i] + da) * m.V[i];
            for (int r = 0; r < 5; ++r) dd[25 * i + 6 * r] = diagS;
            for (int k = 0; k < 5; ++k) {
                double src = m.V[i] * (c.rDeltaT[i] * x[k] - db * c.WOld[5 * i + k] - dcf * c.WOldOld[5 * i + k]);
                sd[5 * i + k] = src - diagS * x[k];
            }
            continue;
        }
        double diagS = c.rDeltaT[i] * m.V[i];
        for (int r = 0; r < 5; ++r) dd[25 * i + 6 * r] = diagS;  // I * rhoEqn.diag()
        // source - diag*psi with source = rDeltaT*psi_old*V and psi_old = psi
        for (int k = 0; k < 5; ++k) sd[5 * i + k] = c.rDeltaT[i] * x[k] * m.V[i] - diagS * x[k];
    }
    // flux matrix
    std::vector<double> fd(25 * n, 0.0);
    c.lower.assign(25 * F, 0.0); c.upper.assign(25 * F, 0.0);
    for (int f = 0; f < F; ++f) {
        for (int k = 0; k < 25; ++k) {
            double lam = (k % 6 == 0) ? c.lambda[f] : 0.0;
            c.lower[25 * f + k] = -beta * (c.ownC[25 * f + k] * m.magSf[f] + lam);
            c.upper[25 * f + k] = beta * (c.neiC[25 * f + k] * m.magSf[f] - lam);
        }
    }
    // negSumDiag
    for (int f = 0; f < F; ++f) {
        int o = m.owner[f], nb = m.neighbour[f];
        for (int k = 0; k < 25; ++k) {
            fd[25 * o + k] -= c.lower[25 * f + k];
            fd[25 * nb + k] -= c.upper[25 * f + k];
        }
    }
    for (int b = 0; b < m.nB; ++b) {
        int ci = m.bCell[b];
        for (int k = 0; k < 25; ++k) fd[25 * ci + k] += beta * (c.patchIC[25 * b + k] + 0.0);  // spectPatch = 0
    }
    std::vector<double> fs(5 * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 5; ++k) fs[5 * i + k] -= c.R[5 * i + k] * m.V[i];
    c.diag.resize(25 * n);
    c.source.resize(5 * n);
    for (int i = 0; i < 25 * n; ++i) c.diag[i] = dd[i] + fd[i];
    for (int i = 0; i < 5 * n; ++i) c.source[i] = sd[i] + fs[i];
    c.hasMatrix = true;
}

// ------------------------------------------------------------------ linear algebra on LDU
static void Amul(const Case& c, const std::vector<double>& x, std::vector<double>& y) {
    const Mesh& m = c.m;
    int n = m.nCells;
    y.assign(5 * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int r = 0; r < 5; ++r) {
            double s = 0;
            for (int k = 0; k < 5; ++k) s += c.diag[25 * i + 5 * r + k] * x[5 * i + k];
            y[5 * i + r] = s;
        }
    for (int f = 0; f < m.nInternal; ++f) {
        int l = m.owner[f], u = m.neighbour[f];
        for (int r = 0; r < 5; ++r) {
            double su = 0, sl = 0;
            for (int k = 0; k < 5; ++k) {
                sl += c.lower[25 * f + 5 * r + k] * x[5 * l + k];
                su += c.upper[25 * f + 5 * r + k] * x[5 * u + k];
            }
            y[5 * u + r] += sl;
            y[5 * l + r] += su;
        }
    }
}

// 5x5 inverse, Gauss-Jordan with partial pivoting
static bool inv5(const double* A, double* Ainv) {
    double a[5][10];
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) { a[i][j] = A[5 * i + j]; a[i][5 + j] = (i == j) ? 1.0 : 0.0; }
    for (int col = 0; col < 5; ++col) {
        int piv = col;
        for (int r = col + 1; r < 5; ++r) if (std::fabs(a[r][col]) > std::fabs(a[piv][col])) piv = r;
        if (a[piv][col] == 0.0) return false;
        if (piv != col) for (int j = 0; j < 10; ++j) std::swap(a[piv][j], a[col][j]);
        double d = 1.0 / a[col][col];
        for (int j = 0; j < 10; ++j) a[col][j] *= d;
        for (int r = 0; r < 5; ++r) {
            if (r == col) continue;
            double fct = a[r][col];
            if (fct == 0.0) continue;
            for (int j = 0; j < 10; ++j) a[r][j] -= fct * a[col][j];
        }
    }
    for (int i = 0; i < 5; ++i) for (int j = 0; j < 5; ++j) Ainv[5 * i + j] = a[i][5 + j];
    return true;
}

// Block ILU(0) on the cell-ordered sparsity of the LDU matrix.  ILU(0)
// factors are unique for a fixed ordering and pattern, so this is the
// factor of foam-extend's ILUC0 (natural order) up to rounding.
struct ILU {
    int n = 0;
    std::vector<int> rowPtr, col;     // full pattern, sorted columns per row
    std::vector<int> diagPos;
    std::vector<double> val;          // [nnz][25]; L strictly lower (unit diag), U upper, diag inverted stored separately
    std::vector<double> dinv;         // [n][25]
};

static void buildILU(const Case& c, ILU& il) {
    const Mesh& m = c.m;
    int n = m.nCells;
    il.n = n;
    std::vector<std::vector<std::pair<int, const double*>>> rows(n);
    for (int i = 0; i < n; ++i) rows[i].push_back({i, &c.diag[25 * i]});
    const bool blockJacobi = c.st.ilu_order == RDST_ILU_PARTITIONED;
    if (blockJacobi && (int)c.part.size() != n) throw std::runtime_error("partitioned ILU needs a cell partition");
    for (int f = 0; f < m.nInternal; ++f) {
        int o = m.owner[f], nb = m.neighbour[f];
        // block-Jacobi: the factor is the ILU(0) of the partition-diagonal blocks
        if (blockJacobi && c.part[o] != c.part[nb]) continue;
        rows[o].push_back({nb, &c.upper[25 * f]});   // owner row, neighbour column
        rows[nb].push_back({o, &c.lower[25 * f]});   // neighbour row, owner column
    }
    il.rowPtr.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        std::sort(rows[i].begin(), rows[i].end(), [](const std::pair<int, const double*>& a, const std::pair<int, const double*>& b) { return a.first < b.first; });
        il.rowPtr[i + 1] = il.rowPtr[i] + (int)rows[i].size();
    }
    il.col.resize(il.rowPtr[n]);
    il.val.resize(25 * (size_t)il.rowPtr[n]);
    il.diagPos.resize(n);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < (int)rows[i].size(); ++k) {
            int pos = il.rowPtr[i] + k;
            il.col[pos] = rows[i][k].first;
            if (rows[i][k].first == i) il.diagPos[i] = pos;
            std::memcpy(&il.val[25 * (size_t)pos], rows[i][k].second, 25 * sizeof(double));
        }
    // IKJ ILU(0)
    il.dinv.assign(25 * (size_t)n, 0.0);
    std::vector<int> where(n, -1);
    for (int i = 0; i < n; ++i) {
        for (int p = il.rowPtr[i]; p < il.rowPtr[i + 1]; ++p) where[il.col[p]] = p;
        for (int p = il.rowPtr[i]; p < il.rowPtr[i + 1]; ++p) {
            int k = il.col[p];
            if (k >= i) break;
            // L_ik = A_ik * inv(U_kk)
            double t[25];
            matmul5(&il.val[25 * (size_t)p], &il.dinv[25 * (size_t)k], t);
            std::memcpy(&il.val[25 * (size_t)p], t, sizeof t);
            for (int q = il.diagPos[k] + 1; q < il.rowPtr[k + 1]; ++q) {
                int j = il.col[q];
                int pj = where[j];
                if (pj < 0) continue;
                double prod[25];
                matmul5(t, &il.val[25 * (size_t)q], prod);
                for (int e = 0; e < 25; ++e) il.val[25 * (size_t)pj + e] -= prod[e];
            }
        }
        if (!inv5(&il.val[25 * (size_t)il.diagPos[i]], &il.dinv[25 * (size_t)i]))
            throw std::runtime_error("singular ILU pivot block");
        for (int p = il.rowPtr[i]; p < il.rowPtr[i + 1]; ++p) where[il.col[p]] = -1;
    }
}

static void mv5(const double* A, const double* x, double* y) {
    for (int r = 0; r < 5; ++r) {
        double s = 0;
        for (int k = 0; k < 5; ++k) s += A[5 * r + k] * x[k];
        y[r] = s;
    }
}

static void iluApply(const ILU& il, const std::vector<double>& r, std::vector<double>& x) {
    int n = il.n;
    std::vector<double> y(r);
    for (int i = 0; i < n; ++i) {
        for (int p = il.rowPtr[i]; p < il.diagPos[i]; ++p) {
            double t[5];
            mv5(&il.val[25 * (size_t)p], &y[5 * (size_t)il.col[p]], t);
            for (int k = 0; k < 5; ++k) y[5 * i + k] -= t[k];
        }
    }
    x.assign(5 * (size_t)n, 0.0);
    for (int i = n - 1; i >= 0; --i) {
        double s[5];
        for (int k = 0; k < 5; ++k) s[k] = y[5 * i + k];
        for (int p = il.diagPos[i] + 1; p < il.rowPtr[i + 1]; ++p) {
            double t[5];
            mv5(&il.val[25 * (size_t)p], &x[5 * (size_t)il.col[p]], t);
            for (int k = 0; k < 5; ++k) s[k] -= t[k];
        }
        mv5(&il.dinv[25 * (size_t)i], s, &x[5 * i]);
    }
}

struct Precond {
    int kind;
    const Case* c;
    ILU il;
    std::vector<double> dinv;
    void init(const Case& cc, int k) {
        c = &cc; kind = k;
        if (k == RDST_PRECOND_ILU0) buildILU(cc, il);
        else if (k == RDST_PRECOND_DIAG) {
            dinv.resize(cc.diag.size());
            for (int i = 0; i < cc.m.nCells; ++i) inv5(&cc.diag[25 * i], &dinv[25 * i]);
        }
    }
    void apply(std::vector<double>& out, const std::vector<double>& in) const {
        if (kind == RDST_PRECOND_ILU0) iluApply(il, in, out);
        else if (kind == RDST_PRECOND_DIAG) {
            out.resize(in.size());
            for (int i = 0; i < c->m.nCells; ++i) mv5(&dinv[25 * i], &in[5 * i], &out[5 * i]);
        } else out = in;
    }
};

static double sumProd(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}
// gSum(cmptMag(r)) per component
static void cmptMagSum(const std::vector<double>& r, double* out) {
    for (int k = 0; k < 5; ++k) out[k] = 0;
    for (size_t i = 0; i < r.size() / 5; ++i)
        for (int k = 0; k < 5; ++k) out[k] += std::fabs(r[5 * i + k]);
}
static double mag5(const double* v) {
    double s = 0;
    for (int k = 0; k < 5; ++k) s += v[k] * v[k];
    return std::sqrt(s);
}

// BlockLduSolver::normFactor
static double normFactor(const Case& c, const std::vector<double>& x, const std::vector<double>& b) {
    int n = c.m.nCells;
    double xRef[5] = {0};
    for (int i = 0; i < n; ++i) for (int k = 0; k < 5; ++k) xRef[k] += x[5 * i + k];
    for (int k = 0; k < 5; ++k) xRef[k] /= n;
    std::vector<double> wA, pA, xr(5 * (size_t)n);
    Amul(c, x, wA);
    for (int i = 0; i < n; ++i) for (int k = 0; k < 5; ++k) xr[5 * i + k] = xRef[k];
    Amul(c, xr, pA);
    double s = 0;
    for (int i = 0; i < n; ++i) {
        double d1[5], d2[5];
        for (int k = 0; k < 5; ++k) { d1[k] = wA[5 * i + k] - pA[5 * i + k]; d2[k] = b[5 * i + k] - pA[5 * i + k]; }
        s += mag5(d1) + mag5(d2);
    }
    return s + BLOCK_SMALL;
}

struct Perf { double init[5], fin[5]; int nIter; bool conv; };

static bool checkConvergence(const rdst_settings& st, Perf& p) {
    double fm = mag5(p.fin), im = mag5(p.init);
    p.conv = (fm < st.tolerance) || (st.rel_tol > SMALL && fm <= st.rel_tol * im);
    return p.conv;
}
static bool converged(const rdst_settings& st, Perf& p) {
    return p.nIter >= st.min_iter && (p.nIter >= st.max_iter || checkConvergence(st, p));
}

// Givens rotation as in foam-extend's GMRES
static void givens(double h, double beta, double& cs, double& sn) {
    if (beta == 0) { cs = 1; sn = 0; }
    else if (std::fabs(beta) > std::fabs(h)) { double tau = -h / beta; sn = 1.0 / std::sqrt(1.0 + tau * tau); cs = sn * tau; }
    else { double tau = -beta / h; cs = 1.0 / std::sqrt(1.0 + tau * tau); sn = cs * tau; }
}

// foam-extend BlockGMRESSolver<vector5>::solve: left-preconditioned
// restarted GMRES(nDirections), modified Gram-Schmidt, one "iteration" per restart
static Perf gmres(const Case& c, const Precond& M, std::vector<double>& x, const std::vector<double>& b) {
    const rdst_settings& st = c.st;
    size_t N5 = x.size();
    int nDirs = st.n_directions;
    Perf perf = {};
    double norm = normFactor(c, x, b);
    std::vector<double> wA, rA(N5);
    Amul(c, x, wA);
    for (size_t i = 0; i < N5; ++i) rA[i] = b[i] - wA[i];
    cmptMagSum(rA, perf.init);
    for (int k = 0; k < 5; ++k) { perf.init[k] /= norm; perf.fin[k] = perf.init[k]; }
    if (!converged(st, perf)) {
        std::vector<double> H((size_t)nDirs * nDirs, 0.0), yh(nDirs, 0.0), bh(nDirs + 1, 0.0), cs(nDirs, 0.0), sn(nDirs, 0.0);
        std::vector<std::vector<double>> Vk(nDirs + 1, std::vector<double>(N5, 0.0));
        do {
            M.apply(wA, rA);
            double beta = std::sqrt(sumProd(wA, wA));
            std::fill(bh.begin(), bh.end(), 0.0);
            bh[0] = beta;
            for (int i = 0; i < nDirs; ++i) {
                for (size_t q = 0; q < N5; ++q) Vk[i][q] = wA[q] / beta;
                Amul(c, Vk[i], rA);
                M.apply(wA, rA);
                if (st.gram_schmidt == RDST_GS_CLASSICAL) {
                    // classical Gram-Schmidt: all projections of the same w, then one update
                    for (int j = 0; j <= i; ++j) H[(size_t)j * nDirs + i] = sumProd(wA, Vk[j]);
                    for (int j = 0; j <= i; ++j)
                        for (size_t q = 0; q < N5; ++q) wA[q] -= H[(size_t)j * nDirs + i] * Vk[j][q];
                } else {
                    // modified Gram-Schmidt, as foam-extend's GMRES
                    for (int j = 0; j <= i; ++j) {
                        beta = sumProd(wA, Vk[j]);
                        H[(size_t)j * nDirs + i] = beta;
                        for (size_t q = 0; q < N5; ++q) wA[q] -= beta * Vk[j][q];
                    }
                }
                beta = std::sqrt(sumProd(wA, wA));
                for (int j = 0; j < i; ++j) {
                    double Hji = H[(size_t)j * nDirs + i];
                    H[(size_t)j * nDirs + i] = cs[j] * Hji - sn[j] * H[(size_t)(j + 1) * nDirs + i];
                    H[(size_t)(j + 1) * nDirs + i] = sn[j] * Hji + cs[j] * H[(size_t)(j + 1) * nDirs + i];
                }
                givens(H[(size_t)i * nDirs + i], beta, cs[i], sn[i]);
                double bhi = bh[i];
                bh[i] = cs[i] * bhi - sn[i] * bh[i + 1];
                bh[i + 1] = sn[i] * bhi + cs[i] * bh[i + 1];
                H[(size_t)i * nDirs + i] = cs[i] * H[(size_t)i * nDirs + i] - sn[i] * beta;
            }
            for (int i = nDirs - 1; i >= 0; --i) {
                double s = bh[i];
                for (int j = i + 1; j < nDirs; ++j) s -= H[(size_t)i * nDirs + j] * yh[j];
                yh[i] = s / H[(size_t)i * nDirs + i];
            }
            for (int i = 0; i < nDirs; ++i)
                for (size_t q = 0; q < N5; ++q) x[q] += yh[i] * Vk[i][q];
            Amul(c, x, wA);
            for (size_t q = 0; q < N5; ++q) rA[q] = b[q] - wA[q];
            cmptMagSum(rA, perf.fin);
            for (int k = 0; k < 5; ++k) perf.fin[k] /= norm;
            perf.nIter++;
        } while (!converged(st, perf));
    }
    return perf;
}

// foam-extend BlockBiCGStabSolver<vector5>::solve (right-preconditioned)
static Perf bicgstab(const Case& c, const Precond& M, std::vector<double>& x, const std::vector<double>& b) {
    const rdst_settings& st = c.st;
    size_t N5 = x.size();
    Perf perf = {};
    double norm = normFactor(c, x, b);
    std::vector<double> p, r(N5);
    Amul(c, x, p);
    for (size_t i = 0; i < N5; ++i) r[i] = b[i] - p[i];
    cmptMagSum(r, perf.init);
    for (int k = 0; k < 5; ++k) { perf.init[k] /= norm; perf.fin[k] = perf.init[k]; }
    if (!converged(st, perf)) {
        double rho = BLOCK_GREAT, rhoOld = rho, alpha = 0, omega = BLOCK_GREAT, beta;
        p.assign(N5, 0.0);
        std::vector<double> ph(N5, 0.0), v(N5, 0.0), t(N5, 0.0), th(N5, 0.0), rw(r);
        do {
            rhoOld = rho;
            rho = sumProd(rw, r);
            beta = rho / rhoOld * (alpha / omega);
            for (size_t i = 0; i < N5; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
            M.apply(ph, p);
            Amul(c, ph, v);
            alpha = rho / sumProd(rw, v);
            for (size_t i = 0; i < N5; ++i) r[i] -= alpha * v[i];
            M.apply(th, r);
            Amul(c, th, t);
            omega = sumProd(t, r) / sumProd(t, t);
            for (size_t i = 0; i < N5; ++i) x[i] += alpha * ph[i] + omega * th[i];
            for (size_t i = 0; i < N5; ++i) r[i] -= omega * t[i];
            cmptMagSum(r, perf.fin);
            for (int k = 0; k < 5; ++k) perf.fin[k] /= norm;
            perf.nIter++;
        } while (!converged(st, perf));
    }
    return perf;
}

// fvBlockMatrixDST::solve (fvBlockMatrixDST.C:98-155)
static Perf solve(Case& c) {
    int n = c.m.nCells;
    Precond M;
    M.init(c, c.st.precond);
    Perf perf;
    if (c.st.dependent == RDST_DEP_W) {
        std::vector<double> sourceW;
        Amul(c, c.W, sourceW);
        for (int i = 0; i < 5 * n; ++i) sourceW[i] += c.source[i];
        perf = (c.st.solver == RDST_SOLVER_GMRES) ? gmres(c, M, c.W, sourceW) : bicgstab(c, M, c.W, sourceW);
    } else {
        std::vector<double> dW(5 * (size_t)n, 0.0);
        perf = (c.st.solver == RDST_SOLVER_GMRES) ? gmres(c, M, dW, c.source) : bicgstab(c, M, dW, c.source);
        for (int i = 0; i < 5 * n; ++i) c.W[i] += dW[i];
    }
    return perf;
}

// ------------------------------------------------------------------ updateFields.H
static inline double clampv(double v, double lo, double hi) { return std::min(std::max(v, lo), hi); }

static void updateFields(Case& c) {
    Mesh& m = c.m;
    int n = m.nCells, nB = m.nB;
    double Cp = c.th.Cp, Cv = c.Cv, gamma = Cp / Cv;
    // rho.oldTime() of the turbulence ddt: the density before this update
    c.tb.rhoPrev = c.rho;
    c.tb.rhoPrevValid = true;
    // old values the BCs read before thermo.correct()
    std::vector<double> pOld = c.p, TOld = c.T, psiOld = c.psi;
    std::vector<double> bpsiOld = c.bpsi, brhoOld = c.brho;
    // update limit (an extension, off in the reference): scale each cell's
    // change of W so that rho and rhoE move by at most update_limit
    if (c.st.update_limit > 0.0) {
        const double lim = c.st.update_limit;
        for (int i = 0; i < n; ++i) {
            const double o[5] = {c.rho[i], c.rhoU[i].x, c.rhoU[i].y, c.rhoU[i].z, c.rhoE[i]};
            const double dr = std::fabs(c.W[5 * i] - o[0]), de = std::fabs(c.W[5 * i + 4] - o[4]);
            double a = 1.0;
            if (dr > lim * o[0]) a = std::fmin(a, lim * o[0] / dr);
            if (de > lim * o[4]) a = std::fmin(a, lim * o[4] / de);
            if (a < 1.0)
                for (int q = 0; q < 5; ++q) c.W[5 * i + q] = o[q] + a * (c.W[5 * i + q] - o[q]);
        }
    }
    // DST::retrieve (updateFields.H:40-42), boundMinMax(rho) as a clamp
    for (int i = 0; i < n; ++i) {
        c.rho[i] = clampv(c.W[5 * i], c.st.rho_min, c.st.rho_max);
        c.rhoU[i] = v3(c.W[5 * i + 1], c.W[5 * i + 2], c.W[5 * i + 3]);
        c.rhoE[i] = c.W[5 * i + 4];
        c.U[i] = c.rhoU[i] / c.rho[i];
    }
    // U.correctBoundaryConditions()
    std::vector<double> TrefH(nB, 0.0);
    for (int b = 0; b < nB; ++b) {
        const rdst_bc& bc = c.bc[m.bPatch[b]];
        int ci = m.bCell[b], f = m.bFace[b];
        V3 nrm = m.Sf[f] / m.magSf[f];
        double phib = c.phi[m.nInternal + b];
        V3 UI = c.U[ci];
        switch (bc.U_type) {
            case RDST_BC_FIXED_VALUE: c.bU[b] = v3(bc.U_value[0], bc.U_value[1], bc.U_value[2]); break;
            case RDST_BC_ZERO_GRADIENT: c.bU[b] = UI; break;
            case RDST_BC_SYMMETRY: c.bU[b] = symmetricValue(nrm, UI); break;
            case RDST_BC_INLET_OUTLET: {
                double vf = 1.0 - fe_pos(phib);
                c.bcs.vf_U[b] = vf;
                V3 ref = v3(bc.U_value[0], bc.U_value[1], bc.U_value[2]);
                c.bcs.refU[b] = ref;
                c.bU[b] = vf * ref + (1.0 - vf) * UI;
                break;
            }
            case RDST_BC_FARFIELD: {
                // farFieldVelocityDST::updateCoeffs (U.C:144-217)
                double cb = std::sqrt(Cp / Cv / bpsiOld[b]);
                double cI = std::sqrt(Cp / Cv / psiOld[ci]);
                double mach = phib / brhoOld[b] / m.magSf[f] / cb;
                double supIn = neg0(mach + 1), supOut = pos0(mach - 1);
                double subIn = neg(mach) - supIn, subOut = pos0(mach) - supOut;
                double rho0c0 = c.rho[ci] * cI;
                V3 Uinf = v3(bc.U_inf[0], bc.U_inf[1], bc.U_inf[2]);
                double pb = 0.5 * (bc.p_inf + pOld[ci] - rho0c0 * dot(nrm, Uinf - UI));
                V3 subsInValue = Uinf - nrm * (bc.p_inf - pb) / rho0c0;
                V3 subsOutValue = nrm * (pOld[ci] - pb) / rho0c0;
                V3 ref = supIn * Uinf + subIn * subsInValue + subOut * subsOutValue;
                double vf = 1 - supOut - subOut;
                c.bcs.vf_U[b] = vf;
                c.bcs.refU[b] = ref;
                c.bU[b] = vf * ref + (1.0 - vf) * UI;
                break;
            }
            default: break;
        }
        // T boundary coefficients evaluated inside h.correctBoundaryConditions()
        // (mixedEnthalpy calls Tw.evaluate() with the not-yet-corrected T, p)
        if (bc.T_type == RDST_BC_INLET_OUTLET) {
            c.bcs.vf_T[b] = 1.0 - fe_pos(phib);
            TrefH[b] = bc.T_value;
        } else if (bc.T_type == RDST_BC_FARFIELD) {
            // farFieldTemperatureDST::updateCoeffs (T.C:144-221)
            double cb = std::sqrt(Cp / Cv / bpsiOld[b]);
            double cI = std::sqrt(Cp / Cv / psiOld[ci]);
            double mach = phib / brhoOld[b] / m.magSf[f] / cb;
            double supIn = neg0(mach + 1), supOut = pos0(mach - 1);
            double subIn = neg(mach) - supIn, subOut = pos0(mach) - supOut;
            double rho0c0 = c.rho[ci] * cI;
            V3 Uinf = v3(bc.U_inf[0], bc.U_inf[1], bc.U_inf[2]);
            double pI = pOld[ci], TI = TOld[ci];
            double pb = 0.5 * (bc.p_inf + pI - rho0c0 * dot(nrm, Uinf - UI));
            double RI = pI / c.rho[ci] / TI;
            double rhoa = bc.p_inf / bc.T_inf / RI;
            double subsInValue = pb / (rhoa + (pb - bc.p_inf) / sqr(cI)) / RI;
            double subsOutVF = pb / (pI + RI * TI * (pb - pI) / sqr(cI));
            TrefH[b] = supIn * bc.T_inf + subIn * subsInValue;
            c.bcs.vf_T[b] = 1 - supOut - subsOutVF * subOut;
        }
    }
    // h = Cp/Cv*(rhoE/rho - 0.5|U|^2)  (updateFields.H:60)
    for (int i = 0; i < n; ++i) c.h[i] = gamma * (c.rhoE[i] / c.rho[i] - 0.5 * magSqr(c.U[i]));
    // h.correctBoundaryConditions() via the enthalpy BCs derived from T's
    for (int b = 0; b < nB; ++b) {
        const rdst_bc& bc = c.bc[m.bPatch[b]];
        int ci = m.bCell[b];
        switch (bc.T_type) {
            case RDST_BC_FIXED_VALUE: c.bh[b] = Cp * bc.T_value; break;
            case RDST_BC_INLET_OUTLET: case RDST_BC_FARFIELD: {
                double vf = c.bcs.vf_T[b];
                c.bh[b] = vf * (Cp * TrefH[b]) + (1.0 - vf) * c.h[ci];
                break;
            }
            default: c.bh[b] = c.h[ci];
        }
    }
    // boundMinMax(h, CpMin*TMin, CpMax*TMax)
    double hMin = Cp * c.st.T_min, hMax = Cp * c.st.T_max;
    for (int i = 0; i < n; ++i) c.h[i] = clampv(c.h[i], hMin, hMax);
    // thermo.correct(): T = h/Cp, psi = 1/(R T), mu
    for (int i = 0; i < n; ++i) {
        c.T[i] = c.h[i] / Cp;
        c.psi[i] = 1.0 / (c.th.R * c.T[i]);
        c.mu[i] = muOf(c, c.T[i]);
        c.alpha[i] = alphaOf(c, c.T[i]);
    }
    for (int b = 0; b < nB; ++b) {
        const rdst_bc& bc = c.bc[m.bPatch[b]];
        if (bc.T_type == RDST_BC_FIXED_VALUE) {
            c.bT[b] = bc.T_value;
            c.bh[b] = Cp * c.bT[b];
        } else {
            c.bT[b] = c.bh[b] / Cp;
        }
        c.bpsi[b] = 1.0 / (c.th.R * c.bT[b]);
        c.bmu[b] = muOf(c, c.bT[b]);
        c.balpha[b] = alphaOf(c, c.bT[b]);
    }
    // p = rho/psi; bound; p.correctBoundaryConditions()
    for (int i = 0; i < n; ++i) c.p[i] = clampv(c.rho[i] / c.psi[i], c.st.p_min, c.st.p_max);
    for (int b = 0; b < nB; ++b) {
        const rdst_bc& bc = c.bc[m.bPatch[b]];
        int ci = m.bCell[b], f = m.bFace[b];
        double phib = c.phi[m.nInternal + b];
        switch (bc.p_type) {
            case RDST_BC_FIXED_VALUE: c.bp[b] = bc.p_value; break;
            case RDST_BC_INLET_OUTLET: {
                double vf = 1.0 - fe_pos(phib);
                c.bcs.vf_p[b] = vf;
                c.bp[b] = vf * bc.p_value + (1.0 - vf) * c.p[ci];
                break;
            }
            case RDST_BC_FARFIELD: {
                // farFieldPressureDST::updateCoeffs (P.C:144-215)
                V3 nrm = m.Sf[f] / m.magSf[f];
                double cb = std::sqrt(Cp / Cv / c.bpsi[b]);
                double cI = std::sqrt(Cp / Cv / c.psi[ci]);
                double mach = phib / brhoOld[b] / m.magSf[f] / cb;
                double supIn = neg0(mach + 1), supOut = pos0(mach - 1);
                double subIn = neg(mach) - supIn, subOut = pos0(mach) - supOut;
                double rho0c0 = c.rho[ci] * cI;
                V3 Uinf = v3(bc.U_inf[0], bc.U_inf[1], bc.U_inf[2]);
                double pb = bc.p_inf - rho0c0 * dot(nrm, Uinf - c.U[ci]);
                double ref = supIn * bc.p_inf + subIn * pb + subOut * bc.p_inf;
                double vf = 1 - supOut - 0.5 * subIn;
                c.bcs.vf_p[b] = vf;
                c.bp[b] = vf * ref + (1.0 - vf) * c.p[ci];
                break;
            }
            default: c.bp[b] = c.p[ci];
        }
        c.bp[b] = clampv(c.bp[b], c.st.p_min, c.st.p_max);
        // rho_b = p_b psi_b, rhoU_b = rho_b U_b, rhoE_b = rho_b (h_b + 0.5|U_b|^2) - p_b
        c.brho[b] = c.bp[b] * c.bpsi[b];
        c.brhoU[b] = c.brho[b] * c.bU[b];
        c.brhoE[b] = c.brho[b] * (c.bh[b] + 0.5 * magSqr(c.bU[b])) - c.bp[b];
    }
}

// ------------------------------------------------------------------ timeDST
// setSteadyStateDeltaT (timeDST.C:218-273)
static void setSteadyStateDeltaT(Case& c) {
    const Mesh& m = c.m;
    int n = m.nCells;
    double gam = c.th.Cp / c.Cv;
    if (!c.st.neighbouring) {
        for (int i = 0; i < n; ++i) {
            double cc = std::sqrt(c.th.Cp / c.Cv / c.psi[i]);
            double lam = (std::fabs(c.U[i].x) + cc) * c.dS[3 * i] + (std::fabs(c.U[i].y) + cc) * c.dS[3 * i + 1]
                       + (std::fabs(c.U[i].z) + cc) * c.dS[3 * i + 2];
            c.rDeltaT[i] = lam / (c.courantScalar * m.V[i]);
        }
        return;
    }
    std::vector<V3> Uf;
    linearInterpV(m, c.U, c.bU, Uf);
    std::vector<double> cc(n), bcc(m.nB), cf;
    for (int i = 0; i < n; ++i) cc[i] = std::sqrt(gam / c.psi[i]);
    for (int b = 0; b < m.nB; ++b) bcc[b] = std::sqrt(gam / c.bpsi[b]);
    linearInterp(m, cc, bcc, cf);
    for (int i = 0; i < n; ++i) c.rDeltaT[i] = 0.0;
    for (int f = 0; f < m.nInternal; ++f) {
        double dtSf = (mag(Uf[f]) + cf[f]) * m.dc[f];
        int o = m.owner[f], nb = m.neighbour[f];
        c.rDeltaT[o] = std::max(c.rDeltaT[o], dtSf);
        c.rDeltaT[nb] = std::max(c.rDeltaT[nb], dtSf);
    }
    for (int b = 0; b < m.nB; ++b) {
        if (m.patches[m.bPatch[b]].type != RDST_PATCH_WALL) continue;
        int ci = m.bCell[b];
        double pLambda = 0.5 * m.dc[m.nInternal + b] * (std::sqrt(gam / c.psi[ci])) + mag(c.U[ci]);
        c.rDeltaT[ci] = std::max(c.rDeltaT[ci], pLambda);
    }
    for (int i = 0; i < n; ++i) c.rDeltaT[i] /= c.courant[i];
}

// setTransientDeltaT (timeDST.C:275-314): the physical deltaT, the minimum over
// the internal faces of |Sf| / ((|U_f . Sf| + |Sf| c_f) deltaCoeffs) courant
// (neighbouring) or over the cells of courant V / ((|U| + c) . dS)
static void setTransientDeltaT(Case& c) {
    const Mesh& m = c.m;
    int n = m.nCells;
    double dt = 1e300;
    if (!c.st.neighbouring) {
        for (int i = 0; i < n; ++i) {
            double cc = std::sqrt(c.th.Cp / c.Cv / c.psi[i]);
            double lam = (std::fabs(c.U[i].x) + cc) * c.dS[3 * i] + (std::fabs(c.U[i].y) + cc) * c.dS[3 * i + 1]
                       + (std::fabs(c.U[i].z) + cc) * c.dS[3 * i + 2];
            dt = std::min(dt, c.courantScalar * m.V[i] / lam);
        }
    } else {
        double gam = c.th.Cp / c.Cv;
        std::vector<V3> Uf;
        linearInterpV(m, c.U, c.bU, Uf);
        std::vector<double> cc(n), bcc(m.nB), cf;
        for (int i = 0; i < n; ++i) cc[i] = std::sqrt(gam / c.psi[i]);
        for (int b = 0; b < m.nB; ++b) bcc[b] = std::sqrt(gam / c.bpsi[b]);
        linearInterp(m, cc, bcc, cf);
        for (int f = 0; f < m.nInternal; ++f) {
            double amax = std::fabs(dot(Uf[f], m.Sf[f])) + m.magSf[f] * cf[f];
            double dtSf = m.magSf[f] / (amax * m.dc[f]);
            dt = std::min(dt, dtSf * c.courantScalar);
        }
    }
    c.deltaT = dt;
}

// setCourantDST (timeDST.C:316-443).  residualR: |solverPerf.initialResidual()|
// when lin is given (the inner iterations of Dual), else |R|_2.  pseudo: the
// pseudo-time branch (steady state, or dual time while the inner iterations
// are not converged): ramp of the per-cell courant and setSteadyStateDeltaT;
// otherwise the physical branch: with adjustTimeStep the scalar courant
// follows the residual ratio and setTransientDeltaT sets the next deltaT.
static double setCourantDST(Case& c, const double* lin, bool pseudo) {
    const Mesh& m = c.m;
    int n = m.nCells;
    double residualR;
    if (lin) {
        residualR = mag5(lin);
    } else {
        double ss = 0;
        for (int i = 0; i < n; ++i) {
            double mg = mag5(&c.R[5 * i]);
            ss += mg * mg;
        }
        residualR = std::sqrt(ss);
    }
    if (c.residual != -1) c.residualRatio = std::min(std::max(c.residual / residualR, 0.1), 2.0);
    c.residual = residualR;
    if (!pseudo) {
        if (c.st.adjust_time_step) {
            c.courantScalar = std::min(std::max(c.courantScalar * c.residualRatio, c.st.min_co), c.st.max_co);
            setTransientDeltaT(c);
        }
        return residualR;
    }
    std::vector<double> eNew(n), factor(n, 1.0);
    for (int i = 0; i < n; ++i) eNew[i] = c.rhoE[i] / c.rho[i] - 0.5 * magSqr(c.rhoU[i] / c.rho[i]);
    for (int f = 0; f < m.nInternal; ++f) {
        int o = m.owner[f], nb = m.neighbour[f];
        if (c.rho[o] < 0.95 * c.rhoOld[o] || eNew[o] < 0.95 * c.eOld[o] || eNew[o] < SMALL) {
            factor[o] = std::min(factor[o], 0.5);
            factor[nb] = std::min(factor[nb], 0.75);
        }
        if (c.rho[nb] < 0.95 * c.rhoOld[nb] || eNew[nb] < 0.95 * c.eOld[nb] || eNew[nb] < SMALL) {
            factor[nb] = std::min(factor[nb], 0.5);
            factor[o] = std::min(factor[o], 0.75);
        }
    }
    c.rhoOld = c.rho;
    c.eOld = eNew;
    for (int i = 0; i < n; ++i) {
        c.courant[i] *= c.residualRatio * factor[i];
        c.courant[i] = std::min(std::max(c.courant[i], c.st.min_co), c.st.max_co);
    }
    setSteadyStateDeltaT(c);
    return residualR;
}

static double setCourant(Case& c) { return setCourantDST(c, nullptr, true); }

// Single::iterate (Single.C:80-104)
static void saCorrect(Case& c, Perf* perf);
static double iterate(Case& c, Perf* perfOut) {
    computeFlux(c);
    computeJacobian(c);
    assemble(c);
    Perf perf = solve(c);
    if (perfOut) *perfOut = perf;
    updateFields(c);
    if (c.tb.on) saCorrect(c, nullptr);
    return setCourant(c);
}

// residualControls + steadyCheck (residualControls.C:37-64, steadyCheck.C:37-98):
// the inner residual of Dual is the mean of the five initial residuals of
// the inner solve; converged when more than maxInnerIter inner iterations
// ran, the residual fell below minResudual, or (from iterationStart on) the
// scaled slope of log(residual) over the last 20 % of the history fell below
// magConvergedSlope.  converged() refreshes the slope; the const form (used
// by setCourantDST) reads the last one.
struct SteadyCheck {
    std::vector<double> residuals;
    int iterationCount = 0, iterationStart = 10, maxInnerIter = 20;
    double residual = 1, magScaledSlope = 1, magConvergedSlope = 0.1, minResidual = 1e-5;
    void addResidual(const double* res) {
        double meanResidual = 0;
        for (int i = 0; i < 5; ++i) meanResidual += res[i];
        meanResidual /= 5;
        iterationCount++;
        residual = meanResidual;
        residuals.push_back(meanResidual);
    }
    bool convergedConst() const {
        return iterationCount > maxInnerIter || (residual < minResidual || magScaledSlope < magConvergedSlope);
    }
    bool converged() {
        if (iterationCount >= iterationStart) {
            double sumXi = 0, sumYi = 0, sumXiSqr = 0, sumYiXi = 0;
            int size = (int)residuals.size();
            int N = 0.2 * size;
            double maxY = std::log(*std::max_element(residuals.begin(), residuals.end()));
            double minY = std::log(*std::min_element(residuals.begin(), residuals.end()));
            double deltaY = maxY - minY;
            for (int i = 0; i < N; i++) {
                int xi = size - 1 - i;
                double yi = std::log(residuals[xi]);
                sumXi += xi;
                sumXiSqr += std::pow(xi, 2);
                sumYi += yi;
                sumYiXi += yi * xi;
            }
            magScaledSlope = (size - 1) / deltaY *
                             std::fabs((sumYiXi - sumXi * sumYi / N) / (sumXiSqr - std::pow(sumXi, 2) / N));
        }
        return convergedConst();
    }
};

static SteadyCheck makeSteadyCheck(const rdst_settings& st) {
    SteadyCheck sc;
    sc.iterationStart = st.iteration_start;
    sc.maxInnerIter = st.max_inner_iter;
    sc.magConvergedSlope = st.mag_converged_slope;
    sc.minResidual = st.min_residual;
    return sc;
}

// runTime++ (rhoDST.C:73; foam Time: deltaT0 = the deltaT of the previous
// step, the controlDict deltaT before the first) and Dual::iterate
// (Dual.C:80-124) for a laminar case
static double dualStep(Case& c, int* nInner, double* innerRes) {
    const int n = c.m.nCells;
    c.deltaT0 = c.deltaTSave;
    c.deltaTSave = c.deltaT;
    c.time += c.deltaT;
    c.timeIndex++;
    std::vector<double> Wc(5 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        Wc[5 * i] = c.rho[i]; Wc[5 * i + 1] = c.rhoU[i].x; Wc[5 * i + 2] = c.rhoU[i].y; Wc[5 * i + 3] = c.rhoU[i].z;
        Wc[5 * i + 4] = c.rhoE[i];
    }
    c.WOldOld = (c.timeIndex == 1) ? Wc : c.WOld;  // oldTime().oldTime() of the first step is oldTime()
    c.WOld = Wc;
    SteadyCheck sc = makeSteadyCheck(c.st);
    int k = 0;
    do {
        computeFlux(c);
        computeJacobian(c);
        assemble(c);
        Perf perf = solve(c);
        updateFields(c);
        sc.addResidual(perf.init);
        if (innerRes) innerRes[k] = sc.residual;
        ++k;
        setCourantDST(c, perf.init, !sc.convergedConst());
    } while (!sc.converged());
    if (nInner) *nInner = k;
    return setCourantDST(c, nullptr, !sc.convergedConst());
}


// ======================================================================= Spalart-Allmaras
// SAneg (libraries/turbulence/SAneg/SAneg.C) and SARC2
// (libraries/turbulence/SARC2/SARC2.C) correct(), field-at-a-time.
// foam-extend helpers restated: pow3(s) = s*sqr(s), pow6(s) = pow3(sqr(s)),
// pos(s) = s >= 0, neg(s) = s < 0.
static inline double pow3(double s) { return s * sqr(s); }
static inline double pow6(double s) { return pow3(sqr(s)); }

struct SAField {
    std::vector<double> chi, fv1, fv2, fv3, Omega, Stilda, fw, ft2, fr1, Gamma, bGamma;
};

// chi = rho nuTilda / mu (SAneg.C:62-65), fv1 (:68-72), fv2 (:75-83), fv3 (:86-100)
static inline double saFv1(const rdst_turbulence& t, double chi) {
    double chi3 = pow3(chi);
    return chi3 / (chi3 + pow3(t.Cv1));
}
static inline double saFv2(const rdst_turbulence& t, double chi) { return 1.0 / pow3(1.0 + chi / t.Cv2); }
static inline double saFv3(const rdst_turbulence& t, double chi, double fv1) {
    double cb = (1.0 / t.Cv2) * chi;
    return (1.0 + chi * fv1) * (1.0 / t.Cv2) * (3.0 * (1.0 + cb) + sqr(cb)) / pow3(1.0 + cb);
}
// fw (SAneg.C:103-120; SARC2.C:103-118)
static inline double saFw(const rdst_turbulence& t, double Stilda, double nut, double d) {
    double r = std::min(nut / (std::max(Stilda, SMALL) * sqr(t.kappa * d)), 10.0);
    double g = r + t.Cw2 * (pow6(r) - r);
    return g * std::pow((1.0 + pow6(t.Cw3)) / (pow6(g) + pow6(t.Cw3)), 1.0 / 6.0);
}
static inline double saCw1(const rdst_turbulence& t) { return t.Cb1 / sqr(t.kappa) + (1.0 + t.Cb2) / t.sigma_nut; }
// DnuTildaEff: SAneg (SAneg.C:375-393) (rho nuTilda fn + mu)/sigmaNut with
// fn = pos(chi) + neg(chi)(Cn1 + chi^3)/(Cn1 - chi^3); SARC2 (SARC2.H:160-170)
// rho nuTilda/sigmaNut + mu
static inline double saGamma(const rdst_turbulence& t, double rho, double nut, double mu) {
    if (t.model == RDST_TURB_SANEG) {
        double chi = rho * nut / mu;
        double p3 = std::pow(chi, 3);
        double fn = pos0(chi) + neg(chi) * (t.Cn1 + p3) / (t.Cn1 - p3);
        return (rho * nut * fn + mu) / t.sigma_nut;
    }
    return rho * nut / t.sigma_nut + mu;
}
// magSqr of skew(G) and symm(G), G[3i+j] = dU_j/dx_i
static inline double magSqrSkew(const T9& g) {
    double s = 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) s += sqr(0.5 * (g.a[3 * i + j] - g.a[3 * j + i]));
    return s;
}
static inline double magSqrSymm(const T9& g) {
    double s = 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) s += sqr(0.5 * (g.a[3 * i + j] + g.a[3 * j + i]));
    return s;
}
// SARC2::fr1 (SARC2.C:121-135)
static inline double saFr1(const rdst_turbulence& t, const T9& g) {
    if (!t.rcm_correction) return 1.0;
    double sqrS = 2.0 * magSqrSymm(g), sqrW = 2.0 * magSqrSkew(g);
    double rStar = std::sqrt(sqrS / std::max(sqrW, SMALL));
    double rrStar = std::sqrt(sqrW / std::max(sqrS, SMALL));
    double rTilda = rrStar * (rrStar - 1.0);
    return (1.0 + t.Cr1) * 2.0 * rStar / (1.0 + rStar) * (1.0 - t.Cr3 * std::atan(t.Cr2 * rTilda)) - t.Cr1;
}

// SARC1::fr1 (SARC1.C:125-176): Spalart-Shur rotation/curvature function.
// S = symm(grad U), W = skew(grad U) per cell; DSDt[9] = ddt(rho, S_ij) +
// div(phi, S_ij) per unit volume; sqrS = 2|S|^2, sqrW = 2|W|^2,
// sqrD = (sqrS + sqrW)/2, rStar = sqrt(sqrS / max(sqrW, SMALL)),
// rTilda = 2 / (rho max(sqrD, SMALL)^2) sum_ij (sum_k W_ik S_jk) DSDt_ij.
static inline double saFr1SS(const rdst_turbulence& t, const T9& g, double rho, const double* DSDt) {
    if (!t.rcm_correction) return 1.0;
    double S[9], W[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            S[3 * i + j] = 0.5 * (g.a[3 * i + j] + g.a[3 * j + i]);
            W[3 * i + j] = 0.5 * (g.a[3 * i + j] - g.a[3 * j + i]);
        }
    double sqrS = 0, sqrW = 0;
    for (int q = 0; q < 9; ++q) { sqrS += S[q] * S[q]; sqrW += W[q] * W[q]; }
    sqrS *= 2.0; sqrW *= 2.0;
    const double sqrD = 0.5 * (sqrS + sqrW);
    const double rStar = std::sqrt(sqrS / std::max(sqrW, SMALL));
    double sum = 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double ws = 0;
            for (int k = 0; k < 3; ++k) ws += W[3 * i + k] * S[3 * j + k];
            sum += ws * DSDt[3 * i + j];
        }
    const double rTilda = 2.0 / (rho * sqr(std::max(sqrD, SMALL))) * sum;
    return (1.0 + t.Cr1) * 2.0 * rStar / (1.0 + rStar) * (1.0 - t.Cr3 * std::atan(t.Cr2 * rTilda)) - t.Cr1;
}
static inline T9 symmT(const T9& g) {
    T9 s;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) s.a[3 * i + j] = 0.5 * (g.a[3 * i + j] + g.a[3 * j + i]);
    return s;
}

// nuTilda boundary values from 0/nuTilda (fixedValue, zeroGradient,
// symmetryPlane = zeroGradient for a scalar, inletOutlet on phi)
static void saNutBoundary(Case& c) {
    const Mesh& m = c.m;
    Turb& tb = c.tb;
    tb.bnut.resize(m.nB);
    for (int b = 0; b < m.nB; ++b) {
        const rdst_scalar_bc& bc = tb.bc[m.bPatch[b]];
        double vI = tb.nut[m.bCell[b]];
        switch (bc.type) {
            case RDST_BC_FIXED_VALUE: tb.bnut[b] = bc.value; break;
            case RDST_BC_INLET_OUTLET: {
                double vf = 1.0 - fe_pos(c.phi[m.nInternal + b]);
                tb.bnut[b] = vf * bc.value + (1.0 - vf) * vI;
                break;
            }
            default: tb.bnut[b] = vI;
        }
    }
}
// mut on the boundary: 0 on wall patches (the low-Re wall), elsewhere the
// calculated rho_b nuTilda_b |fv1(chi_b)|, bounded at 0
static void saMutBoundary(Case& c) {
    const Mesh& m = c.m;
    Turb& tb = c.tb;
    tb.bmut.resize(m.nB);
    tb.balphat.resize(m.nB);
    for (int b = 0; b < m.nB; ++b) {
        double v = 0.0;
        if (m.patches[m.bPatch[b]].type != RDST_PATCH_WALL) {
            double chi = c.brho[b] * tb.bnut[b] / c.bmu[b];
            v = std::max(std::fabs(saFv1(tb.t, chi)) * tb.bnut[b] * c.brho[b], 0.0);
        }
        tb.bmut[b] = v;
        tb.balphat[b] = v / tb.t.Prt;
    }
}

static void gradScalarB(const Mesh& m, const std::vector<double>& v, const std::vector<double>& bv,
                        std::vector<V3>& g) {
    int n = m.nCells;
    g.assign(n, v3(0, 0, 0));
    for (int f = 0; f < m.nInternal; ++f) {
        int o = m.owner[f], nb = m.neighbour[f];
        double vf = m.w[f] * (v[o] - v[nb]) + v[nb];
        V3 t = m.Sf[f] * vf;
        g[o] = g[o] + t;
        g[nb] = g[nb] - t;
    }
    for (int b = 0; b < m.nB; ++b) g[m.bCell[b]] = g[m.bCell[b]] + m.Sf[m.bFace[b]] * bv[b];
    for (int i = 0; i < n; ++i) g[i] = g[i] / m.V[i];
}

// Gauss linear gradient of a scalar with boundary values
static void gradScalar(const Mesh& m, const std::vector<double>& v, const std::vector<double>& bv,
                       std::vector<V3>& g) {
    int n = m.nCells;
    g.assign(n, v3(0, 0, 0));
    for (int f = 0; f < m.nInternal; ++f) {
        int o = m.owner[f], nb = m.neighbour[f];
        double vf = m.w[f] * (v[o] - v[nb]) + v[nb];
        V3 t = m.Sf[f] * vf;
        g[o] = g[o] + t;
        g[nb] = g[nb] - t;
    }
    for (int b = 0; b < m.nB; ++b) g[m.bCell[b]] = g[m.bCell[b]] + m.Sf[m.bFace[b]] * bv[b];
    for (int i = 0; i < n; ++i) g[i] = g[i] / m.V[i];
}

// foam bound(vsf, vsf0): where min(vsf) < vsf0, replace
// vsf = max(max(vsf, average(max(vsf, vsf0)) pos(vsf0 - vsf)), vsf0), with
// average() the |Sf|-weighted mean of the linear face values over each
// cell's faces; the caller re-evaluates the boundary values, then
// boundary = max(boundary, vsf0)
static bool saBound(const Mesh& m, std::vector<double>& v, const std::vector<double>& bv, double v0) {
    double mn = 1e300;
    for (double x : v) mn = std::min(mn, x);
    for (double x : bv) mn = std::min(mn, x);
    if (!(mn < v0)) return false;
    int n = m.nCells;
    std::vector<double> mc(n), num(n, 0.0), den(n, 0.0);
    for (int i = 0; i < n; ++i) mc[i] = std::max(v[i], v0);
    for (int f = 0; f < m.nInternal; ++f) {
        int o = m.owner[f], nb = m.neighbour[f];
        double vf = m.w[f] * (mc[o] - mc[nb]) + mc[nb];
        num[o] += m.magSf[f] * vf; num[nb] += m.magSf[f] * vf;
        den[o] += m.magSf[f]; den[nb] += m.magSf[f];
    }
    for (int b = 0; b < m.nB; ++b) {
        int ci = m.bCell[b];
        double ms = m.magSf[m.bFace[b]];
        num[ci] += ms * std::max(bv[b], v0);
        den[ci] += ms;
    }
    for (int i = 0; i < n; ++i) {
        double avg = num[i] / den[i];
        v[i] = std::max(std::max(v[i], avg * pos0(v0 - v[i])), v0);
    }
    return true;
}

// turbulence.correct(): SAneg.C:490-605 / SARC2.C:418-488
static void saCorrect(Case& c, Perf* perfOut) {
    const Mesh& m = c.m;
    Turb& tb = c.tb;
    const rdst_turbulence& t = tb.t;
    const int n = m.nCells, nI = m.nInternal, nB = m.nB;
    const bool isNeg = t.model == RDST_TURB_SANEG;
    const std::vector<double>& rhoOld = (tb.nCorrect > 0 && tb.rhoPrevValid) ? tb.rhoPrev : c.rho;
    const double Cw1 = saCw1(t);
    // fields of the current state
    std::vector<T9> g, gb;
    gradU(c, g, gb);
    std::vector<V3> gn;
    gradScalar(m, tb.nut, tb.bnut, gn);
    std::vector<double> fv1(n), diag(n), src(n);
    std::vector<double> G(n), bG(nB);
    for (int i = 0; i < n; ++i) G[i] = saGamma(t, c.rho[i], tb.nut[i], c.mu[i]);
    for (int b = 0; b < nB; ++b) bG[b] = saGamma(t, c.brho[b], tb.bnut[b], c.bmu[b]);
    // fvm::ddt(rho, nuTilda), steadyStateDST.C:230-255
    for (int i = 0; i < n; ++i) {
        diag[i] = c.rDeltaT[i] * c.rho[i] * m.V[i];
        src[i] = c.rDeltaT[i] * rhoOld[i] * tb.nut[i] * m.V[i];
    }
    std::vector<double> up(nI, 0.0), lo(nI, 0.0), ic(nB, 0.0), bcf(nB, 0.0);
    // fvm::div(phi, nuTilda) Gauss upwind: weights pos(phi) (>= 0)
    for (int f = 0; f < nI; ++f) {
        double phi = c.phi[f];
        double w = fe_pos(phi);
        double l = -w * phi, u = l + phi;
        lo[f] += l; up[f] += u;
        diag[m.owner[f]] -= l;
        diag[m.neighbour[f]] -= u;
    }
    for (int b = 0; b < nB; ++b) {
        const rdst_scalar_bc& bc = tb.bc[m.bPatch[b]];
        double phib = c.phi[nI + b];
        double vic, vbc;
        switch (bc.type) {
            case RDST_BC_FIXED_VALUE: vic = 0.0; vbc = bc.value; break;
            case RDST_BC_INLET_OUTLET: {
                double vf = 1.0 - fe_pos(phib);
                vic = 1.0 - vf; vbc = vf * bc.value;
                break;
            }
            default: vic = 1.0; vbc = 0.0;
        }
        ic[b] += phib * vic;
        bcf[b] += -phib * vbc;
    }
    // - fvm::laplacian(DnuTildaEff, nuTilda), Gauss linear with the
    // laplacianSchemes snGrad (corrected / uncorrected / limited psi)
    const double psiL = snGradPsi(c.st.laplacian_sngrad, c.st.laplacian_limit);
    for (int f = 0; f < nI; ++f) {
        int o = m.owner[f], nb = m.neighbour[f];
        double gf = m.w[f] * (G[o] - G[nb]) + G[nb];
        double gms = gf * m.magSf[f];
        double coef = gms * m.dc[f];
        up[f] -= coef; lo[f] -= coef;
        diag[o] += coef; diag[nb] += coef;
        V3 gfn = m.w[f] * (gn[o] - gn[nb]) + gn[nb];
        double q = dot(m.corr[f], gfn);
        double un = m.dc[f] * (tb.nut[nb] - tb.nut[o]);
        double corr = gms * (snGradLimiter(psiL, un * un, q * q) * q);
        src[o] += corr;
        src[nb] -= corr;
        if (t.convection == 1) {
            // div(phi,nuTilda) Gauss linearUpwind grad(nuTilda): the explicit
            // correction phi (Cf - C_u).grad(nuTilda)_u of the upwind cell u
            // (gaussConvectionScheme::fvmDiv adds fvc::surfaceIntegrate of it)
            const double phi = c.phi[f];
            const int u = phi > 0.0 ? o : nb;
            const V3 d = m.Cf[f] - m.C[u];
            const double lc = phi * (d.x * gn[u].x + d.y * gn[u].y + d.z * gn[u].z);
            src[o] -= lc;
            src[nb] += lc;
        }
    }
    for (int b = 0; b < nB; ++b) {
        const rdst_scalar_bc& bc = tb.bc[m.bPatch[b]];
        double gms = bG[b] * m.magSf[m.bFace[b]];
        double dcb = m.dc[nI + b];
        switch (bc.type) {
            case RDST_BC_FIXED_VALUE: ic[b] += gms * dcb; bcf[b] += gms * dcb * bc.value; break;
            case RDST_BC_INLET_OUTLET: {
                double vf = 1.0 - fe_pos(c.phi[nI + b]);
                ic[b] += gms * vf * dcb;
                bcf[b] += gms * vf * dcb * bc.value;
                break;
            }
            default: break;
        }
    }
    // SARC1: DS/Dt = ddt(rho, S) + div(phi, S) of the nine strain components
    // (steadyStateDST::fvcDdt: rDeltaT (rho S - rho.oldTime() S.oldTime()),
    // S.oldTime() the previous correct()'s S, the current S on the first;
    // div Gauss upwind with the face mass flux; boundary faces carry symm of
    // the corrected boundary gradient)
    std::vector<T9> Scell, DSDt;
    if (t.model == RDST_TURB_SARC1) {
        Scell.resize(n);
        DSDt.assign(n, T9{});
        for (int i = 0; i < n; ++i) Scell[i] = symmT(g[i]);
        const std::vector<T9>& So = (tb.nCorrect > 0 && (int)tb.Sold.size() == n) ? tb.Sold : Scell;
        for (int f = 0; f < nI; ++f) {
            const double phi = c.phi[f];
            const double w = fe_pos(phi);
            const int o = m.owner[f], nb = m.neighbour[f];
            for (int q = 0; q < 9; ++q) {
                const double sf = w * Scell[o].a[q] + (1.0 - w) * Scell[nb].a[q];
                DSDt[o].a[q] += phi * sf;
                DSDt[nb].a[q] -= phi * sf;
            }
        }
        for (int b = 0; b < nB; ++b) {
            const T9 Sb = symmT(gb[b]);
            const int ci = m.bCell[b];
            for (int q = 0; q < 9; ++q) DSDt[ci].a[q] += c.phi[nI + b] * Sb.a[q];
        }
        for (int i = 0; i < n; ++i)
            for (int q = 0; q < 9; ++q)
                DSDt[i].a[q] = c.rDeltaT[i] * (c.rho[i] * Scell[i].a[q] - rhoOld[i] * So[i].a[q]) + DSDt[i].a[q] / m.V[i];
    }
    // cell sources
    for (int i = 0; i < n; ++i) {
        const double V = m.V[i], rho = c.rho[i], nut = tb.nut[i], d = tb.d[i];
        const double chi = rho * nut / c.mu[i];
        fv1[i] = saFv1(t, chi);
        const double fv2 = saFv2(t, chi);
        const double Omega = std::sqrt(2.0) * std::sqrt(magSqrSkew(g[i]));
        // - Cb2/sigmaNut rho magSqr(grad nuTilda)
        src[i] += t.Cb2 / t.sigma_nut * rho * magSqr(gn[i]) * V;
        if (isNeg) {
            double Stilda;
            if (t.stilda_modification) {
                double Sbar = fv2 * nut / sqr(t.kappa * d);
                double a = t.Cn2 * Omega + Sbar;
                Stilda = Omega + pos0(a) * Sbar
                       + neg(a) * (Omega * (sqr(t.Cn2) * Omega + t.Cn3 * Sbar))
                             / std::max((t.Cn3 - 2.0 * t.Cn2) * Omega - Sbar, SMALL);
            } else {
                Stilda = saFv3(t, chi, fv1[i]) * std::sqrt(2.0) * std::sqrt(magSqrSkew(g[i])) + fv2 * nut / sqr(t.kappa * d);
            }
            const double ft2 = t.negative_nu_tilda ? t.Ct3 * std::exp(-t.Ct4 * sqr(chi)) : 0.0;
            const double fw = saFw(t, Stilda, nut, d);
            if (pos0(nut) > 0) {
                src[i] += t.Cb1 * (1.0 - ft2) * rho * Stilda * nut * V;
                diag[i] += (Cw1 * fw - t.Cb1 / sqr(t.kappa) * ft2) * nut * rho / sqr(d) * V;
            } else {
                src[i] += t.Cb1 * (1.0 - t.Ct3) * rho * Omega * nut * V;
                diag[i] -= Cw1 * nut * rho / sqr(d) * V;
            }
        } else {
            const double Stilda = saFv3(t, chi, fv1[i]) * std::sqrt(2.0) * std::sqrt(magSqrSkew(g[i]))
                                + fv2 * nut / sqr(t.kappa * d);
            const double fw = saFw(t, Stilda, nut, d);
            const double fr1 = (t.model == RDST_TURB_SARC1) ? saFr1SS(t, g[i], rho, DSDt[i].a) : saFr1(t, g[i]);
            src[i] += t.Cb1 * rho * Stilda * nut * fr1 * V;
            diag[i] += Cw1 * fw * nut * rho / sqr(d) * V;
        }
    }
    // nuTildaEqn.relax(): fvMatrix::relax(alpha)
    const double alpha = t.relax;
    if (alpha > 0) {
        std::vector<double> D0 = diag, sumOff(n, 0.0);
        for (int f = 0; f < nI; ++f) {
            sumOff[m.owner[f]] += std::fabs(up[f]);
            sumOff[m.neighbour[f]] += std::fabs(lo[f]);
        }
        for (int b = 0; b < nB; ++b) {
            int ci = m.bCell[b];
            double ic0 = ic[b];
            ic[b] = std::fabs(ic[b]);
            sumOff[ci] -= ic[b];
            ic[b] /= alpha;
            src[ci] += (ic[b] - ic0) * tb.nut[ci];
        }
        for (int i = 0; i < n; ++i) {
            diag[i] = std::max(diag[i], sumOff[i]);
            diag[i] /= alpha;
        }
        for (int i = 0; i < n; ++i) src[i] += (diag[i] - D0[i]) * tb.nut[i];
    }
    // solve form: addBoundaryDiag / addBoundarySource
    for (int b = 0; b < nB; ++b) {
        diag[m.bCell[b]] += ic[b];
        src[m.bCell[b]] += bcf[b];
    }
    tb.Ad = diag; tb.Au = up; tb.Al = lo; tb.As = src;
    // smoothSolver + symGaussSeidel; block-Jacobi over the cell partition
    // (one partition = the serial sweep): couplings across partitions use
    // the values at the start of each sweep
    std::vector<int> rs(n + 1, 0), cc;
    std::vector<double> cv;
    {
        std::vector<int> cnt(n, 0);
        for (int f = 0; f < nI; ++f) { cnt[m.owner[f]]++; cnt[m.neighbour[f]]++; }
        for (int i = 0; i < n; ++i) rs[i + 1] = rs[i] + cnt[i];
        cc.resize(rs[n]); cv.resize(rs[n]);
        std::vector<int> pos(rs.begin(), rs.end() - 1);
        for (int f = 0; f < nI; ++f) {
            int o = m.owner[f], nb = m.neighbour[f];
            cc[pos[o]] = nb; cv[pos[o]++] = up[f];
            cc[pos[nb]] = o; cv[pos[nb]++] = lo[f];
        }
    }
    std::vector<double>& x = tb.nut;
    auto residual = [&](std::vector<double>& Ax) {
        Ax.assign(n, 0.0);
        for (int i = 0; i < n; ++i) {
            double s = diag[i] * x[i];
            for (int k = rs[i]; k < rs[i + 1]; ++k) s += cv[k] * x[cc[k]];
            Ax[i] = s;
        }
    };
    std::vector<double> Ax;
    residual(Ax);
    double xRef = 0;
    for (int i = 0; i < n; ++i) xRef += x[i];
    xRef /= n;
    double nf = 0;
    for (int i = 0; i < n; ++i) {
        double sumA = diag[i];
        for (int k = rs[i]; k < rs[i + 1]; ++k) sumA += cv[k];
        double tmp = sumA * xRef;
        nf += std::fabs(Ax[i] - tmp) + std::fabs(src[i] - tmp);
    }
    nf += 1.0e-20;
    auto resNorm = [&]() {
        residual(Ax);
        double r = 0;
        for (int i = 0; i < n; ++i) r += std::fabs(src[i] - Ax[i]);
        return r / nf;
    };
    Perf pf{};
    pf.init[0] = resNorm();
    pf.fin[0] = pf.init[0];
    pf.nIter = 0;
    auto conv = [&]() { return pf.fin[0] < t.tolerance || (t.rel_tol > 0 && pf.fin[0] < t.rel_tol * pf.init[0]); };
    const std::vector<int>& part = c.saPart;
    auto samePart = [&](int a, int b) { return part.empty() || part[a] == part[b]; };
    if (!conv() || t.min_iter > 0) {
        std::vector<double> x0, bp(n);
        do {
            x0 = x;
            for (int i = 0; i < n; ++i) {
                double s = src[i];
                for (int k = rs[i]; k < rs[i + 1]; ++k)
                    if (!samePart(i, cc[k])) s -= cv[k] * x0[cc[k]];
                bp[i] = s;
            }
            for (int i = 0; i < n; ++i) {
                double s = bp[i];
                for (int k = rs[i]; k < rs[i + 1]; ++k)
                    if (samePart(i, cc[k])) s -= cv[k] * x[cc[k]];
                x[i] = s / diag[i];
            }
            for (int i = n - 1; i >= 0; --i) {
                double s = bp[i];
                for (int k = rs[i]; k < rs[i + 1]; ++k)
                    if (samePart(i, cc[k])) s -= cv[k] * x[cc[k]];
                x[i] = s / diag[i];
            }
            pf.fin[0] = resNorm();
            pf.nIter++;
        } while ((pf.nIter < t.max_iter && !conv()) || pf.nIter < t.min_iter);
    }
    pf.conv = conv();
    // bound, boundary values, mut, alphat
    if (!isNeg || !t.negative_nu_tilda) {
        saNutBoundary(c);
        if (saBound(m, x, tb.bnut, 0.0)) {
            saNutBoundary(c);
            for (double& v : tb.bnut) v = std::max(v, 0.0);
        }
    } else {
        saNutBoundary(c);
    }
    for (int i = 0; i < n; ++i) tb.mut[i] = (isNeg ? std::fabs(fv1[i]) : fv1[i]) * x[i] * c.rho[i];
    saMutBoundary(c);
    if (isNeg && t.negative_nu_tilda) {
        if (saBound(m, tb.mut, tb.bmut, 0.0)) saMutBoundary(c);
    }
    for (int i = 0; i < n; ++i) tb.alphat[i] = tb.mut[i] / t.Prt;
    if (t.model == RDST_TURB_SARC1) tb.Sold = Scell;
    tb.nCorrect++;
    if (perfOut) *perfOut = pf;
}

static void saSet(Case& c, const rdst_turbulence* t, const rdst_scalar_bc* bcs, const double* d, const double* nut) {
    Turb& tb = c.tb;
    const Mesh& m = c.m;
    int n = m.nCells;
    tb = Turb();
    tb.t = *t;
    // foam-extend's stock SpalartAllmaras::correct is SARC2::correct with
    // fr1 = 1 (SARC2.C:121-140: RCMCorrection off returns 1)
    if (tb.t.model == RDST_TURB_SA) {
        tb.t.model = RDST_TURB_SARC2;
        tb.t.rcm_correction = 0;
    }
    tb.on = t->model != RDST_TURB_LAMINAR;
    if (!tb.on) return;
    if (!c.hasState) throw std::runtime_error("set_turbulence: no state");
    tb.bc.assign(bcs, bcs + m.patches.size());
    tb.d.assign(d, d + n);
    tb.nut.assign(nut, nut + n);
    tb.mut.assign(n, 0.0);
    tb.alphat.assign(n, 0.0);
    saNutBoundary(c);
    const bool isNeg = t->model == RDST_TURB_SANEG;
    for (int i = 0; i < n; ++i) {
        double chi = c.rho[i] * tb.nut[i] / c.mu[i];
        double f = saFv1(*t, chi);
        tb.mut[i] = std::max((isNeg ? std::fabs(f) : f) * tb.nut[i] * c.rho[i], 0.0);
        tb.alphat[i] = tb.mut[i] / t->Prt;
    }
    saMutBoundary(c);
}

// forceCoeffs: pressure and viscous force on one patch; devRhoReff =
// -muEff dev(twoSymm(grad U)) with the corrected boundary gradient
static void forces(const Case& c, int patch, const double* dragDir, const double* liftDir, double pRef,
                   double rhoInf, double magUInf, double aRef, double* out) {
    const Mesh& m = c.m;
    std::vector<T9> g, gb;
    gradU(c, g, gb);
    std::vector<double> muE, bmuE;
    muAlphaEff(c, muE, bmuE, nullptr, nullptr);
    V3 Fp = v3(0, 0, 0), Fv = v3(0, 0, 0);
    for (int b = 0; b < m.nB; ++b) {
        if (m.bPatch[b] != patch) continue;
        V3 S = m.Sf[m.bFace[b]];
        Fp = Fp + S * (c.bp[b] - pRef);
        if (c.st.viscous) {
            const T9& G = gb[b];
            double tr = G.a[0] + G.a[4] + G.a[8];
            T9 D;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double ts = G.a[3 * i + j] + G.a[3 * j + i];
                    D.a[3 * i + j] = -bmuE[b] * (ts - (i == j ? (2.0 / 3.0) * tr : 0.0));
                }
            Fv = Fv + dotVT(S, D);
        }
    }
    V3 F = Fp + Fv;
    double q = 0.5 * rhoInf * magUInf * magUInf * aRef;
    out[0] = (F.x * dragDir[0] + F.y * dragDir[1] + F.z * dragDir[2]) / q;
    out[1] = (F.x * liftDir[0] + F.y * liftDir[1] + F.z * liftDir[2]) / q;
    out[2] = F.x; out[3] = F.y; out[4] = F.z;
    out[5] = Fv.x; out[6] = Fv.y; out[7] = Fv.z;
}

}  // namespace orc

// ======================================================================= C API
using namespace orc;

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

#define ORC_TRY(...) try { __VA_ARGS__; return 0; } catch (const std::exception& e) { g_err = e.what(); return 1; }

// geometry of a polyMesh, the fvMesh quantities the operators read
int orc_geometry(const rdst_mesh_desc* d, double* Sf, double* magSf, double* Cf, double* C, double* V,
                 double* w, double* dc, int32_t* nB_out) {
    ORC_TRY({
        Mesh m;
        buildMesh(m, d);
        int nf = m.nInternal + m.nB;
        for (int k = 0; k < nf; ++k) {
            int f = k < m.nInternal ? k : m.bFace[k - m.nInternal];
            if (Sf) { Sf[3 * k] = m.Sf[f].x; Sf[3 * k + 1] = m.Sf[f].y; Sf[3 * k + 2] = m.Sf[f].z; }
            if (magSf) magSf[k] = m.magSf[f];
            if (Cf) { Cf[3 * k] = m.Cf[f].x; Cf[3 * k + 1] = m.Cf[f].y; Cf[3 * k + 2] = m.Cf[f].z; }
            if (w) w[k] = m.w[k];
            if (dc) dc[k] = m.dc[k];
        }
        for (int i = 0; i < m.nCells; ++i) {
            if (C) { C[3 * i] = m.C[i].x; C[3 * i + 1] = m.C[i].y; C[3 * i + 2] = m.C[i].z; }
            if (V) V[i] = m.V[i];
        }
        if (nB_out) *nB_out = m.nB;
    })
}

// single-face flux of a scheme (golden vectors and invariant tests)
// state arrays: [rho, rhoUx, rhoUy, rhoUz, e, rPsi, aStar, c] for each side
int orc_face_flux(int scheme, const double* L, const double* R, const double* Sf, double gamma,
                  double fHLLCP, int lowMach, double Mainf, double* out /* phi, phiU[3], phiE, Up[3] */) {
    ORC_TRY({
        FaceIn in;
        in.rho_o = L[0]; in.rhoU_o = v3(L[1], L[2], L[3]); in.e_o = L[4]; in.rPsi_o = L[5]; in.a_o = L[6]; in.c_o = L[7];
        in.rho_n = R[0]; in.rhoU_n = v3(R[1], R[2], R[3]); in.e_n = R[4]; in.rPsi_n = R[5]; in.a_n = R[6]; in.c_n = R[7];
        in.Sf = v3(Sf[0], Sf[1], Sf[2]); in.magSf = mag(in.Sf); in.gam_f = gamma; in.fHLLCP = fHLLCP;
        rdst_settings st = {};
        st.low_mach = lowMach; st.Mainf = Mainf;
        FaceOut o = faceFlux(scheme, in, st);
        out[0] = o.phi; out[1] = o.phiU.x; out[2] = o.phiU.y; out[3] = o.phiU.z; out[4] = o.phiE;
        out[5] = o.Up.x; out[6] = o.Up.y; out[7] = o.Up.z;
    })
}

int orc_euler_jacobian(const double* U, double gamma, double rhoE, double rho, const double* n, double* J) {
    ORC_TRY({ eulerJacobian(v3(U[0], U[1], U[2]), gamma, rhoE, rho, v3(n[0], n[1], n[2]), J); })
}

int orc_case_create(const rdst_mesh_desc* d, const rdst_thermo* th, const rdst_bc* bcs,
                    const rdst_settings* st, void** out) {
    ORC_TRY({
        if (st->jacobian != RDST_JACOBIAN_ROE) throw std::runtime_error("the oracle restates the Roe Jacobian only");
        Case* c = new Case();
        buildMesh(c->m, d);
        c->th = *th;
        c->bc.assign(bcs, bcs + d->n_patches);
        c->st = *st;
        c->Cv = th->Cp - th->R;
        c->gamma = th->Cp / c->Cv;
        *out = c;
    })
}
int orc_case_destroy(void* h) { delete (Case*)h; return 0; }
int orc_case_set_state(void* h, const double* p, const double* T, const double* U) {
    ORC_TRY({ setState(*(Case*)h, p, T, U); })
}
int orc_flux_residual(void* h) { ORC_TRY({ computeFlux(*(Case*)h); }) }
int orc_jacobian_assemble(void* h) {
    ORC_TRY({ Case& c = *(Case*)h; computeJacobian(c); assemble(c); })
}
int orc_block_solve(void* h, rdst_solver_perf* perf) {
    ORC_TRY({
        Perf p = solve(*(Case*)h);
        if (perf) {
            for (int k = 0; k < 5; ++k) { perf->initial_residual[k] = p.init[k]; perf->final_residual[k] = p.fin[k]; }
            perf->n_iterations = p.nIter; perf->converged = p.conv;
        }
    })
}
int orc_update_fields(void* h) { ORC_TRY({ updateFields(*(Case*)h); }) }
int orc_set_courant(void* h, double* res) { ORC_TRY({ double r = setCourant(*(Case*)h); if (res) *res = r; }) }
int orc_dual_step(void* h, int32_t* nInner, double* innerRes, double* res) {
    ORC_TRY({
        Case& c = *(Case*)h;
        if (c.st.time_integration != RDST_TIME_DUAL) throw std::runtime_error("orc_dual_step: not a Dual case");
        int k = 0;
        double r = dualStep(c, &k, innerRes);
        if (nInner) *nInner = k;
        if (res) *res = r;
    })
}
int orc_time_info(void* h, double* out) {
    ORC_TRY({
        const Case& c = *(Case*)h;
        out[0] = c.time; out[1] = c.deltaT; out[2] = c.deltaT0; out[3] = c.courantScalar;
    })
}
// steadyCheck on a given history: residuals are added one at a time and
// converged() is asked after each, as Dual::iterate does; returns the
// 1-based iteration at which it first converged (0: never) and the scaled
// slope at that point
int orc_steady_check(const double* residuals, int n, const rdst_settings* st, int32_t* at, double* slope) {
    ORC_TRY({
        SteadyCheck sc = makeSteadyCheck(*st);
        *at = 0;
        for (int i = 0; i < n; ++i) {
            double r5[5] = {residuals[i], residuals[i], residuals[i], residuals[i], residuals[i]};
            sc.addResidual(r5);
            if (sc.converged()) { *at = i + 1; break; }
        }
        *slope = sc.magScaledSlope;
    })
}
int orc_iterate(void* h, int nSteps, double* res, rdst_solver_perf* perfs) {
    ORC_TRY({
        Case& c = *(Case*)h;
        for (int s = 0; s < nSteps; ++s) {
            Perf p;
            double r = iterate(c, &p);
            if (res) res[s] = r;
            if (perfs) {
                for (int k = 0; k < 5; ++k) { perfs[s].initial_residual[k] = p.init[k]; perfs[s].final_residual[k] = p.fin[k]; }
                perfs[s].n_iterations = p.nIter; perfs[s].converged = p.conv;
            }
        }
    })
}

// LDU access: which = 0 diag [n][25], 1 upper [F][25], 2 lower [F][25], 3 source [n][5],
// 4 lambda [F], 5 ownerCoeffs [F][25], 6 neighbourCoeffs [F][25], 7 patchInternalCoeffs [nB][25]
int orc_get_matrix(void* h, int which, double* out) {
    ORC_TRY({
        Case& c = *(Case*)h;
        const std::vector<double>* v = nullptr;
        switch (which) {
            case 0: v = &c.diag; break; case 1: v = &c.upper; break; case 2: v = &c.lower; break;
            case 3: v = &c.source; break; case 4: v = &c.lambda; break; case 5: v = &c.ownC; break;
            case 6: v = &c.neiC; break; case 7: v = &c.patchIC; break;
            default: throw std::runtime_error("bad matrix id");
        }
        std::memcpy(out, v->data(), v->size() * sizeof(double));
    })
}

// apply the assembled matrix / the preconditioner (for solver tests)
int orc_amul(void* h, const double* x, double* y) {
    ORC_TRY({
        Case& c = *(Case*)h;
        std::vector<double> xv(x, x + 5 * (size_t)c.m.nCells), yv;
        Amul(c, xv, yv);
        std::memcpy(y, yv.data(), yv.size() * sizeof(double));
    })
}
int orc_ilu_apply(void* h, const double* r, double* x) {
    ORC_TRY({
        Case& c = *(Case*)h;
        ILU il;
        buildILU(c, il);
        std::vector<double> rv(r, r + 5 * (size_t)c.m.nCells), xv;
        iluApply(il, rv, xv);
        std::memcpy(x, xv.data(), xv.size() * sizeof(double));
    })
}
// solve A x = b with the configured solver from x0 (for solver tests)
int orc_linear_solve(void* h, const double* b, double* x, rdst_solver_perf* perf) {
    ORC_TRY({
        Case& c = *(Case*)h;
        Precond M;
        M.init(c, c.st.precond);
        std::vector<double> bv(b, b + 5 * (size_t)c.m.nCells), xv(x, x + 5 * (size_t)c.m.nCells);
        Perf p = (c.st.solver == RDST_SOLVER_GMRES) ? gmres(c, M, xv, bv) : bicgstab(c, M, xv, bv);
        std::memcpy(x, xv.data(), xv.size() * sizeof(double));
        if (perf) {
            for (int k = 0; k < 5; ++k) { perf->initial_residual[k] = p.init[k]; perf->final_residual[k] = p.fin[k]; }
            perf->n_iterations = p.nIter; perf->converged = p.conv;
        }
    })
}
// overwrite the assembled matrix (diag/upper/lower) for solver tests
int orc_set_matrix(void* h, const double* diag, const double* upper, const double* lower) {
    ORC_TRY({
        Case& c = *(Case*)h;
        int n = c.m.nCells, F = c.m.nInternal;
        c.diag.assign(diag, diag + 25 * (size_t)n);
        c.upper.assign(upper, upper + 25 * (size_t)F);
        c.lower.assign(lower, lower + 25 * (size_t)F);
        c.source.assign(5 * (size_t)n, 0.0);
        c.hasMatrix = true;
    })
}

// cell partition of the block-Jacobi ILU(0) (settings.ilu_order = partitioned)
int orc_set_partition(void* h, const int32_t* part) {
    ORC_TRY({
        Case& c = *(Case*)h;
        c.part.assign(part, part + c.m.nCells);
    })
}

// cell partition of the block-Jacobi symmetric Gauss-Seidel sweep of nuTilda
int orc_set_sa_partition(void* h, const int32_t* part) {
    ORC_TRY({
        Case& c = *(Case*)h;
        c.saPart.assign(part, part + c.m.nCells);
    })
}

// the limited reconstruction alone: cell [n] (or [n][3]) and boundary values
// [nB] ([nB][3]) -> owner-side and neighbour-side face values [nI + nB] (x3)
int orc_reconstruct(void* h, int vector, const double* cell, const double* bnd, double* own, double* nei) {
    ORC_TRY({
        const Mesh& m = ((Case*)h)->m;
        int n = m.nCells, nB = m.nB, nf = m.nInternal + m.nB;
        if (!vector) {
            std::vector<double> c(cell, cell + n), b(bnd, bnd + nB), o, e;
            reconstructOSPRE(m, c, b, o, e);
            std::memcpy(own, o.data(), nf * sizeof(double));
            std::memcpy(nei, e.data(), nf * sizeof(double));
        } else {
            std::vector<V3> c(n), b(nB), o, e;
            for (int i = 0; i < n; ++i) c[i] = v3(cell[3 * i], cell[3 * i + 1], cell[3 * i + 2]);
            for (int i = 0; i < nB; ++i) b[i] = v3(bnd[3 * i], bnd[3 * i + 1], bnd[3 * i + 2]);
            reconstructOSPREV(m, c, b, o, e);
            for (int k = 0; k < nf; ++k) {
                own[3 * k] = o[k].x; own[3 * k + 1] = o[k].y; own[3 * k + 2] = o[k].z;
                nei[3 * k] = e[k].x; nei[3 * k + 1] = e[k].y; nei[3 * k + 2] = e[k].z;
            }
        }
    })
}
int orc_set_turbulence(void* h, const rdst_turbulence* t, const rdst_scalar_bc* bcs, const double* d,
                       const double* nut) {
    ORC_TRY({ saSet(*(Case*)h, t, bcs, d, nut); })
}
int orc_turbulence_correct(void* h, rdst_solver_perf* perf) {
    ORC_TRY({
        Case& c = *(Case*)h;
        if (!c.tb.on) throw std::runtime_error("turbulence_correct: no model");
        Perf p{};
        saCorrect(c, &p);
        if (perf) {
            *perf = rdst_solver_perf{};
            perf->initial_residual[0] = p.init[0];
            perf->final_residual[0] = p.fin[0];
            perf->n_iterations = p.nIter;
            perf->converged = p.conv ? 1 : 0;
        }
    })
}
int orc_forces(void* h, int patch, const double* drag, const double* lift, double pRef, double rhoInf,
               double magUInf, double aRef, double* out) {
    ORC_TRY({ forces(*(Case*)h, patch, drag, lift, pRef, rhoInf, magUInf, aRef, out); })
}
int orc_case_get(void* h, int field, double* out) {
    ORC_TRY({
        Case& c = *(Case*)h;
        const Mesh& m = c.m;
        auto putS = [&](const std::vector<double>& v) { std::memcpy(out, v.data(), v.size() * sizeof(double)); };
        auto putV = [&](const std::vector<V3>& v) {
            for (size_t i = 0; i < v.size(); ++i) { out[3 * i] = v[i].x; out[3 * i + 1] = v[i].y; out[3 * i + 2] = v[i].z; }
        };
        switch (field) {
            case RDST_F_RHO: putS(c.rho); break; case RDST_F_RHOU: putV(c.rhoU); break;
            case RDST_F_RHOE: putS(c.rhoE); break; case RDST_F_P: putS(c.p); break;
            case RDST_F_T: putS(c.T); break; case RDST_F_U: putV(c.U); break;
            case RDST_F_H: putS(c.h); break; case RDST_F_PSI: putS(c.psi); break;
            case RDST_F_R: putS(c.R); break; case RDST_F_W: putS(c.W); break;
            case RDST_F_RDELTAT: putS(c.rDeltaT); break; case RDST_F_COURANT: putS(c.courant); break;
            case RDST_F_PHI: putS(c.phi); break; case RDST_F_PHIUP: putV(c.phiUp); break;
            case RDST_F_PHIEP: putS(c.phiEp); break; case RDST_F_UP: putV(c.Up); break;
            case RDST_F_BRHO: putS(c.brho); break; case RDST_F_BU: putV(c.bU); break;
            case RDST_F_BP: putS(c.bp); break; case RDST_F_BT: putS(c.bT); break;
            case RDST_F_BRHOU: putV(c.brhoU); break; case RDST_F_BRHOE: putS(c.brhoE); break;
            case RDST_F_BH: putS(c.bh); break; case RDST_F_BPSI: putS(c.bpsi); break;
            case RDST_F_SOURCE: putS(c.source); break; case RDST_F_MU: putS(c.mu); break;
            case RDST_F_RHO_OLD: putS(c.rhoOld); break; case RDST_F_E_OLD: putS(c.eOld); break;
            case RDST_F_NU_TILDA: putS(c.tb.nut); break; case RDST_F_MUT: putS(c.tb.mut); break;
            case RDST_F_ALPHAT: putS(c.tb.alphat); break; case RDST_F_WALL_DIST: putS(c.tb.d); break;
            case RDST_F_BNU_TILDA: putS(c.tb.bnut); break; case RDST_F_BMUT: putS(c.tb.bmut); break;
            case RDST_F_SA_DIAG: putS(c.tb.Ad); break; case RDST_F_SA_UPPER: putS(c.tb.Au); break;
            case RDST_F_SA_LOWER: putS(c.tb.Al); break; case RDST_F_SA_SOURCE: putS(c.tb.As); break;
            case RDST_F_GRAD_U: {  // grad(U) of the current U and boundary values
                std::vector<T9> g, gb;
                gradU(c, g, gb);
                for (size_t i = 0; i < g.size(); ++i)
                    for (int q = 0; q < 9; ++q) out[9 * i + q] = g[i].a[q];
                break;
            }
            default: throw std::runtime_error("bad field id");
        }
        (void)m;
    })
}

int orc_case_put(void* h, int field, const double* in) {
    ORC_TRY({
        Case& c = *(Case*)h;
        int n = c.m.nCells;
        switch (field) {
            case RDST_F_RHO: c.rho.assign(in, in + n); break;
            case RDST_F_RHOE: c.rhoE.assign(in, in + n); break;
            case RDST_F_RHOU: for (int i = 0; i < n; ++i) c.rhoU[i] = v3(in[3 * i], in[3 * i + 1], in[3 * i + 2]); break;
            case RDST_F_W: c.W.assign(in, in + 5 * (size_t)n); break;
            case RDST_F_RDELTAT: c.rDeltaT.assign(in, in + n); break;
            // the stored mass flux, which the boundary conditions' masks read
            // (a state restored before a differenced residual)
            case RDST_F_PHI: c.phi.assign(in, in + c.phi.size()); break;
            // scripted inputs of setCourantDST (timeDST.C:316-414): the
            // residual R_, the time-control state and the previous rho, e
            case RDST_F_R: c.R.assign(in, in + 5 * (size_t)n); break;
            case RDST_F_COURANT: c.courant.assign(in, in + n); break;
            case RDST_F_RHO_OLD: c.rhoOld.assign(in, in + n); break;
            case RDST_F_E_OLD: c.eOld.assign(in, in + n); break;
            case RDST_F_NU_TILDA:
                if (!c.tb.on) throw std::runtime_error("no turbulence model");
                c.tb.nut.assign(in, in + n);
                saNutBoundary(c);
                break;
            // an eddy viscosity held fixed (a differenced residual with the
            // turbulence frozen); alphat_b = mut_b / Prt as saMutBoundary sets it
            case RDST_F_MUT:
                if (!c.tb.on) throw std::runtime_error("no turbulence model");
                c.tb.mut.assign(in, in + n);
                break;
            case RDST_F_ALPHAT:
                if (!c.tb.on) throw std::runtime_error("no turbulence model");
                c.tb.alphat.assign(in, in + n);
                break;
            case RDST_F_BMUT:
                if (!c.tb.on) throw std::runtime_error("no turbulence model");
                c.tb.bmut.assign(in, in + c.m.nB);
                c.tb.balphat.resize(c.m.nB);
                for (int b = 0; b < c.m.nB; ++b) c.tb.balphat[b] = c.tb.bmut[b] / c.tb.t.Prt;
                break;
            default: throw std::runtime_error("field not writable");
        }
    })
}

}  // extern "C"
