// rdst_internal.hpp -- host-side structures of the MI355X block-coupled
// implicit step: the uploaded mesh, the device-resident case state and the
// kernel launch helpers shared by the .hip translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/rhodst.h"

namespace rdst {

void set_error(const std::string& msg);

#define RDST_HIP(call)                                                                  \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            throw ::rdst::HipError(std::string(#call) + ": " + hipGetErrorString(e_)); \
        }                                                                               \
    } while (0)

struct HipError : std::exception {
    std::string msg;
    explicit HipError(std::string m) : msg(std::move(m)) {}
    const char* what() const noexcept override { return msg.c_str(); }
};
struct ArgError : std::exception {
    std::string msg;
    int code;
    ArgError(std::string m, int c = RDST_ERR_ARG) : msg(std::move(m)), code(c) {}
    const char* what() const noexcept override { return msg.c_str(); }
};

// A device buffer with RAII.
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) RDST_HIP(hipMalloc((void**)&p, count * sizeof(T)));
    }
    void upload(const T* h, size_t count) {
        alloc(count);
        if (count) RDST_HIP(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice));
    }
    void upload(const std::vector<T>& h) { upload(h.data(), h.size()); }
    void zero(hipStream_t s = 0) {
        if (n) RDST_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s));
    }
    void download(T* h, size_t count) const {
        if (count) RDST_HIP(hipMemcpy(h, p, count * sizeof(T), hipMemcpyDeviceToHost));
    }
};

// ELL cell->face entry kinds (bits 30-31 of the code word)
enum : uint32_t { ELL_OWNER = 0u, ELL_NEIGH = 1u, ELL_BOUNDARY = 2u, ELL_NONE = 3u };
static constexpr uint32_t ELL_IDX_MASK = 0x3fffffffu;

// ---------------------------------------------------------------------------
// Mesh: polyMesh primitives, fvMesh geometry (host), device copies.
// One sweep of the partitioned ILU(0) apply laid out for coalesced loads.
// The rows of a partition-local level are independent; a level is cut into
// sub-levels of at most kPiluSubRows rows, and inside a sub-level the rows are
// ordered by their block count, largest first, so that block k of every row
// that has one forms a contiguous "jagged diagonal": diagonal k of sub-level
// s holds n_k blocks at stream positions blk_k .. blk_k + n_k - 1, row j's at
// blk_k + j.  Block values are stored per diagonal as five planes (column c of
// the block) of 5 n_k values, entry (j, r, c) at
// 25 blk_k + 5 n_k c + 5 j + r, so the five lanes of a row and the rows of a
// wavefront read contiguous memory.  Sub-level s is described by
// sub[kPiluSubW s ..]: rowOff (into ord), nRows, kmax, 0, blk_0..blk_7,
// n_0..n_7.  The forward stream holds each row's in-partition L blocks (rows
// without one are left out); the backward stream holds, per row, Dinv first
// (its column is the row itself) and then Dinv U for each in-partition U
// block, so x = Dinv y - sum (Dinv U) x needs no separate diagonal solve.
constexpr int kPiluKD = 8;                  // most blocks of a row in one sweep
constexpr int kPiluSubW = 4 + 2 * kPiluKD;  // ints per sub-level descriptor
constexpr int kPiluSubRows = 48;            // 4 wavefronts x 12 rows of 5 lanes
struct PiluStream {
    int kmax = 0;                  // largest row block count (0: not built)
    long nBlk = 0;
    std::vector<int32_t> subStart;  // [nParts + 1]
    std::vector<int32_t> sub;       // [nSub + 1][kPiluSubW] (one zero descriptor of padding)
    std::vector<int32_t> ord;       // partition-local row of each sub-level position
    std::vector<int32_t> lc;        // [nBlk] partition-local column
    std::vector<int32_t> src;       // [nBlk] block-CSR slot, or -(cell + 1) for the row's Dinv
    std::vector<int32_t> row;       // [nBlk] cell of the block's row
    std::vector<int32_t> jn;        // [nBlk] j | n_k << 8
    DBuf<int32_t> d_subStart, d_sub, d_ord, d_lc, d_src, d_row, d_jn;
};

// A set of compact cell partitions in sweep order: rows of partition p at
// rows[rowStart[p] .. rowStart[p+1]) in (partition-local level, cell) order,
// level L's rows at rows[levPtr[L] .. levPtr[L+1]) for L in
// [levStart[p], levStart[p+1] - 1), localCol = partition-local column of
// each block-CSR slot (-1 outside the partition).
struct LevelRun {
    int L0, L1;   // levels [L0, L1)
    bool fused;   // all of them at most kIluSegRows rows: one single-workgroup launch
};
constexpr int kIluSegRows = 64;

struct PartSet {
    int nParts = 0, maxRows = 0, maxLevels = 0;
    std::vector<int32_t> rowStart, rows, levStart, levPtr, localCol;
    DBuf<int32_t> d_rowStart, d_rows, d_levStart, d_levPtr, d_localCol;
};

struct Mesh {
    int nC = 0, nI = 0, nB = 0, nF = 0;  // nF = nI + nB surface-field length
    int nPatches = 0;
    std::vector<rdst_patch> patches;
    // host geometry (surface fields: internal faces then compacted boundary faces)
    std::vector<double> Sf, Cf, magSf, w, dc;  // Sf/Cf [nF][3]
    std::vector<double> corr;                  // [nI][3] non-orthogonal correction vectors n - d deltaCoeffs
    std::vector<double> C, V;                  // C [nC][3]
    std::vector<int32_t> owner, neigh;         // owner [nF] (boundary: face cell), neigh [nI]
    std::vector<int32_t> bPatch, bPolyFace;    // [nB]
    // ELL cell->face map (faces of each cell in face order: internal ascending, then boundary)
    int ellW = 0;
    std::vector<uint32_t> ellCode;   // [ellW][nC]
    std::vector<int32_t> ellFace;    // [ellW][nC]  surface-field index or -1
    std::vector<int32_t> ellSlot;    // [ellW][nC]  BCSR slot of the off-diagonal block this cell writes, or -1
    // block-CSR, rows = cells, columns sorted
    int nnzb = 0;
    std::vector<int32_t> rowPtr, cols, slotUpper, slotLower, slotDiag;
    // face colouring (no two faces of one colour share a cell)
    int nColours = 0;
    std::vector<int32_t> faceColour;
    std::vector<int32_t> cellColour;
    int nCellColours = 0;
    // level sets of the natural-order lower/upper triangular sweeps
    std::vector<int32_t> levL, levU;           // level of each row
    std::vector<int32_t> levLPtr, levLRows, levUPtr, levURows;
    int nLevL = 0, nLevU = 0;
    // the levels cut into runs: consecutive levels of at most kIluSegRows rows
    // form one fused run (one single-workgroup launch), larger levels one
    // launch each
    std::vector<LevelRun> runL, runU;
    // compact partitions (recursive coordinate bisection of the cell centres)
    // for the block-Jacobi ILU(0): rows of each partition in (local level,
    // cell) order, level offsets, and per block the partition-local column
    int nParts = 0, maxPartRows = 0, maxPartLevels = 0;
    std::vector<int32_t> part, partRowStart, partRows, partLevStart, partLevPtr, localCol;
    // the RCB boxes before the level cut: the partitions of the
    // Spalart-Allmaras Gauss-Seidel sweep
    std::vector<int32_t> saPart;
    PartSet sa;
    // coalesced streams of the partitioned ILU(0) apply (forward: the
    // in-partition L blocks; backward: Dinv and Dinv U): see PiluStream
    PiluStream pf, pb;
    // sliced-ELL copy of the block-CSR pattern for the SpMV: slices of 64
    // rows, slice s holds sellWidth[s] block columns; block p of row r sits at
    // block position sellPtr[s] + 64 p + (r - 64 s), its entry e at value
    // index 25 * (sellPtr[s] + 64 p) + 64 e + (r - 64 s)
    int nSlices = 0;
    long sellBlocks = 0;
    std::vector<int64_t> sellPtr;
    std::vector<int32_t> sellWidth, sellCol;
    std::vector<int64_t> sellOfSlot;  // value index of entry 0 of each block-CSR slot
    // wall-patch faces fanned from the mean of their points: [nT][3 corners][3]
    std::vector<double> wallTri;
    bool uploaded = false;  // device copies are made by the first rdst_case_create
    // second-neighbour block pattern of the exact viscous linearisation (built
    // on first use): row r holds r, its face neighbours and theirs, columns
    // ascending; wideT[e] is the position of the transposed block of entry e,
    // wideOfSlot[s] the position of block-CSR slot s; wideColor colours the
    // cells so that two cells of one colour are more than four faces apart,
    // hence no row holds two columns of one colour
    bool wideBuilt = false;
    int wideColors = 0;
    std::vector<int32_t> wideRowPtr, wideCols, wideT, wideOfSlot, wideColor;

    // ---- device copies
    DBuf<int32_t> d_owner, d_neigh, d_bcell, d_bpatch;
    DBuf<double> d_Sf, d_magSf, d_w, d_dc, d_V;  // Sf SoA [3][nF]
    DBuf<uint32_t> d_ellCode;
    DBuf<int32_t> d_ellFace, d_ellSlot;
    DBuf<double> d_ellSf, d_ellMagSf, d_ellWt, d_ellDc;  // [3][ellW][nC], [ellW][nC]
    DBuf<double> d_ellCorr;                               // [3][ellW][nC]
    DBuf<int32_t> d_rowPtr, d_cols, d_slotUpper, d_slotLower, d_slotDiag;
    DBuf<int32_t> d_levLRows, d_levURows, d_levLPtr, d_levUPtr;
    DBuf<int32_t> d_partRowStart, d_partRows, d_partLevStart, d_partLevPtr, d_localCol;
    DBuf<int64_t> d_sellPtr, d_sellOfSlot;
    DBuf<int32_t> d_sellWidth, d_sellCol;
    DBuf<int32_t> d_bWall;  // [nB] 1 where the patch is a wall
    DBuf<double> d_C;       // cell centres SoA [3][nC]
    DBuf<double> d_Cf;      // face centres SoA [3][nF]
    DBuf<int32_t> d_wideRowPtr, d_wideCols, d_wideT, d_wideOfSlot, d_wideColor;
};

// largest partition of the block-Jacobi ILU(0); its FP64 vector (40 B/cell)
// is held in LDS by one workgroup
constexpr int kPartCells = 1024;
// deepest partition-local level chain kept; deeper partitions are cut in two
constexpr int kMaxPartLevels = 32;
void mesh_build(Mesh& m, const rdst_mesh_desc* d, const rdst_mesh_geometry* g = nullptr);
void build_partitions(Mesh& m, int maxCells);
void mesh_upload(Mesh& m);

// ---------------------------------------------------------------------------
// Device-side views passed to kernels by value.
struct BcPatchDev {
    int p_type, T_type, U_type, is_wall;
    double p_value, T_value, U_value[3];
    double p_inf, T_inf, U_inf[3];
};

struct Thermo {
    double Cp, R, Cv, gamma;
    int transport;
    double mu, Pr, As, Ts;
};

struct Fields {
    // cells (SoA)
    double *rho, *rhoUx, *rhoUy, *rhoUz, *rhoE, *p, *T, *h, *psi, *mu, *alpha;
    double *Ux, *Uy, *Uz;
    // boundary faces
    double *brho, *bUx, *bUy, *bUz, *bp, *bT, *bh, *bpsi, *bmu, *balpha, *brhoUx, *brhoUy, *brhoUz, *brhoE;
    double *vf_p, *vf_T, *vf_U, *TrefH;
    double *refUx, *refUy, *refUz;  // mixed U BCs: refValue
    // viscous terms: grad(U) [9][nC] (T[3i+j] = dU_j/dx_i), corrected boundary gradient [9][nB]
    double *gradU, *bgradU;
    // the unlimited Gauss linear grad(U) [9][nC] that the non-orthogonal
    // correction of snGrad(U) interpolates: gradU itself unless grad(U) is
    // cellLimited
    double* gradUc;
    // surface fields (internal then boundary)
    double *phi, *phiUx, *phiUy, *phiUz, *phiE, *Upx, *Upy, *Upz;
    // residual R_ [5][nC]
    double* R;
    // W [nC][5] (AoS, the linear-solve vector)
    double* W;
    // time control
    double *rDeltaT, *courant, *rhoOld, *eOld, *dS;  // dS [3][nC]
    double* fH;      // HLLCP pressure-sensor cell values pow3(f) [nC]
    // turbulence viscosity and diffusivity (zero when laminar): muEff = mut + mu,
    // alphaEff = alphat + alpha
    double *mut, *alphat, *bmut, *balphat;
    // limited reconstruction: derived cell fields rPsi, e, a*, c [4][nC] and
    // boundary values [4][nB]; Gauss gradients of rho, rhoU, rPsi, e, a*, c
    // [24][nC] (component 3q + i = d(field q)/dx_i)
    double *rec, *brec, *rgrad;
    // the face loop's copy of the cell state, AoS [nC][8], 64 B per cell, so
    // a neighbour gather touches one cache line instead of eight SoA lines;
    // written by the kernels that produce the state (k_init_cells,
    // k_update_cells).  CELL8_STATE: rho, rhoUx, rhoUy, rhoUz, rhoE, p, h,
    // psi.  CELL8_SIDE (the schemes whose flux reads only rho, U, 1/rho,
    // 1/psi, e and a* of a side): those eight quantities themselves, formed
    // once per cell instead of once per face side.
    double* cell8;
    int cell8Kind;
};
enum { CELL8_STATE = 0, CELL8_SIDE = 1 };
inline int cell8_kind(int fluxScheme) {
    return (fluxScheme == RDST_FLUX_AUSMPLUS || fluxScheme == RDST_FLUX_AUSMPLUSUP || fluxScheme == RDST_FLUX_HLL)
               ? CELL8_SIDE
               : CELL8_STATE;
}

// Spalart-Allmaras state of a case (SAneg/SARC2 nuTilda_, mut_, alphat_, d_)
struct SaState {
    bool on = false;
    rdst_turbulence t{};
    std::vector<rdst_scalar_bc> bc;
    DBuf<rdst_scalar_bc> d_bc;
    DBuf<double> buf;      // nut, d, rhoPrev, fv1o, B, x2 [n each], gn [3][n], bnut [nB]
    DBuf<double> A;        // nuTilda matrix values per block-CSR slot [nnzb]
    DBuf<double> part;     // reduction partials
    DBuf<double> scal;     // device scalars
    double *nut = nullptr, *d = nullptr, *rhoPrev = nullptr, *fv1o = nullptr, *B = nullptr, *x2 = nullptr,
           *gn = nullptr, *bnut = nullptr;
    double* Sold = nullptr;  // SARC1: symm(grad U) of the previous correct() [9][nC]
    bool rhoPrevValid = false;
    int nCorrect = 0;
};

// cellLimited grad(U) in effect (k below 1e-15 leaves the gradient unlimited,
// as cellLimitedGrad::calcGrad returns the Gauss linear one then)
inline bool grad_u_limited(const rdst_settings& s) {
    return s.grad_u == RDST_GRAD_CELL_LIMITED && s.grad_u_limit >= 1.0e-15;
}

struct MeshDev {
    int nC, nI, nB, nF, ellW;
    const int32_t *owner, *neigh, *bcell, *bpatch, *bwall;
    const double *Sf, *magSf, *w, *dc, *V;
    const uint32_t* ellCode;
    const int32_t *ellFace, *ellSlot;
    const int32_t* slotDiag;
    const int32_t *slotUpper, *slotLower;  // per internal face: (owner row, neighbour col), (neighbour row, owner col)
    const double* C;                       // cell centres, SoA [3][nC]
    const double* Cf;                      // face centres, SoA [3][nF]
    const double *ellSf, *ellMagSf, *ellWt, *ellDc, *ellCorr;
};

// ---------------------------------------------------------------------------
struct Case {
    Mesh* mesh = nullptr;
    Thermo th{};
    rdst_settings st{};
    std::vector<rdst_bc> bcs;
    DBuf<BcPatchDev> d_bc;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float lastFluxMs = 0.f;

    // field storage
    DBuf<double> cellBuf, bndBuf, faceBuf, RBuf, WBuf, timeBuf, fHBuf, gradBuf, turbBuf, recBuf, cell8Buf;
    bool packDirty = true;  // cell8 is stale (a field was put from the host)
    Fields f{};
    // matrix
    DBuf<double> A64, src;          // [nnzb][25], [nC][5]
    DBuf<float> A32;                // FP32 copy for mixed precision
    // exact-Jacobian (Newton) mode: A holds the exact operator, P the Roe
    // operator whose ILU(0) preconditions it (FP64 / FP32)
    DBuf<double> P64;
    DBuf<float> P32;
    DBuf<double> S64;               // sliced-ELL copies for the SpMV
    DBuf<float> S32;
    bool sellFresh = false;  // S64/S32 hold the assembled operator (written by the assembly itself)
    bool csrStale = false;   // A64 does not hold the operator (the Newton operator is written sliced-ELL only)
    bool hasState = false, hasMatrix = false;
    // solver scratch owned by the solver module
    struct SolverWork* work = nullptr;
    // device copies of x and b of rdst_ldu_solve (allocated on first use)
    DBuf<double> lduX, lduB;
    // time-control scalars
    double residual = -1, residualRatio = 1, courantScalar = 1;
    // dual time stepping: old time levels of W [nC][5] and the physical time
    DBuf<double> WOld, WOldOld;
    double time = 0, deltaT = 0, deltaT0 = 0, deltaTSave = 0;
    int timeIndex = 0;
    std::vector<double> hostScratch;
    SaState sa;
    // discrete adjoint: exact J = dR/dW and its transpose on the block-CSR slots
    DBuf<double> adjJ, adjJT;
    DBuf<float> adjJT32;
    bool hasAdjJ = false;
    // viscous residual: J and J^T on the second-neighbour pattern (FP64, and
    // J^T in FP32 for the mixed solve); adjJT then holds the block-CSR part of
    // J^T, whose ILU(0) preconditions the solve
    DBuf<double> adjW, adjWT;
    DBuf<double> adjPT;  // the preconditioner's matrix with a pseudo-time shift
    DBuf<float> adjWT32;
    DBuf<double> avGrad, avR, avJv, avObj;  // scratch of the forward-mode passes
    bool adjWide = false;                   // the last linearisation is the second-neighbour one

    MeshDev meshDev() const;
};

// kernels / steps (defined in the .hip units)
// face-field store modes of the face loop
enum { STORE_BPHI = 0 /* boundary phi only */, STORE_PHI = 1 /* phi on every face */, STORE_ALL = 2 };
// exact distance of every cell centre to the wall-patch triangles (wall.hip)
void wall_distance(const Mesh& m, hipStream_t stream, double* d_host);
void flux_residual(Case& c, int storeMode);
void flux_launch(Case& c, int storeMode);
void grad_u_launch(Case& c);
void jacobian_assemble(Case& c);
void block_solve(Case& c, rdst_solver_perf* perf);
void ldu_solve(Case& c, const double* diag, const double* upper, const double* lower, const double* b, double* x,
               int mode, rdst_solver_perf* perf);
void update_fields(Case& c);
double set_courant(Case& c);
// setCourantDST with the inner-iteration residual |lin| (lin != nullptr) or
// |R|_2; pseudo: the pseudo-time branch, else the physical one
double set_courant_dst(Case& c, const double* lin, bool pseudo);
void set_transient_deltaT(Case& c);
void store_conserved(Case& c, double* out);  // rho, rhoU, rhoE as [nC][5]
void set_steady_state_deltaT(Case& c);
void set_state(Case& c, const double* p, const double* T, const double* U);
void solver_free(Case& c);
double flux_time(Case& c, int n);
// force integral of one patch; work: device scratch of forces_work_size()
// doubles, or nullptr to allocate it for the call
void forces(Case& c, int patch, const double* dragDir, const double* liftDir, double pRef, double rhoInf,
            double magUInf, double aRef, double* out, double* work);
size_t forces_work_size();
// the per-device address of the force integral's static device scratch,
// looked up once per device with that device current
void forces_scratch_setup(int device);
double* forces_scratch(int device);

void sa_set(Case& c, const rdst_turbulence* t, const rdst_scalar_bc* bcs, const double* wallDist,
            const double* nuTilda);
void sa_correct(Case& c, rdst_solver_perf* perf);
void sa_put_nut(Case& c, const double* nuTilda);
void adjoint_check_scheme(const Case& c);
void adjoint_jacobian(Case& c);   // J = dR/dW only
void newton_operator(Case& c);
void newton_operator_csr(Case& c);  // block-CSR copy of the FP64 Newton operator (read-back)    // V dR/dW + rDeltaTDST V I into the case's A64 / A32
void adjoint_linearise(Case& c);  // J and its transpose
void build_wide(Mesh& m);         // the second-neighbour pattern of the viscous linearisation
void adjoint_residual(Case& c, double* R);
void adjoint_objective_gradient(Case& c, int patch, const double* dir, double q, double* g);
void adjoint_objective_cells(Case& c, int patch, const double* dir, double q, double pRef, double* o);
void adjoint_apply(Case& c, bool transpose, const double* x, double* y);
void adjoint_solve(Case& c, const double* g, double* lambda, const rdst_adjoint_settings& as, rdst_solver_perf* perf);

inline int blocks_for(long n, int bs) { return (int)((n + bs - 1) / bs); }

// XCD-aware block order.  The dispatcher hands workgroups to the 8 XCDs
// round-robin (workgroup b runs on XCD b % 8), and each XCD has its own L2;
// mapping the workgroups of XCD x onto one contiguous range of logical
// blocks keeps the cells (and their mesh neighbours) that one XCD touches
// together in its L2.  A bijection of [0, nb).
// Surface-normal gradient scheme as one number for the kernels: -1 for
// corrected, 0 for uncorrected, psi for limited psi (rdst_sngrad_kind).
inline double sngrad_psi(int kind, double limit) {
    return kind == RDST_SNGRAD_CORRECTED ? -1.0 : (kind == RDST_SNGRAD_UNCORRECTED ? 0.0 : limit);
}

#ifdef __HIPCC__
// Factor on the non-orthogonal correction of a face gradient (foam-extend
// limitedSnGrad::correction): 1 when corrected (psi < 0), else
// min(psi |u| / ((1 - psi) |c| + SMALL), 1) with |u|^2 = u2 the squared
// magnitude of the orthogonal part dc (x_N - x_P), |c|^2 = c2 that of the
// correction; psi = 0 gives 0 (uncorrected).
__device__ inline double sngrad_limiter(double psi, double u2, double c2) {
    if (psi < 0.0) return 1.0;
    return fmin(psi * sqrt(u2) / ((1.0 - psi) * sqrt(c2) + 1.0e-15), 1.0);
}

__device__ inline int xcd_block(int b, int nb) {
    const int x = b & 7, k = b >> 3;
    const int per = nb >> 3, rem = nb & 7;
    return x * per + (x < rem ? x : rem) + k;
}
#endif

}  // namespace rdst
