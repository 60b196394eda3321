/*
 * rhodst.h -- C-ABI of the MI355X block-coupled implicit step.
 *
 * The entry points below replace the run-time-selected pieces that
 * rhoDST's Single::iterate calls (libraries/finiteVolume/iteration/Single/Single.C:80-104):
 *
 *   baseFluxDST::computeFluxDST + calcFields      -> rdst_flux_residual
 *     (fluxSchemes/baseFluxDST/baseFluxDST.H:176, baseFluxDST.C:223-255,
 *      selector newbaseFluxDST.C:49-85)
 *   JacobianDST<tensor5>::computeJacobianDST       -> rdst_jacobian_assemble
 *   + fvmDST::ddt + fluxDST::integrate
 *     (jacobian/Roe/Roe.C:152-379, fvmDST/fvmDST.C:59-143,
 *      timeIntegration/fluxDST/fluxDST/fluxDST.C:73-127)
 *   fvBlockMatrixDST<vector5>::solve               -> rdst_block_solve
 *     (rhoDST/fvBlockMatrixDST/fvBlockMatrixDST.C:98-155; the foam-extend
 *      BlockLduSolver<vector5> GMRES/BiCGStab + ILUC0 it selects)
 *   #include "updateFields.H" + BC updateCoeffs    -> rdst_update_fields
 *     (libraries/include/updateFields.H:40-94,
 *      boundaryConditions/farFieldDST/ * /updateCoeffs)
 *   timeDST::setCourantDST / setSteadyStateDeltaT  -> rdst_set_courant
 *     (rhoDST/timeDST/timeDST.C:218-273, :316-443)
 *   the whole Single::iterate                     -> rdst_iterate
 *
 * Conventions: every call returns 0 on success and a nonzero RDST_ERR_*
 * code otherwise; rdst_last_error() gives the message (the reference
 * aborts through FatalError instead, e.g. newbaseFluxDST.C:72-82).
 * Host arrays are caller-owned and only read during the call.  Device
 * state lives behind the opaque handles.  Calls are synchronous on return.
 * A handle is used from one host thread.
 *
 * Component order of the 5-vector W is (rho, rhoUx, rhoUy, rhoUz, rhoE)
 * (rhoDST/geometricFieldDST/geometricFieldDST.C:156-158); 5x5 blocks are
 * row-major, entry (r,c) at 5*r+c (rhoDST/DST/DST.C:240-242).
 */
#ifndef RHODST_H
#define RHODST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- codes */
enum rdst_status {
    RDST_OK = 0,
    RDST_ERR_ARG = 1,          /* bad argument / size */
    RDST_ERR_HIP = 2,          /* a HIP runtime call failed */
    RDST_ERR_UNKNOWN_NAME = 3, /* unknown run-time-selection name */
    RDST_ERR_STATE = 4,        /* call out of order (e.g. solve before assembly) */
    RDST_ERR_NO_DEVICE = 5     /* no GPU visible / extension unusable */
};

/* fluxScheme names, the baseFluxDST RTS table (newbaseFluxDST.C:64-84) */
enum rdst_flux_scheme {
    RDST_FLUX_AUSMPLUS = 0,   /* "AUSM+"    AUSMPLDST.C:127-217   */
    RDST_FLUX_AUSMPLUSUP = 1, /* "AUSM+up"  AUSMPLUPDST.C:209-320 */
    RDST_FLUX_HLL = 2,        /* "HLL"      HLLDST.C:91-204       */
    RDST_FLUX_HLLC = 3,       /* "HLLC"     HLLCDST.C:135-272     */
    RDST_FLUX_HLLCP = 4,      /* "HLLCP"    HLLCPDST.C:166-317    */
    RDST_FLUX_KURGANOV = 5    /* "Kurganov" KurganovDST.C:93-167  */
};

/* polyMesh patch geometric types (constant/polyMesh/boundary "type") */
enum rdst_patch_type {
    RDST_PATCH_GENERIC = 0,   /* patch */
    RDST_PATCH_WALL = 1,      /* wall (timeDST.C:254 treats these specially) */
    RDST_PATCH_SYMMETRY = 2,  /* symmetryPlane */
    RDST_PATCH_EMPTY = 3      /* empty: fvPatch of size 0 */
};

/* boundary-condition types for p, T and U */
enum rdst_bc_type {
    RDST_BC_FIXED_VALUE = 0,  /* fixedValue (wall no-slip = fixedValue 0) */
    RDST_BC_ZERO_GRADIENT = 1,
    RDST_BC_INLET_OUTLET = 2, /* mixed, valueFraction 1-pos(phi) */
    RDST_BC_SYMMETRY = 3,     /* symmetryPlane; for U also slip */
    RDST_BC_FARFIELD = 4,     /* farField{Pressure,Temperature,Velocity}DST */
    RDST_BC_EMPTY = 5
};

enum rdst_solver_kind { RDST_SOLVER_GMRES = 0, RDST_SOLVER_BICGSTAB = 1 };
enum rdst_precond_kind { RDST_PRECOND_ILU0 = 0, RDST_PRECOND_NONE = 1, RDST_PRECOND_DIAG = 2 };
enum rdst_dependent { RDST_DEP_W = 0, RDST_DEP_DELTAW = 1 };
/* arithmetic of the linear solve: plain FP64, or FP32 Jacobian/ILU/Krylov
 * inside an FP64 defect-correction loop */
enum rdst_precision { RDST_PREC_FP64 = 0, RDST_PREC_MIXED = 1 };
/* ordering of the block-ILU(0): natural cell order over the whole mesh
 * (the reference's serial ILUC0 factor, run level-scheduled), or
 * block-Jacobi ILU(0) over compact cell partitions of at most 1024 cells
 * (what the reference's processor-local ILUC0 computes when the case is
 * decomposed, one partition per rank) */
enum rdst_ilu_order { RDST_ILU_NATURAL = 0, RDST_ILU_PARTITIONED = 1 };
/* GMRES orthogonalisation: modified Gram-Schmidt as foam-extend's
 * BlockGMRESSolver, or classical Gram-Schmidt with one fused multi-dot
 * (same basis in exact arithmetic, about half the Krylov-vector traffic) */
enum rdst_gram_schmidt { RDST_GS_MODIFIED = 0, RDST_GS_CLASSICAL = 1 };
/* interpolationSchemes reconstruct(rho|U|T) (e.g. forwardStep/system/fvSchemes:41-44):
 * upwind (first order), or the limitedScheme OSPRE for the scalars with OSPREV
 * for rhoU (r-factor from the Gauss-linear cell gradients) */
enum rdst_reconstruction { RDST_RECON_UPWIND = 0, RDST_RECON_OSPRE = 1 };

typedef struct rdst_patch {
    char name[64];
    int32_t type;   /* rdst_patch_type */
    int32_t start;  /* first polyMesh face of the patch */
    int32_t size;   /* number of polyMesh faces */
    int32_t pad;
} rdst_patch;

/* polyMesh primitives, as in constant/polyMesh/{points,faces,owner,neighbour,boundary} */
typedef struct rdst_mesh_desc {
    int32_t n_points;
    int32_t n_faces;
    int32_t n_internal_faces;
    int32_t n_cells;
    int32_t n_patches;
    int32_t pad;
    const double* points;        /* [n_points][3] */
    const int32_t* face_offsets; /* [n_faces+1] */
    const int32_t* face_points;  /* [face_offsets[n_faces]] */
    const int32_t* owner;        /* [n_faces] */
    const int32_t* neighbour;    /* [n_internal_faces] */
    const rdst_patch* patches;   /* [n_patches] */
} rdst_mesh_desc;

/* per-patch boundary conditions for p, T and U (0/{p,T,U}) */
typedef struct rdst_bc {
    int32_t p_type, T_type, U_type, pad;
    double p_value, T_value;      /* fixedValue value / inletOutlet inletValue */
    double U_value[3];
    double p_inf, T_inf;          /* farField*DST pInf, TInf (P.C:80-82) */
    double U_inf[3];              /* farField*DST UInf */
} rdst_bc;

/* hPsiThermo<pureMixture<{const,sutherland}Transport<specieThermo<hConstThermo<perfectGas>>>>> */
typedef struct rdst_thermo {
    double Cp;      /* J/kg/K, constant */
    double R;       /* specific gas constant RR/W */
    int32_t transport; /* 0 = constTransport (mu), 1 = sutherlandTransport (As, Ts) */
    int32_t pad;
    double mu;      /* constTransport mu */
    double Pr;      /* Prandtl number (alpha = mu/Pr) */
    double As, Ts;  /* sutherland coefficients: mu = As*sqrt(T)/(1+Ts/T) */
} rdst_thermo;

/* system/rhoDSTDict, fvSolution.solvers.W and fvSolution.fieldBounds */
typedef struct rdst_settings {
    int32_t flux_scheme;      /* rdst_flux_scheme */
    int32_t low_mach;         /* AUSM+up lowMach (AUSMPLUPDST.C:195) */
    double Mainf;             /* AUSM+up Mainf, default 0.5 */
    int32_t dependent;        /* rdst_dependent, default W (IOdictDST.C:147-150) */
    int32_t neighbouring;     /* timeControls.neighbouring (IOdictDST.C:197-200) */
    double beta;              /* timeIntegration.beta, default 1 */
    double initial_co, min_co, max_co;
    int32_t solver;           /* rdst_solver_kind */
    int32_t precond;          /* rdst_precond_kind */
    int32_t n_directions;     /* GMRES restart length */
    int32_t min_iter;
    int32_t max_iter;
    int32_t precision;        /* rdst_precision */
    double tolerance, rel_tol;
    double rho_min, rho_max, p_min, p_max, T_min, T_max; /* fieldBounds */
    int32_t ilu_order;        /* rdst_ilu_order */
    int32_t viscous;          /* evaluate the viscous terms of calcFields */
    int32_t gram_schmidt;     /* rdst_gram_schmidt: GMRES orthogonalisation */
    int32_t reconstruction;   /* rdst_reconstruction */
    int32_t jacobian;         /* rdst_jacobian_kind: the implicit operator */
    int32_t time_integration; /* rdst_time_integration: timeIntegration.timeIntegration (IOdictDST.C:206-209) */
    int32_t adjust_time_step; /* timeControls.adjustTimeStep, default 1 (IOdictDST.C:137-140) */
    int32_t max_inner_iter;   /* timeIntegration.maxInnerIter, default 20 (IOdictDST.C:222-225) */
    int32_t iteration_start;  /* residualControls.iterationStart, default 10 when unsteady (IOdictDST.C:202-210) */
    double mag_converged_slope; /* residualControls.magConvergedSlope, default 0.1 (IOdictDST.C:212-215) */
    double min_residual;      /* residualControls.minResudual, default 1e-5 (IOdictDST.C:217-220) */
    double delta_t;           /* controlDict deltaT: the deltaT0 of the first physical step */
    /* fvSchemes: the surface-normal gradient inside laplacianSchemes (used by
     * fvc::laplacian(muEff, U) of calcFields, baseFluxDST.C:248, and by
     * fvm::laplacian(DnuTildaEff, nuTilda) of the Spalart-Allmaras models)
     * and snGradSchemes (fvc::snGrad(U) of sigmaDotU, baseFluxDST.C:240-246);
     * rdst_sngrad_kind, 0 = corrected (every reference tutorial's snGrad,
     * tutorials/oneraM6Turb/system/fvSchemes:44 uses "limited 0.33" for the
     * laplacian).  *_limit is psi of "limited psi". */
    int32_t laplacian_sngrad;
    int32_t sngrad;
    double laplacian_limit;
    double sngrad_limit;
    /* Heat conduction in the energy residual (no counterpart in the
     * reference: calcFields, baseFluxDST.C:223-255, has no heat-flux term).
     * 1 adds -laplacian(alphaEff, h) with its orthogonal part,
     * alphaEff_f |Sf| dc (h_N - h_P) (boundary faces: h_b - h_P), alphaEff =
     * alpha + alphat; 0 (default) is the reference's energy equation. */
    int32_t heat_conduction;
    /* Update limit (no counterpart in the reference; default 0 = off): when
     * > 0, each cell's update of W in updateFields is scaled so that rho and
     * rhoE change by at most this fraction of their values before the update
     * (a per-cell damping for the exact-Jacobian operator at large CFL). */
    double update_limit;
    /* fvSchemes gradSchemes grad(U) (rdst_grad_kind; tutorials/RAE2822/
     * system/fvSchemes:27 selects "cellLimited Gauss linear 1"): the gradient
     * of U that tauMC (baseFluxDST.C:233-236), the Spalart-Allmaras strain and
     * vorticity (SAneg.C:524, SARC2.C:458) and the force coefficients read.
     * grad_u_limit is the k of cellLimited, in [0, 1]; k below 1e-15 leaves
     * the gradient unlimited.  The non-orthogonal correction of snGrad(U) is
     * formed from the unlimited Gauss linear gradient (foam-extend 4.1
     * correctedSnGrad reads the gradScheme of "snGradCorr(U)", the default). */
    int32_t grad_u;
    int32_t grad_u_pad;
    double grad_u_limit;
} rdst_settings;

/* foam-extend 4.1 gradSchemes for grad(U).  GAUSS_LINEAR: sum_f Sf U_f / V,
 * U_f linearly interpolated (the boundary value on boundary faces).
 * CELL_LIMITED (cellLimitedGrad): the Gauss linear gradient with column j
 * (dU_j/dx_i, i = 0..2) scaled by l_j in [0, 1], the largest factor for which
 * U_j + (Cf - C) & grad stays within [min, max] of U_j over the cell, its
 * face neighbours and its boundary-face values at every face centre Cf of
 * the cell; for k < 1 that interval is widened on both sides by
 * (1/k - 1)(max - min).  Per face and component: where the extrapolated
 * change e exceeds max - U_j (by more than 1e-300), l_j = min(l_j,
 * (max - U_j)/e); where it is below min - U_j, l_j = min(l_j, (min - U_j)/e).
 * The boundary gradient is then corrected as for Gauss linear
 * (gaussGrad::correctBoundaryConditions). */
enum rdst_grad_kind { RDST_GRAD_GAUSS_LINEAR = 0, RDST_GRAD_CELL_LIMITED = 1 };

/* Non-orthogonal correction of a surface-normal gradient (foam-extend 4.1
 * snGradSchemes).  CORRECTED: dc (x_N - x_P) + corr & interp(grad x), with
 * corr = n - dc d.  UNCORRECTED: dc (x_N - x_P).  LIMITED psi (limitedSnGrad):
 * the correction scaled by min(psi |dc (x_N - x_P)| / ((1 - psi) |correction|
 * + 1e-15), 1), |.| the magnitude of the (vector) face value, so the
 * correction never exceeds psi / (1 - psi) of the orthogonal part; psi in
 * [0, 1]. */
enum rdst_sngrad_kind {
    RDST_SNGRAD_CORRECTED = 0,
    RDST_SNGRAD_UNCORRECTED = 1,
    RDST_SNGRAD_LIMITED = 2
};

/* Time integration of the case (timeIntegration.timeIntegration).  SINGLE:
 * the steady pseudo-time step Single::iterate (steadyStateDST ddt, local time
 * stepping).  DUAL: dual time stepping Dual::iterate (Dual.C:80-124): each
 * physical step of dual2ndDST (second-order backward differences in physical
 * time, dual2ndDST.C:377-411) is converged by inner pseudo-time iterations
 * until steadyCheck (steadyCheck.C:69-98) declares them converged; the
 * physical time step follows setTransientDeltaT (timeDST.C:275-314).  DUAL
 * runs with precision FP64, the Roe Jacobian and no turbulence model. */
enum rdst_time_integration {
    RDST_TIME_SINGLE = 0,
    RDST_TIME_DUAL = 1
};

/* The implicit operator of the step.  ROE is the reference's Jacobian
 * (Roe<tensor5>, the only entry of the JacobianDST table, JacobianDST.C:57-92).
 * EXACT (no counterpart in the reference) assembles V/dtau I + V dR/dW with
 * the exact linearisation of the first-order inviscid residual (the
 * adjoint's, rdst_adjoint_linearise) plus, with viscous terms on, the compact
 * (orthogonal, muEff frozen) linearisation of the viscous fluxes: a Newton
 * step as dtau grows; upwind reconstruction only, no HLLCP. */
enum rdst_jacobian_kind {
    RDST_JACOBIAN_ROE = 0,
    RDST_JACOBIAN_EXACT = 1
};

/* compressible::RASModel selected by constant/RASProperties RASModel */
enum rdst_turbulence_model {
    RDST_TURB_LAMINAR = 0,
    RDST_TURB_SANEG = 1,   /* SAneg (libraries/turbulence/SAneg/SAneg.C) */
    RDST_TURB_SARC2 = 2,   /* SARC2 (libraries/turbulence/SARC2/SARC2.C) */
    RDST_TURB_SARC1 = 3,   /* SARC1 (libraries/turbulence/SARC1/SARC1.C): Spalart-Shur fr1 with DS/Dt */
    RDST_TURB_SA = 4       /* foam-extend's stock SpalartAllmaras (every turbulent tutorial,
                              e.g. RAE2822/constant/RASProperties:19): the fv3 form SARC2.C:418-488 is
                              built on, without the rotation-curvature multiplier (fr1 = 1) */
};

/* RASProperties <model>Coeffs (SAneg.C:168-307, SARC2.C:172-285), plus
 * fvSolution solvers.nuTilda (smoothSolver, symGaussSeidel) and
 * relaxationFactors.nuTilda (tutorials/RAE2822/system/fvSolution:43-57) */
typedef struct rdst_turbulence {
    int32_t model;                /* rdst_turbulence_model */
    int32_t negative_nu_tilda;    /* SAneg negativeNuTilda (default on) */
    int32_t stilda_modification;  /* SAneg StildaModification (default on) */
    int32_t rcm_correction;       /* SARC2 RCMCorrection / SARC1 spalartShurCorrection (default on) */
    double sigma_nut, kappa, Prt; /* 0.66666, 0.41, 1 */
    double Cb1, Cb2, Cw2, Cw3, Cv1, Cv2;  /* 0.1355 0.622 0.3 2 7.1 5 */
    double Cn1, Cn2, Cn3, Ct3, Ct4;       /* SAneg: 16 0.7 0.9 1.2 0.5 */
    double Cr1, Cr2, Cr3;                 /* SARC2: 1 2 1; SARC1: 1 12 1 */
    double relax;                 /* nuTilda relaxation factor (<= 0: none) */
    double tolerance, rel_tol;    /* smoothSolver tolerance / relTol */
    int32_t min_iter, max_iter;   /* symmetric Gauss-Seidel sweeps */
    int32_t convection;           /* divSchemes div(phi,nuTilda): 0 Gauss upwind, 1 Gauss linearUpwind
                                     grad(nuTilda) (upwind matrix, explicit (Cf - C_upwind).grad(nuTilda)
                                     correction in the source, Gauss linear gradient) */
    int32_t pad;
} rdst_turbulence;

/* 0/nuTilda boundary condition of one patch (the wall patches get mut = 0) */
typedef struct rdst_scalar_bc {
    int32_t type;   /* RDST_BC_FIXED_VALUE, _ZERO_GRADIENT, _INLET_OUTLET, _SYMMETRY, _EMPTY */
    int32_t pad;
    double value;   /* fixedValue value / inletOutlet inletValue */
} rdst_scalar_bc;

/* BlockSolverPerformance<vector5>: per-component initial/final residuals */
typedef struct rdst_solver_perf {
    double initial_residual[5];
    double final_residual[5];
    int32_t n_iterations;
    int32_t converged;
} rdst_solver_perf;

/* fields addressable through rdst_case_get / rdst_case_put */
enum rdst_field {
    RDST_F_RHO = 0, RDST_F_RHOU = 1, RDST_F_RHOE = 2,   /* cells; rhoU is [n][3] */
    RDST_F_P = 3, RDST_F_T = 4, RDST_F_U = 5, RDST_F_H = 6, RDST_F_PSI = 7,
    RDST_F_R = 8,                 /* residual R_ [n_cells][5] */
    RDST_F_W = 9,                 /* solution vector W [n_cells][5] */
    RDST_F_RDELTAT = 10, RDST_F_COURANT = 11,
    RDST_F_PHI = 12,              /* mass flux, internal then boundary faces */
    RDST_F_PHIUP = 13,            /* [nf][3] */
    RDST_F_PHIEP = 14,
    RDST_F_UP = 15,               /* [nf][3] */
    RDST_F_BRHO = 16, RDST_F_BU = 17, RDST_F_BP = 18, RDST_F_BT = 19, /* boundary faces */
    RDST_F_BRHOU = 20, RDST_F_BRHOE = 21, RDST_F_BH = 22, RDST_F_BPSI = 23,
    RDST_F_BCSR_VALUES = 24,      /* assembled FP64 blocks [nnzb][25] */
    RDST_F_SOURCE = 25,           /* assembled source [n_cells][5] */
    RDST_F_MU = 26,
    RDST_F_RHO_OLD = 27, RDST_F_E_OLD = 28,
    RDST_F_NU_TILDA = 29,         /* cells */
    RDST_F_MUT = 30,              /* cells */
    RDST_F_ALPHAT = 31,           /* cells */
    RDST_F_WALL_DIST = 32,        /* cells: the wallDist d (supplied, or rdst_mesh_wall_distance) */
    RDST_F_BNU_TILDA = 33,        /* boundary faces */
    RDST_F_BMUT = 34,             /* boundary faces */
    RDST_F_SA_DIAG = 35,          /* nuTilda matrix after relax(): diagonal incl. boundary coefficients [n_cells] */
    RDST_F_SA_UPPER = 36,         /* [n_internal] (owner row, neighbour column) */
    RDST_F_SA_LOWER = 37,         /* [n_internal] (neighbour row, owner column) */
    RDST_F_SA_SOURCE = 38,        /* [n_cells] incl. boundary coefficients and relaxation */
    RDST_F_GRAD_U = 39            /* cells [n][9]: fvc::grad(U) of a viscous case as last formed (by the
                                     residual evaluation or the turbulence model's correct step), with
                                     settings.grad_u applied; dU_j/dx_i at 3 i + j */
};

/* integer arrays of a mesh */
enum rdst_mesh_array {
    RDST_M_OWNER = 0,       /* [n_internal + n_bfaces]: internal owners, then boundary face cells */
    RDST_M_NEIGHBOUR = 1,   /* [n_internal] */
    RDST_M_BCSR_ROWPTR = 2, /* [n_cells+1] */
    RDST_M_BCSR_COLS = 3,   /* [nnzb] */
    RDST_M_SLOT_UPPER = 4,  /* [n_internal] block slot of (owner row, neighbour col) */
    RDST_M_SLOT_LOWER = 5,  /* [n_internal] block slot of (neighbour row, owner col) */
    RDST_M_SLOT_DIAG = 6,   /* [n_cells] */
    RDST_M_FACE_COLOUR = 7, /* [n_internal] */
    RDST_M_ILU_LEVEL = 8,   /* [n_cells] level of the lower-triangular sweep */
    RDST_M_BFACE_PATCH = 9, /* [n_bfaces] patch index of each non-empty boundary face */
    RDST_M_CELL_COLOUR = 10, /* [n_cells] */
    RDST_M_PARTITION = 11,  /* [n_cells] partition of each cell (partitioned ILU) */
    RDST_M_SA_PARTITION = 12 /* [n_cells] partition of the nuTilda Gauss-Seidel sweep (the RCB boxes of at
                                most 1024 cells before the ILU's level cut; one box = the serial sweep) */
};

/* double arrays of a mesh */
enum rdst_mesh_geom {
    RDST_G_SF = 0,       /* [n_internal + n_bfaces][3] */
    RDST_G_MAGSF = 1,
    RDST_G_CF = 2,       /* [..][3] */
    RDST_G_C = 3,        /* [n_cells][3] */
    RDST_G_V = 4,        /* [n_cells] */
    RDST_G_WEIGHTS = 5,  /* [n_internal + n_bfaces] (1 on boundary faces) */
    RDST_G_DELTACOEFFS = 6
};

/* integer mesh sizes returned by rdst_mesh_info */
enum rdst_mesh_info_idx {
    RDST_I_NCELLS = 0, RDST_I_NINTERNAL = 1, RDST_I_NBFACES = 2, RDST_I_NNZB = 3,
    RDST_I_NCOLOURS = 4, RDST_I_NLEVELS_L = 5, RDST_I_NLEVELS_U = 6, RDST_I_NCELL_COLOURS = 7,
    RDST_I_NPARTS = 8, RDST_I_MAX_PART_LEVELS = 9,
    RDST_I_COUNT = 10
};

typedef struct rdst_mesh rdst_mesh;
typedef struct rdst_case rdst_case;

/* ----------------------------------------------------------- library */
const char* rdst_last_error(void);
int rdst_version(void);
/* select the HIP device used by later calls */
int rdst_init(int device);

/* ----------------------------------------------------------- mesh
 * Copies the polyMesh primitives, computes the finite-volume geometry on
 * the host (fvMesh: Sf, magSf, Cf, C, V, weights, deltaCoeffs), the
 * owner/neighbour-sorted block-CSR with its face->slot maps, a face
 * colouring and the ILU level sets (host only); the first rdst_case_create
 * on the mesh uploads it to the GPU.  Replaces the fvMesh the reference's
 * operators read through mesh_. */
int rdst_mesh_create(const rdst_mesh_desc* desc, rdst_mesh** out);

/* The finite-volume geometry of the host's fvMesh, in polyMesh face order
 * (internal faces, then the faces of every patch; entries of empty-patch
 * faces are ignored): what the reference's operators read through mesh_
 * (Sf, magSf, weights, deltaCoeffs: Roe.C:195-199; V: timeDST.C:237). */
typedef struct rdst_mesh_geometry {
    const double* Sf;           /* [n_faces][3] face area vectors */
    const double* mag_sf;       /* [n_faces] |Sf|, or NULL to form it from Sf */
    const double* Cf;           /* [n_faces][3] face centres */
    const double* C;            /* [n_cells][3] cell centres */
    const double* V;            /* [n_cells] cell volumes */
    const double* weights;      /* [n_faces] interpolation weights (boundary entries ignored: 1) */
    const double* delta_coeffs; /* [n_faces] deltaCoeffs (boundary: 1/|Cf - C|) */
} rdst_mesh_geometry;

/* rdst_mesh_create with the caller's geometry kept bit for bit instead of
 * the one computed from the points (the non-orthogonal correction vectors
 * n - d deltaCoeffs are formed from it).  The points still define the wall
 * triangles of rdst_mesh_wall_distance. */
int rdst_mesh_create_geometry(const rdst_mesh_desc* desc, const rdst_mesh_geometry* geom, rdst_mesh** out);
int rdst_mesh_destroy(rdst_mesh* mesh);
int rdst_mesh_info(const rdst_mesh* mesh, int64_t* info /* [RDST_I_COUNT] */);
int rdst_mesh_get_int(const rdst_mesh* mesh, int which, int32_t* out);
int rdst_mesh_get_geom(const rdst_mesh* mesh, int which, double* out);
/* d[n_cells]: exact distance of every cell centre to the faces of the wall
 * patches (RDST_PATCH_WALL), computed on the device.  Replaces foam-extend's
 * wallDist (meshWave; held as d_ by SAneg.C:160 and SARC2.C:158), whose
 * propagated distance is exact only next to the wall. */
int rdst_mesh_wall_distance(const rdst_mesh* mesh, double* d);

/* ----------------------------------------------------------- case
 * A case = mesh + thermo + per-patch BCs + rhoDSTDict/fvSolution settings
 * + the device-resident state.  rdst_case_create validates the names the
 * reference selects at run time (flux scheme, solver, preconditioner) and
 * fails with RDST_ERR_UNKNOWN_NAME on an unknown id. */
int rdst_case_create(rdst_mesh* mesh, const rdst_thermo* thermo,
                     const rdst_bc* bcs /* [n_patches] */,
                     const rdst_settings* settings, rdst_case** out);
int rdst_case_destroy(rdst_case* c);
/* createFields.H:21-72 + timeDST::initialize (timeDST.C:67-208):
 * from cell p, T, U ([n][3]) build rho, rhoU, rhoE, h, psi, boundary
 * values, phi = interpolate(rho*U)&Sf, W, courant and rDeltaTDST */
int rdst_case_set_state(rdst_case* c, const double* p, const double* T, const double* U);
int rdst_case_get(rdst_case* c, int field, double* out);
/* overwrite a field: cells RDST_F_RHO, _RHOU, _RHOE, _W, _RDELTAT, _NU_TILDA,
 * _MUT, _ALPHAT; faces RDST_F_PHI (the stored mass flux the boundary
 * conditions read); boundary faces RDST_F_BMUT (alphat_b = mut_b / Prt with a
 * turbulence model on) -- the last four restore a state or carry one, with
 * the turbulence frozen, to another mesh */
int rdst_case_put(rdst_case* c, int field, const double* in);

/* computeFluxDST + calcFields: fills phi, phiUp, phiEp, Up and R_ */
int rdst_flux_residual(rdst_case* c);
/* computeJacobianDST + fvmDST::ddt + fluxDST::integrate into the BCSR
 * (FP64, and the FP32 copy when precision is mixed) and the source */
int rdst_jacobian_assemble(rdst_case* c);
/* fvBlockMatrixDST::solve: W mode solves A.W = A.W + s; deltaW mode
 * solves A.dW = s then W += dW */
int rdst_block_solve(rdst_case* c, rdst_solver_perf* perf);

/* The linear solve on a system the caller assembled, in foam-extend's LDU
 * order: what fvBlockMatrixDST<vector5>::solve hands to
 * BlockLduSolver<vector5>::New(name, A, dict)->solve(x, b)
 * (fvBlockMatrixDST.C:112-123, :136-144; the GMRES/BiCGStab + ILUC0 solvers
 * of Vector5typedefs.C:1596-1711).  diag [n_cells][25], upper and lower
 * [n_internal_faces][25] (tensor5 row-major, entry (r,c) at 5r+c) in
 * polyMesh internal-face order: the owner row holds upper x x_neighbour,
 * the neighbour row lower x x_owner (foam's LDU convention).  b and x are
 * [n_cells][5] (vector5).  The case's settings choose the solver,
 * preconditioner, nDirections, tolerances and precision (mixed: FP32 ILU and
 * Krylov cycles in an FP64 defect correction on this FP64 matrix).  The
 * coefficients replace the case's assembled matrix.  Host arrays cross
 * PCIe on every call (the matrix is ~2.8 GB at 2 M cells).
 *   RDST_LDU_PLAIN:  A x = b from the initial guess x (BlockLduSolver::solve)
 *   RDST_LDU_W:      A x' = A x + b from x, x holding W and b the source
 *                    (fvBlockMatrixDST::solve, dependent W)
 *   RDST_LDU_DELTAW: A d = b from 0, then x += d (dependent deltaW) */
enum rdst_ldu_mode { RDST_LDU_PLAIN = 0, RDST_LDU_W = 1, RDST_LDU_DELTAW = 2 };
int rdst_ldu_solve(rdst_case* c, const double* diag, const double* upper, const double* lower, const double* b,
                   double* x, int mode, rdst_solver_perf* perf);
/* updateFields.H: primitives from W, bounding, BC updates */
int rdst_update_fields(rdst_case* c);
/* timeDST::setCourantDST: ||R||_2, CFL ramp, setSteadyStateDeltaT;
 * returns ||R||_2 (timeDST.C:328-331) in *res_norm */
int rdst_set_courant(rdst_case* c, double* res_norm);
/* Single::iterate, n_steps times; res_norm[n_steps] receives ||R||_2 of
 * each step and perf[n_steps] (may be NULL) the solver performance */
int rdst_iterate(rdst_case* c, int n_steps, double* res_norm, rdst_solver_perf* perf);

/* One physical time step of a DUAL case: runTime++ (rhoDST.C:73: time +=
 * deltaT, deltaT0 = the previous deltaT, the old time levels of W shifted),
 * then Dual::iterate (Dual.C:80-124): inner iterations of flux, dual2ndDST +
 * flux Jacobian assembly, block solve, updateFields and the pseudo-time
 * setCourantDST(&solverPerf), until steadyCheck converges; then
 * setCourantDST() with |R|_2, which sets the next physical deltaT when
 * adjust_time_step is on.  n_inner receives the number of inner iterations,
 * inner_residuals (NULL or max_inner_iter + 1 entries) the mean of the five
 * initial residuals of each inner solve, res_norm (NULL or one value) the
 * |R|_2 of the last setCourantDST.  RDST_ERR_STATE for a SINGLE case.
 * When the call fails part-way, the time, deltaT0 and time index go back to
 * their values before the call and the case has no valid state until
 * rdst_case_set_state (later calls return RDST_ERR_STATE).  An FP64 case
 * with dependent deltaW starts each inner solve from zero, as the reference
 * does, so the normalised initial residuals steadyCheck reads stay near 1
 * (a flat history, not the decrease the W form shows); dependent W is the
 * usual choice with Dual.  Mixed precision is refused for the same reason
 * (its correction form also starts from zero). */
int rdst_dual_step(rdst_case* c, int32_t* n_inner, double* inner_residuals, double* res_norm);

/* time, deltaT, deltaT0 and the physical courant number of the case */
int rdst_time_info(const rdst_case* c, double* out4);

/* Spalart-Allmaras turbulence on the case: model coefficients, the 0/nuTilda
 * boundary conditions [n_patches], the wall distance d [n_cells] (wallDist;
 * NULL: rdst_mesh_wall_distance on the case's mesh) and the initial nuTilda
 * [n_cells]; computes mut
 * and alphat = mut/Prt, which then enter muEff/alphaEff of calcFields
 * (baseFluxDST.C:231) and of the Roe spectral radius (Roe.C:187-199).
 * RDST_TURB_LAMINAR switches the model off (mut = 0). */
int rdst_case_set_turbulence(rdst_case* c, const rdst_turbulence* turb, const rdst_scalar_bc* nu_tilda_bcs,
                             const double* wall_dist, const double* nu_tilda);
/* RASModel::correct() (SAneg.C:490-605, SARC2.C:418-488): assemble the
 * nuTilda equation (steadyStateDST ddt with rDeltaTDST, upwind convection
 * with phi, laplacian of DnuTildaEff, Cb2 gradient term, production and
 * implicit destruction), relax, solve by symmetric Gauss-Seidel sweeps to
 * tolerance/relTol, bound, update the boundary values, mut and alphat.
 * perf->initial_residual[0]/final_residual[0] and n_iterations describe the
 * scalar solve.  rdst_iterate calls it after rdst_update_fields when a model
 * is set (Single.C:102). */
int rdst_turbulence_correct(rdst_case* c, rdst_solver_perf* perf);

/* ----------------------------------------------------------- discrete adjoint
 * New (no counterpart in the reference): exact linearisation of the
 * first-order residual R(W) = (1/V) sum_f s_f F_f (baseFluxDST.C:223-255 with
 * the W -> primitive chain of updateFields.H:40-94 and the boundary
 * conditions at the same state), its transpose, and the adjoint solve for
 * force-coefficient sensitivities dJ/dalpha = dJ/dalpha|_W - lambda^T
 * dR/dalpha|_W with J^T lambda = dJ/dW.  Flux schemes AUSM+, AUSM+up, HLL,
 * HLLC, Kurganov (HLLCP's sensor couples second neighbours).  With
 * settings.viscous the viscous terms of calcFields are included exactly
 * (Gauss gradient of U and its corrected boundary value, the laplacian with
 * its non-orthogonal correction, Sf & tauMC, the work on the interface
 * velocity, laminar mu(T) and alpha(T)); J then couples second neighbours and
 * is formed on that pattern, its face-neighbour part feeding the ILU(0).  A
 * turbulence model's mut and alphat are frozen at their stored values, and
 * the objective's force includes the viscous part. */
typedef struct rdst_adjoint_settings {
    int32_t precision;      /* RDST_PREC_FP64, or RDST_PREC_MIXED: FP32 transposed ILU and Krylov cycles
                               inside an FP64 defect correction on the exact FP64 J^T */
    int32_t n_directions;   /* GMRES restart length */
    int32_t max_cycles;     /* restart cycles */
    int32_t inner_operator; /* mixed precision: 0 the inner FP32 cycles solve with J^T itself; 1 with the
                               preconditioner's matrix (J^T's face-neighbour part plus the precond_shift
                               diagonal): a defect correction by that pseudo-time operator, each outer
                               FP64 residual against the exact J^T */
    double tolerance, rel_tol;  /* on the normalised residual |r| as for the primal solve */
    double precond_shift;       /* the ILU(0) is taken of J^T (viscous: of its face-neighbour part) plus
                                   precond_shift x the case's pseudo-time diagonal rDeltaT (the primal's
                                   implicit operator at its current CFL for 1); 0: of J^T itself */
} rdst_adjoint_settings;

/* J = dR/dW at the current state (and J^T), kept on the device */
int rdst_adjoint_linearise(rdst_case* c);
/* y = J x (transpose = 0) or y = J^T x (transpose = 1) with the last
 * linearisation; x, y host [n_cells][5] */
int rdst_adjoint_apply(rdst_case* c, int transpose, const double* x, double* y);
/* R(W) [n_cells][5] by the residual chain that rdst_adjoint_linearise differentiates */
int rdst_adjoint_residual(rdst_case* c, double* R);
/* dJ/dW [n_cells][5] of J = sum_{faces of patch} (p_b - p_ref) (Sf . dir) / q
 * (+ Sf & devRhoReff_b . dir / q with settings.viscous, as rdst_forces) */
int rdst_adjoint_objective_gradient(rdst_case* c, int patch, const double* dir, double q, double* g);
/* each cell's share of J (values, [n_cells]): its faces on the patch,
 * (p_b - p_ref)(Sf . dir)/q (+ the viscous force with settings.viscous); the
 * shares of cells near a displaced wall node give that node's dJ/dx|_W by
 * differences without summing the whole patch */
int rdst_adjoint_objective_cells(rdst_case* c, int patch, const double* dir, double q, double p_ref, double* out);
/* linearises at the current state, then solves J^T lambda = dJ/dW; lambda
 * [n_cells][5] holds the initial guess (zeros for none) and receives the solution.
 * With settings.viscous GMRES is preconditioned from the right (it minimises
 * the true residual, the ILU's matrix being only the face-neighbour part). */
int rdst_adjoint_solve(rdst_case* c, int patch, const double* dir, double q, const rdst_adjoint_settings* s,
                       double* lambda, rdst_solver_perf* perf);

/* forceCoeffs-style integral over one patch: force = sum (p_b - p_ref) Sf
 * + sum Sf & devRhoReff_b with devRhoReff = -muEff dev(twoSymm(grad U))
 * (the viscous part when settings.viscous); out[0]=Cd, out[1]=Cl,
 * out[2..4]=force, out[5..7]=its viscous part */
int rdst_forces(rdst_case* c, int patch, const double* drag_dir, const double* lift_dir,
                double p_ref, double rho_inf, double mag_u_inf, double a_ref, double* out);

/* rdst_forces without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory during
 * the call (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments, kernels and result as
 * rdst_forces; the block partials and sums live in a static device array of
 * 6 x 1024 + 8 doubles (49 216 bytes) that each GPU holds until the process
 * ends, found with hipGetSymbolAddress by rdst_init (or by the first call on
 * a device).  The partials are grid-stride sums over at most 1024 blocks, so
 * every patch size fits and none is rejected.  All calls on one GPU share
 * that array: calls that overlap on one GPU (two host threads, or two cases
 * on different streams) race and must be serialised by the caller. */
int rdst_forces_noalloc(rdst_case* c, int patch, const double* drag_dir, const double* lift_dir,
                        double p_ref, double rho_inf, double mag_u_inf, double a_ref, double* out);

/* device time of the last rdst_flux_residual face kernel(s), in ms (HIP events) */
int rdst_last_flux_ms(rdst_case* c, double* ms);
/* run the face-flux + residual kernel n times back to back on the case's
 * stream and report its mean device time per launch (HIP events); the
 * gradients it reads (viscous terms, limited reconstruction) are formed
 * once before the timed launches */
int rdst_time_flux(rdst_case* c, int n, double* ms_per_call);

#ifdef __cplusplus
}
#endif

#endif /* RHODST_H */
