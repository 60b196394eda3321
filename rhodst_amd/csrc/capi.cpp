// capi.cpp -- the extern "C" boundary declared in include/rhodst.h.
//
// Every entry point converts exceptions to an RDST_ERR_* code and keeps the
// message for rdst_last_error(); the reference aborts through FatalError
// instead (e.g. newbaseFluxDST.C:72-82 lists the valid fluxScheme names).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "rdst_internal.hpp"

namespace rdst {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

MeshDev Case::meshDev() const {
    const Mesh& m = *mesh;
    MeshDev d;
    d.nC = m.nC; d.nI = m.nI; d.nB = m.nB; d.nF = m.nF; d.ellW = m.ellW;
    d.owner = m.d_owner.p; d.neigh = m.d_neigh.p; d.bcell = m.d_bcell.p; d.bpatch = m.d_bpatch.p; d.bwall = m.d_bWall.p;
    d.Sf = m.d_Sf.p; d.magSf = m.d_magSf.p; d.w = m.d_w.p; d.dc = m.d_dc.p; d.V = m.d_V.p;
    d.ellCode = m.d_ellCode.p; d.ellFace = m.d_ellFace.p; d.ellSlot = m.d_ellSlot.p; d.slotDiag = m.d_slotDiag.p;
    d.slotUpper = m.d_slotUpper.p; d.slotLower = m.d_slotLower.p;
    d.C = m.d_C.p;
    d.Cf = m.d_Cf.p;
    d.ellSf = m.d_ellSf.p; d.ellMagSf = m.d_ellMagSf.p; d.ellWt = m.d_ellWt.p; d.ellDc = m.d_ellDc.p;
    d.ellCorr = m.d_ellCorr.p;
    return d;
}

static void alloc_fields(Case& c) {
    const Mesh& m = *c.mesh;
    const size_t n = m.nC, nB = std::max(m.nB, 1), nF = m.nF;
    c.cellBuf.alloc(14 * n);
    c.bndBuf.alloc(21 * nB);
    const bool lim = grad_u_limited(c.st);
    c.gradBuf.alloc(9 * n + 9 * nB + (lim ? 9 * n : 0));
    c.faceBuf.alloc(8 * nF);
    c.RBuf.alloc(5 * n);
    c.WBuf.alloc(5 * n);
    c.timeBuf.alloc(7 * n + 4096);
    c.fHBuf.alloc(n);
    c.src.alloc(5 * n);
    if (c.st.time_integration == RDST_TIME_DUAL) {
        c.WOld.alloc(5 * n);
        c.WOldOld.alloc(5 * n);
    }
    c.cellBuf.zero(); c.bndBuf.zero(); c.faceBuf.zero(); c.RBuf.zero(); c.WBuf.zero(); c.timeBuf.zero(); c.fHBuf.zero();
    c.gradBuf.zero();
    c.turbBuf.alloc(2 * n + 2 * nB);
    c.turbBuf.zero();
    c.cell8Buf.alloc(8 * n);
    c.cell8Buf.zero();
    RDST_HIP(hipDeviceSynchronize());
    Fields& f = c.f;
    double* p = c.cellBuf.p;
    double** cells[14] = {&f.rho, &f.rhoUx, &f.rhoUy, &f.rhoUz, &f.rhoE, &f.p, &f.T, &f.h, &f.psi, &f.mu, &f.alpha, &f.Ux, &f.Uy, &f.Uz};
    for (int i = 0; i < 14; ++i) *cells[i] = p + i * n;
    p = c.bndBuf.p;
    double** bnd[21] = {&f.brho, &f.bUx, &f.bUy, &f.bUz, &f.bp, &f.bT, &f.bh, &f.bpsi, &f.bmu, &f.balpha,
                        &f.brhoUx, &f.brhoUy, &f.brhoUz, &f.brhoE, &f.vf_p, &f.vf_T, &f.vf_U, &f.TrefH,
                        &f.refUx, &f.refUy, &f.refUz};
    for (int i = 0; i < 21; ++i) *bnd[i] = p + i * nB;
    f.gradU = c.gradBuf.p;
    f.bgradU = c.gradBuf.p + 9 * n;
    f.gradUc = lim ? c.gradBuf.p + 9 * n + 9 * nB : f.gradU;
    p = c.faceBuf.p;
    double** face[8] = {&f.phi, &f.phiUx, &f.phiUy, &f.phiUz, &f.phiE, &f.Upx, &f.Upy, &f.Upz};
    for (int i = 0; i < 8; ++i) *face[i] = p + i * nF;
    f.R = c.RBuf.p;
    f.W = c.WBuf.p;
    p = c.timeBuf.p;
    f.rDeltaT = p; f.courant = p + n; f.rhoOld = p + 2 * n; f.eOld = p + 3 * n; f.dS = p + 4 * n;
    f.fH = c.fHBuf.p;
    f.cell8 = c.cell8Buf.p;
    f.cell8Kind = cell8_kind(c.st.flux_scheme);
    c.packDirty = true;
    f.mut = c.turbBuf.p; f.alphat = c.turbBuf.p + n;
    f.bmut = c.turbBuf.p + 2 * n; f.balphat = c.turbBuf.p + 2 * n + nB;
    f.rec = f.brec = f.rgrad = nullptr;
    if (c.st.reconstruction == RDST_RECON_OSPRE) {
        c.recBuf.alloc(28 * n + 4 * nB);
        c.recBuf.zero();
        f.rec = c.recBuf.p; f.rgrad = c.recBuf.p + 4 * n; f.brec = c.recBuf.p + 28 * n;
    }
}

static void download_soa(const double* const* comps, int k, size_t n, double* out) {
    std::vector<double> tmp(n);
    for (int d = 0; d < k; ++d) {
        RDST_HIP(hipMemcpy(tmp.data(), comps[d], n * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) out[k * i + d] = tmp[i];
    }
}

}  // namespace rdst

using namespace rdst;

#define RDST_API_BODY(...)                                    \
    try {                                                     \
        __VA_ARGS__;                                          \
        return RDST_OK;                                       \
    } catch (const ArgError& e) {                             \
        set_error(e.msg);                                     \
        return e.code;                                        \
    } catch (const HipError& e) {                             \
        set_error(e.msg);                                     \
        return RDST_ERR_HIP;                                  \
    } catch (const std::exception& e) {                       \
        set_error(e.what());                                  \
        return RDST_ERR_ARG;                                  \
    }

extern "C" {

const char* rdst_last_error(void) { return g_err.c_str(); }
int rdst_version(void) { return 1; }

int rdst_init(int device) {
    RDST_API_BODY({
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw ArgError("no HIP device visible", RDST_ERR_NO_DEVICE);
        if (device < 0 || device >= n) throw ArgError("rdst_init: device index out of range");
        RDST_HIP(hipSetDevice(device));
        forces_scratch_setup(device);
    })
}

int rdst_mesh_create(const rdst_mesh_desc* desc, rdst_mesh** out) {
    RDST_API_BODY({
        if (!out) throw ArgError("rdst_mesh_create: null out");
        Mesh* m = new Mesh();
        try {
            mesh_build(*m, desc);
        } catch (...) {
            delete m;
            throw;
        }
        *out = reinterpret_cast<rdst_mesh*>(m);
    })
}

int rdst_mesh_create_geometry(const rdst_mesh_desc* desc, const rdst_mesh_geometry* geom, rdst_mesh** out) {
    RDST_API_BODY({
        if (!out || !geom) throw ArgError("rdst_mesh_create_geometry: null argument");
        Mesh* m = new Mesh();
        try {
            mesh_build(*m, desc, geom);
        } catch (...) {
            delete m;
            throw;
        }
        *out = reinterpret_cast<rdst_mesh*>(m);
    })
}

int rdst_mesh_destroy(rdst_mesh* mesh) {
    RDST_API_BODY({ delete reinterpret_cast<Mesh*>(mesh); })
}

int rdst_mesh_info(const rdst_mesh* mesh, int64_t* info) {
    RDST_API_BODY({
        const Mesh* m = reinterpret_cast<const Mesh*>(mesh);
        if (!m || !info) throw ArgError("rdst_mesh_info: null argument");
        info[RDST_I_NCELLS] = m->nC;
        info[RDST_I_NINTERNAL] = m->nI;
        info[RDST_I_NBFACES] = m->nB;
        info[RDST_I_NNZB] = m->nnzb;
        info[RDST_I_NCOLOURS] = m->nColours;
        info[RDST_I_NLEVELS_L] = m->nLevL;
        info[RDST_I_NLEVELS_U] = m->nLevU;
        info[RDST_I_NCELL_COLOURS] = m->nCellColours;
        info[RDST_I_NPARTS] = m->nParts;
        info[RDST_I_MAX_PART_LEVELS] = m->maxPartLevels;
    })
}

int rdst_mesh_get_int(const rdst_mesh* mesh, int which, int32_t* out) {
    RDST_API_BODY({
        const Mesh* m = reinterpret_cast<const Mesh*>(mesh);
        if (!m || !out) throw ArgError("rdst_mesh_get_int: null argument");
        const std::vector<int32_t>* v = nullptr;
        std::vector<int32_t> tmp;
        switch (which) {
            case RDST_M_OWNER: v = &m->owner; break;
            case RDST_M_NEIGHBOUR: v = &m->neigh; break;
            case RDST_M_BCSR_ROWPTR: v = &m->rowPtr; break;
            case RDST_M_BCSR_COLS: v = &m->cols; break;
            case RDST_M_SLOT_UPPER: v = &m->slotUpper; break;
            case RDST_M_SLOT_LOWER: v = &m->slotLower; break;
            case RDST_M_SLOT_DIAG: v = &m->slotDiag; break;
            case RDST_M_FACE_COLOUR: v = &m->faceColour; break;
            case RDST_M_ILU_LEVEL: v = &m->levL; break;
            case RDST_M_BFACE_PATCH: v = &m->bPatch; break;
            case RDST_M_CELL_COLOUR: v = &m->cellColour; break;
            case RDST_M_PARTITION: v = &m->part; break;
            case RDST_M_SA_PARTITION: v = &m->saPart; break;
            default: throw ArgError("rdst_mesh_get_int: unknown array");
        }
        std::memcpy(out, v->data(), v->size() * sizeof(int32_t));
    })
}

int rdst_mesh_wall_distance(const rdst_mesh* mesh, double* d) {
    RDST_API_BODY({
        const Mesh* m = reinterpret_cast<const Mesh*>(mesh);
        if (!m || !d) throw ArgError("rdst_mesh_wall_distance: null argument");
        wall_distance(*m, 0, d);
    })
}

int rdst_mesh_get_geom(const rdst_mesh* mesh, int which, double* out) {
    RDST_API_BODY({
        const Mesh* m = reinterpret_cast<const Mesh*>(mesh);
        if (!m || !out) throw ArgError("rdst_mesh_get_geom: null argument");
        const std::vector<double>* v = nullptr;
        switch (which) {
            case RDST_G_SF: v = &m->Sf; break;
            case RDST_G_MAGSF: v = &m->magSf; break;
            case RDST_G_CF: v = &m->Cf; break;
            case RDST_G_C: v = &m->C; break;
            case RDST_G_V: v = &m->V; break;
            case RDST_G_WEIGHTS: v = &m->w; break;
            case RDST_G_DELTACOEFFS: v = &m->dc; break;
            default: throw ArgError("rdst_mesh_get_geom: unknown array");
        }
        std::memcpy(out, v->data(), v->size() * sizeof(double));
    })
}

int rdst_case_create(rdst_mesh* mesh, const rdst_thermo* thermo, const rdst_bc* bcs, const rdst_settings* settings,
                     rdst_case** out) {
    RDST_API_BODY({
        Mesh* m = reinterpret_cast<Mesh*>(mesh);
        if (!m || !thermo || !bcs || !settings || !out) throw ArgError("rdst_case_create: null argument");
        const rdst_settings& s = *settings;
        if (s.flux_scheme < 0 || s.flux_scheme > RDST_FLUX_KURGANOV)
            throw ArgError("Unknown fluxScheme; valid fluxSchemes are: AUSM+ AUSM+up HLL HLLC HLLCP Kurganov",
                           RDST_ERR_UNKNOWN_NAME);
        if (s.solver != RDST_SOLVER_GMRES && s.solver != RDST_SOLVER_BICGSTAB)
            throw ArgError("Unknown solver; valid solvers are: GMRES BiCGStab", RDST_ERR_UNKNOWN_NAME);
        if (s.precond < RDST_PRECOND_ILU0 || s.precond > RDST_PRECOND_DIAG)
            throw ArgError("Unknown preconditioner; valid preconditioners are: ILUC0 none diagonal", RDST_ERR_UNKNOWN_NAME);
        if (s.dependent != RDST_DEP_W && s.dependent != RDST_DEP_DELTAW)
            throw ArgError("Unknown dependentVariable; valid: W deltaW", RDST_ERR_UNKNOWN_NAME);
        if (s.precision != RDST_PREC_FP64 && s.precision != RDST_PREC_MIXED) throw ArgError("Unknown precision");
        if (s.n_directions < 1 || s.max_iter < 0 || s.min_iter < 0) throw ArgError("bad solver controls");
        if (s.gram_schmidt != RDST_GS_MODIFIED && s.gram_schmidt != RDST_GS_CLASSICAL)
            throw ArgError("Unknown Gram-Schmidt variant; valid: modified classical", RDST_ERR_UNKNOWN_NAME);
        if (s.reconstruction != RDST_RECON_UPWIND && s.reconstruction != RDST_RECON_OSPRE)
            throw ArgError("rdst_case_create: unknown reconstruction (valid: upwind OSPRE)", RDST_ERR_UNKNOWN_NAME);
        if (s.jacobian != RDST_JACOBIAN_ROE && s.jacobian != RDST_JACOBIAN_EXACT)
            throw ArgError("rdst_case_create: unknown Jacobian (valid: Roe exact)", RDST_ERR_UNKNOWN_NAME);
        if (s.jacobian == RDST_JACOBIAN_EXACT && (s.reconstruction != RDST_RECON_UPWIND || s.flux_scheme == RDST_FLUX_HLLCP))
            throw ArgError("rdst_case_create: the exact Jacobian covers the first-order residual "
                           "(upwind reconstruction, not HLLCP)");
        if (s.time_integration != RDST_TIME_SINGLE && s.time_integration != RDST_TIME_DUAL)
            throw ArgError("Unknown timeIntegration; valid: Single Dual", RDST_ERR_UNKNOWN_NAME);
        if (s.time_integration == RDST_TIME_DUAL) {
            if (s.jacobian != RDST_JACOBIAN_ROE)
                throw ArgError("rdst_case_create: dual time stepping is provided with the Roe Jacobian");
            // the mixed form solves for the correction from zero, so its
            // normalised initial residual is ~1 on every inner iteration and
            // steadyCheck (which reads it) could never see convergence
            if (s.precision != RDST_PREC_FP64)
                throw ArgError("rdst_case_create: dual time stepping runs in FP64 (the inner convergence test reads the "
                               "W-form initial residual)");
            if (s.max_inner_iter < 1 || !(s.delta_t >= 0) || !(s.min_residual >= 0))
                throw ArgError("rdst_case_create: bad dual-time controls (maxInnerIter >= 1, deltaT >= 0)");
        }
        if (s.ilu_order != RDST_ILU_NATURAL && s.ilu_order != RDST_ILU_PARTITIONED)
            throw ArgError("Unknown ILU ordering; valid: natural partitioned", RDST_ERR_UNKNOWN_NAME);
        for (int q = 0; q < 2; ++q) {
            const int kind = q ? s.sngrad : s.laplacian_sngrad;
            const double psi = q ? s.sngrad_limit : s.laplacian_limit;
            if (kind != RDST_SNGRAD_CORRECTED && kind != RDST_SNGRAD_UNCORRECTED && kind != RDST_SNGRAD_LIMITED)
                throw ArgError("Unknown snGradScheme; valid: corrected uncorrected limited", RDST_ERR_UNKNOWN_NAME);
            if (kind == RDST_SNGRAD_LIMITED && !(psi >= 0.0 && psi <= 1.0))
                throw ArgError("rdst_case_create: limited snGrad coefficient must be in [0, 1]");
        }
        if (s.heat_conduction != 0 && s.heat_conduction != 1)
            throw ArgError("rdst_case_create: heat_conduction is 0 or 1");
        if (!(s.update_limit >= 0.0)) throw ArgError("rdst_case_create: update_limit must be >= 0");
        if (s.grad_u != RDST_GRAD_GAUSS_LINEAR && s.grad_u != RDST_GRAD_CELL_LIMITED)
            throw ArgError("Unknown gradScheme for grad(U); valid: Gauss linear, cellLimited Gauss linear",
                           RDST_ERR_UNKNOWN_NAME);
        if (s.grad_u == RDST_GRAD_CELL_LIMITED && !(s.grad_u_limit >= 0.0 && s.grad_u_limit <= 1.0))
            throw ArgError("rdst_case_create: cellLimited coefficient k must be >= 0 and <= 1");
        if (!(thermo->Cp > thermo->R) || !(thermo->R > 0)) throw ArgError("rdst_case_create: need Cp > R > 0");
        if (!m->uploaded) {
            mesh_upload(*m);
            m->uploaded = true;
        }
        Case* c = new Case();
        try {
            c->mesh = m;
            c->st = s;
            c->th.Cp = thermo->Cp;
            c->th.R = thermo->R;
            c->th.Cv = thermo->Cp - thermo->R;
            c->th.gamma = thermo->Cp / c->th.Cv;
            c->th.transport = thermo->transport;
            c->th.mu = thermo->mu;
            c->th.Pr = thermo->Pr;
            c->th.As = thermo->As;
            c->th.Ts = thermo->Ts;
            c->bcs.assign(bcs, bcs + m->nPatches);
            std::vector<BcPatchDev> hb(std::max(m->nPatches, 1));
            for (int p = 0; p < m->nPatches; ++p) {
                const rdst_bc& b = bcs[p];
                for (int t : {b.p_type, b.T_type, b.U_type})
                    if (t < RDST_BC_FIXED_VALUE || t > RDST_BC_EMPTY) throw ArgError("unknown boundary condition type");
                if (m->patches[p].type != RDST_PATCH_EMPTY && (b.p_type == RDST_BC_EMPTY || b.T_type == RDST_BC_EMPTY || b.U_type == RDST_BC_EMPTY))
                    throw ArgError(std::string("empty BC on non-empty patch ") + m->patches[p].name);
                BcPatchDev& d = hb[p];
                d.p_type = b.p_type; d.T_type = b.T_type; d.U_type = b.U_type;
                d.is_wall = m->patches[p].type == RDST_PATCH_WALL;
                d.p_value = b.p_value; d.T_value = b.T_value;
                d.p_inf = b.p_inf; d.T_inf = b.T_inf;
                for (int k = 0; k < 3; ++k) { d.U_value[k] = b.U_value[k]; d.U_inf[k] = b.U_inf[k]; }
            }
            c->d_bc.upload(hb);
            RDST_HIP(hipStreamCreate(&c->stream));
            RDST_HIP(hipEventCreate(&c->ev0));
            RDST_HIP(hipEventCreate(&c->ev1));
            alloc_fields(*c);
        } catch (...) {
            delete c;
            throw;
        }
        *out = reinterpret_cast<rdst_case*>(c);
    })
}

int rdst_case_destroy(rdst_case* cc) {
    RDST_API_BODY({
        Case* c = reinterpret_cast<Case*>(cc);
        if (!c) return RDST_OK;
        solver_free(*c);
        if (c->stream) (void)hipStreamSynchronize(c->stream);
        if (c->ev0) (void)hipEventDestroy(c->ev0);
        if (c->ev1) (void)hipEventDestroy(c->ev1);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
    })
}

int rdst_case_set_state(rdst_case* cc, const double* p, const double* T, const double* U) {
    RDST_API_BODY({
        Case* c = reinterpret_cast<Case*>(cc);
        if (!c || !p || !T || !U) throw ArgError("rdst_case_set_state: null argument");
        for (int i = 0; i < c->mesh->nC; ++i)
            if (!(p[i] > 0) || !(T[i] > 0)) throw ArgError("rdst_case_set_state: p and T must be positive");
        set_state(*c, p, T, U);
    })
}

int rdst_case_get(rdst_case* cc, int field, double* out) {
    RDST_API_BODY({
        Case* c = reinterpret_cast<Case*>(cc);
        if (!c || !out) throw ArgError("rdst_case_get: null argument");
        RDST_HIP(hipStreamSynchronize(c->stream));
        const Mesh& m = *c->mesh;
        const size_t n = m.nC, nB = m.nB, nF = m.nF;
        const Fields& f = c->f;
        auto one = [&](const double* d, size_t cnt) { RDST_HIP(hipMemcpy(out, d, cnt * sizeof(double), hipMemcpyDeviceToHost)); };
        switch (field) {
            case RDST_F_RHO: one(f.rho, n); break;
            case RDST_F_RHOU: { const double* v[3] = {f.rhoUx, f.rhoUy, f.rhoUz}; download_soa(v, 3, n, out); break; }
            case RDST_F_RHOE: one(f.rhoE, n); break;
            case RDST_F_P: one(f.p, n); break;
            case RDST_F_T: one(f.T, n); break;
            case RDST_F_U: { const double* v[3] = {f.Ux, f.Uy, f.Uz}; download_soa(v, 3, n, out); break; }
            case RDST_F_H: one(f.h, n); break;
            case RDST_F_PSI: one(f.psi, n); break;
            case RDST_F_R: { const double* v[5] = {f.R, f.R + n, f.R + 2 * n, f.R + 3 * n, f.R + 4 * n}; download_soa(v, 5, n, out); break; }
            case RDST_F_W: one(f.W, 5 * n); break;
            case RDST_F_RDELTAT: one(f.rDeltaT, n); break;
            case RDST_F_COURANT: one(f.courant, n); break;
            case RDST_F_PHI: one(f.phi, nF); break;
            case RDST_F_PHIUP: { const double* v[3] = {f.phiUx, f.phiUy, f.phiUz}; download_soa(v, 3, nF, out); break; }
            case RDST_F_PHIEP: one(f.phiE, nF); break;
            case RDST_F_UP: { const double* v[3] = {f.Upx, f.Upy, f.Upz}; download_soa(v, 3, nF, out); break; }
            case RDST_F_BRHO: one(f.brho, nB); break;
            case RDST_F_BU: { const double* v[3] = {f.bUx, f.bUy, f.bUz}; download_soa(v, 3, nB, out); break; }
            case RDST_F_BP: one(f.bp, nB); break;
            case RDST_F_BT: one(f.bT, nB); break;
            case RDST_F_BRHOU: { const double* v[3] = {f.brhoUx, f.brhoUy, f.brhoUz}; download_soa(v, 3, nB, out); break; }
            case RDST_F_BRHOE: one(f.brhoE, nB); break;
            case RDST_F_BH: one(f.bh, nB); break;
            case RDST_F_BPSI: one(f.bpsi, nB); break;
            case RDST_F_BCSR_VALUES:
                if (!c->hasMatrix || !c->A64.p) throw ArgError("no FP64 matrix assembled", RDST_ERR_STATE);
                newton_operator_csr(*c);
                one(c->A64.p, 25 * (size_t)m.nnzb);
                break;
            case RDST_F_SOURCE: one(c->src.p, 5 * n); break;
            case RDST_F_MU: one(f.mu, n); break;
            case RDST_F_RHO_OLD: one(f.rhoOld, n); break;
            case RDST_F_E_OLD: one(f.eOld, n); break;
            case RDST_F_MUT: one(f.mut, n); break;
            case RDST_F_ALPHAT: one(f.alphat, n); break;
            case RDST_F_BMUT: one(f.bmut, nB); break;
            case RDST_F_GRAD_U: {
                if (!c->st.viscous) throw ArgError("rdst_case_get: grad(U) is formed by viscous cases only", RDST_ERR_STATE);
                const double* v[9];
                for (int q = 0; q < 9; ++q) v[q] = f.gradU + (size_t)q * n;
                download_soa(v, 9, n, out);
                break;
            }
            case RDST_F_NU_TILDA: case RDST_F_WALL_DIST: case RDST_F_BNU_TILDA: case RDST_F_SA_DIAG:
            case RDST_F_SA_UPPER: case RDST_F_SA_LOWER: case RDST_F_SA_SOURCE: {
                const SaState& s = c->sa;
                if (!s.on) throw ArgError("rdst_case_get: no turbulence model on the case", RDST_ERR_STATE);
                if (field == RDST_F_NU_TILDA) one(s.nut, n);
                else if (field == RDST_F_WALL_DIST) one(s.d, n);
                else if (field == RDST_F_BNU_TILDA) one(s.bnut, nB);
                else if (field == RDST_F_SA_SOURCE) one(s.B, n);
                else {
                    std::vector<double> A(m.nnzb);
                    RDST_HIP(hipMemcpy(A.data(), s.A.p, A.size() * sizeof(double), hipMemcpyDeviceToHost));
                    if (field == RDST_F_SA_DIAG)
                        for (size_t i = 0; i < n; ++i) out[i] = A[m.slotDiag[i]];
                    else if (field == RDST_F_SA_UPPER)
                        for (int k = 0; k < m.nI; ++k) out[k] = A[m.slotUpper[k]];
                    else
                        for (int k = 0; k < m.nI; ++k) out[k] = A[m.slotLower[k]];
                }
                break;
            }
            default: throw ArgError("rdst_case_get: unknown field");
        }
    })
}

int rdst_case_put(rdst_case* cc, int field, const double* in) {
    RDST_API_BODY({
        Case* c = reinterpret_cast<Case*>(cc);
        if (!c || !in) throw ArgError("rdst_case_put: null argument");
        RDST_HIP(hipStreamSynchronize(c->stream));
        const size_t n = c->mesh->nC;
        Fields& f = c->f;
        c->packDirty = true;
        auto one = [&](double* d, size_t cnt) { RDST_HIP(hipMemcpy(d, in, cnt * sizeof(double), hipMemcpyHostToDevice)); };
        switch (field) {
            case RDST_F_RHO: one(f.rho, n); break;
            case RDST_F_RHOE: one(f.rhoE, n); break;
            case RDST_F_W: one(f.W, 5 * n); break;
            case RDST_F_RDELTAT: one(f.rDeltaT, n); break;
            case RDST_F_NU_TILDA: sa_put_nut(*c, in); break;
            // the stored mass flux (read by the boundary conditions' masks) and
            // the eddy viscosity and conductivity, for a state restored or
            // carried to another mesh with the turbulence frozen
            case RDST_F_PHI: one(f.phi, (size_t)c->mesh->nF); break;
            case RDST_F_MUT: one(f.mut, n); break;
            case RDST_F_ALPHAT: one(f.alphat, n); break;
            case RDST_F_BMUT: {
                const size_t nB = c->mesh->nB;
                one(f.bmut, nB);
                const double prt = c->sa.on ? c->sa.t.Prt : 1.0;  // alphat_b = mut_b / Prt as the model sets it
                std::vector<double> t(in, in + nB);
                for (double& v : t) v /= prt;
                if (c->sa.on) RDST_HIP(hipMemcpy(f.balphat, t.data(), nB * sizeof(double), hipMemcpyHostToDevice));
                break;
            }
            case RDST_F_RHOU: {
                std::vector<double> t(n);
                double* d[3] = {f.rhoUx, f.rhoUy, f.rhoUz};
                for (int k = 0; k < 3; ++k) {
                    for (size_t i = 0; i < n; ++i) t[i] = in[3 * i + k];
                    RDST_HIP(hipMemcpy(d[k], t.data(), n * sizeof(double), hipMemcpyHostToDevice));
                }
                break;
            }
            default: throw ArgError("rdst_case_put: field not writable");
        }
    })
}

int rdst_flux_residual(rdst_case* cc) {
    RDST_API_BODY({
        Case* c = reinterpret_cast<Case*>(cc);
        if (!c) throw ArgError("null case");
        flux_residual(*c, STORE_ALL);
    })
}

int rdst_jacobian_assemble(rdst_case* cc) {
    RDST_API_BODY({
        Case* c = reinterpret_cast<Case*>(cc);
        if (!c) throw ArgError("null case");
        jacobian_assemble(*c);
        RDST_HIP(hipStreamSynchronize(c->stream));
    })
}

int rdst_block_solve(rdst_case* cc, rdst_solver_perf* perf) {
    RDST_API_BODY({
        Case* c = reinterpret_cast<Case*>(cc);
        if (!c) throw ArgError("null case");
        block_solve(*c, perf);
        RDST_HIP(hipStreamSynchronize(c->stream));
    })
}

int rdst_ldu_solve(rdst_case* cc, const double* diag, const double* upper, const double* lower, const double* b,
                   double* x, int mode, rdst_solver_perf* perf) {
    RDST_API_BODY({
        Case* c = reinterpret_cast<Case*>(cc);
        if (!c || !diag || !b || !x || (c->mesh->nI && (!upper || !lower)))
            throw ArgError("rdst_ldu_solve: null argument");
        if (mode != RDST_LDU_PLAIN && mode != RDST_LDU_W && mode != RDST_LDU_DELTAW)
            throw ArgError("rdst_ldu_solve: unknown mode; valid: PLAIN W deltaW", RDST_ERR_UNKNOWN_NAME);
        ldu_solve(*c, diag, upper, lower, b, x, mode, perf);
    })
}

int rdst_update_fields(rdst_case* cc) {
    RDST_API_BODY({
        Case* c = reinterpret_cast<Case*>(cc);
        if (!c) throw ArgError("null case");
        if (!c->hasState) throw ArgError("rdst_update_fields: no state", RDST_ERR_STATE);
        update_fields(*c);
        RDST_HIP(hipStreamSynchronize(c->stream));
    })
}

int rdst_set_courant(rdst_case* cc, double* res_norm) {
    RDST_API_BODY({
        Case* c = reinterpret_cast<Case*>(cc);
        if (!c) throw ArgError("null case");
        if (!c->hasState) throw ArgError("rdst_set_courant: no state", RDST_ERR_STATE);
        double r = set_courant(*c);
        RDST_HIP(hipStreamSynchronize(c->stream));
        if (res_norm) *res_norm = r;
    })
}

int rdst_iterate(rdst_case* cc, int n_steps, double* res_norm, rdst_solver_perf* perf) {
    RDST_API_BODY({
        Case* c = reinterpret_cast<Case*>(cc);
        if (!c) throw ArgError("null case");
        if (!c->hasState) throw ArgError("rdst_iterate: no state", RDST_ERR_STATE);
        if (c->st.time_integration == RDST_TIME_DUAL)
            throw ArgError("rdst_iterate: a Dual case advances with rdst_dual_step", RDST_ERR_STATE);
        for (int s = 0; s < n_steps; ++s) {
            flux_launch(*c, c->sa.on ? STORE_PHI : STORE_BPHI);
            jacobian_assemble(*c);
            rdst_solver_perf p{};
            block_solve(*c, &p);
            update_fields(*c);
            if (c->sa.on) sa_correct(*c, nullptr);
            double r = set_courant(*c);
            if (res_norm) res_norm[s] = r;
            if (perf) perf[s] = p;
        }
        RDST_HIP(hipStreamSynchronize(c->stream));
    })
}

namespace {
// residualControls + steadyCheck (residualControls.C:37-64, steadyCheck.C:37-98):
// the inner residual is the mean of the five initial residuals of the inner
// solve; converged after more than maxInnerIter inner iterations, below
// minResudual, or (from iterationStart on) when the scaled slope of
// log(residual) over the last 20 % of the history drops below
// magConvergedSlope.  converged() refreshes the slope, converged_const()
// (what setCourantDST asks) reads the last one.
struct SteadyCheck {
    std::vector<double> residuals;
    int iterationCount = 0, iterationStart, maxInnerIter;
    double residual = 1, magScaledSlope = 1, magConvergedSlope, minResidual;
    explicit SteadyCheck(const rdst_settings& st)
        : iterationStart(st.iteration_start), maxInnerIter(st.max_inner_iter),
          magConvergedSlope(st.mag_converged_slope), minResidual(st.min_residual) {}
    void add_residual(const double* res) {
        double mean = 0;
        for (int i = 0; i < 5; ++i) mean += res[i];
        mean /= 5;
        ++iterationCount;
        residual = mean;
        residuals.push_back(mean);
    }
    bool converged_const() const {
        return iterationCount > maxInnerIter || residual < minResidual || magScaledSlope < magConvergedSlope;
    }
    bool converged() {
        if (iterationCount >= iterationStart) {
            const int size = (int)residuals.size();
            const int N = 0.2 * size;
            const double maxY = std::log(*std::max_element(residuals.begin(), residuals.end()));
            const double minY = std::log(*std::min_element(residuals.begin(), residuals.end()));
            const double deltaY = maxY - minY;
            double sumXi = 0, sumYi = 0, sumXiSqr = 0, sumYiXi = 0;
            for (int i = 0; i < N; ++i) {
                const int xi = size - 1 - i;
                const double yi = std::log(residuals[xi]);
                sumXi += xi;
                sumXiSqr += std::pow(xi, 2);
                sumYi += yi;
                sumYiXi += yi * xi;
            }
            magScaledSlope = (size - 1) / deltaY *
                             std::fabs((sumYiXi - sumXi * sumYi / N) / (sumXiSqr - std::pow(sumXi, 2) / N));
        }
        return converged_const();
    }
};
}  // namespace

int rdst_dual_step(rdst_case* cc, int32_t* n_inner, double* inner_residuals, double* res_norm) {
    RDST_API_BODY({
        Case* c = reinterpret_cast<Case*>(cc);
        if (!c) throw ArgError("null case");
        if (!c->hasState) throw ArgError("rdst_dual_step: no state", RDST_ERR_STATE);
        if (c->st.time_integration != RDST_TIME_DUAL)
            throw ArgError("rdst_dual_step: the case is not Dual (settings.time_integration)", RDST_ERR_STATE);
        const size_t n5 = 5 * (size_t)c->mesh->nC;
        // An error inside the step leaves the inner iterations' W behind: the
        // time control goes back to the start of the step and the case needs
        // rdst_case_set_state again, so a retry cannot advance time twice or
        // run on a shifted BDF2 history.
        const double time0 = c->time, deltaT00 = c->deltaT0, deltaTSave0 = c->deltaTSave;
        const int timeIndex0 = c->timeIndex;
        struct Rollback {
            Case* c;
            double time, deltaT0, deltaTSave;
            int timeIndex;
            bool armed = true;
            ~Rollback() {
                if (!armed) return;
                c->time = time; c->deltaT0 = deltaT0; c->deltaTSave = deltaTSave; c->timeIndex = timeIndex;
                c->hasState = false;
            }
        } rollback{c, time0, deltaT00, deltaTSave0, timeIndex0};
        // runTime++: time, deltaT0 = the previous deltaT, old time levels
        c->deltaT0 = c->deltaTSave;
        c->deltaTSave = c->deltaT;
        c->time += c->deltaT;
        ++c->timeIndex;
        // oldTime().oldTime() of the first step is oldTime() itself
        if (c->timeIndex == 1) store_conserved(*c, c->WOldOld.p);
        else RDST_HIP(hipMemcpyAsync(c->WOldOld.p, c->WOld.p, sizeof(double) * n5, hipMemcpyDeviceToDevice, c->stream));
        store_conserved(*c, c->WOld.p);
        // Dual::iterate
        SteadyCheck sc(c->st);
        int k = 0;
        do {
            flux_launch(*c, STORE_BPHI);
            jacobian_assemble(*c);
            rdst_solver_perf p{};
            block_solve(*c, &p);
            update_fields(*c);
            sc.add_residual(p.initial_residual);
            if (inner_residuals) inner_residuals[k] = sc.residual;
            ++k;
            set_courant_dst(*c, p.initial_residual, !sc.converged_const());
        } while (!sc.converged());
        const double r = set_courant_dst(*c, nullptr, !sc.converged_const());
        if (n_inner) *n_inner = k;
        if (res_norm) *res_norm = r;
        RDST_HIP(hipStreamSynchronize(c->stream));
        rollback.armed = false;
    })
}

int rdst_time_info(const rdst_case* cc, double* out) {
    RDST_API_BODY({
        const Case* c = reinterpret_cast<const Case*>(cc);
        if (!c || !out) throw ArgError("rdst_time_info: null argument");
        out[0] = c->time;
        out[1] = c->deltaT;
        out[2] = c->deltaT0;
        out[3] = c->courantScalar;
    })
}

int rdst_forces(rdst_case* cc, int patch, const double* drag_dir, const double* lift_dir, double p_ref, double rho_inf,
                double mag_u_inf, double a_ref, double* out) {
    RDST_API_BODY({
        Case* c = reinterpret_cast<Case*>(cc);
        if (!c || !drag_dir || !lift_dir || !out) throw ArgError("rdst_forces: null argument");
        const Mesh& m = *c->mesh;
        if (patch < 0 || patch >= m.nPatches) throw ArgError("rdst_forces: patch index out of range");
        forces(*c, patch, drag_dir, lift_dir, p_ref, rho_inf, mag_u_inf, a_ref, out, nullptr);
    })
}

int rdst_forces_noalloc(rdst_case* cc, int patch, const double* drag_dir, const double* lift_dir, double p_ref,
                        double rho_inf, double mag_u_inf, double a_ref, double* out) {
    RDST_API_BODY({
        Case* c = reinterpret_cast<Case*>(cc);
        if (!c || !drag_dir || !lift_dir || !out) throw ArgError("rdst_forces_noalloc: null argument");
        const Mesh& m = *c->mesh;
        if (patch < 0 || patch >= m.nPatches) throw ArgError("rdst_forces_noalloc: patch index out of range");
        int dev = -1;
        RDST_HIP(hipGetDevice(&dev));
        forces(*c, patch, drag_dir, lift_dir, p_ref, rho_inf, mag_u_inf, a_ref, out, forces_scratch(dev));
    })
}

int rdst_case_set_turbulence(rdst_case* cc, const rdst_turbulence* turb, const rdst_scalar_bc* nu_tilda_bcs,
                             const double* wall_dist, const double* nu_tilda) {
    RDST_API_BODY({
        Case* c = reinterpret_cast<Case*>(cc);
        if (!c || !turb) throw ArgError("rdst_case_set_turbulence: null argument");
        if (turb->model != RDST_TURB_LAMINAR && turb->model != RDST_TURB_SANEG && turb->model != RDST_TURB_SARC2 &&
            turb->model != RDST_TURB_SARC1 && turb->model != RDST_TURB_SA)
            throw ArgError("rdst_case_set_turbulence: unknown RASModel id (valid: laminar SAneg SARC2 SARC1 "
                           "SpalartAllmaras)", RDST_ERR_UNKNOWN_NAME);
        if (turb->model != RDST_TURB_LAMINAR) {
            if (c->st.time_integration == RDST_TIME_DUAL)
                throw ArgError("rdst_case_set_turbulence: the Spalart-Allmaras models run with Single time "
                               "integration (the dual2ndDST nuTilda ddt is not built)");
            if (!nu_tilda_bcs || !nu_tilda) throw ArgError("rdst_case_set_turbulence: null array");
            if (!c->hasState) throw ArgError("rdst_case_set_turbulence: no state (call rdst_case_set_state)", RDST_ERR_STATE);
            if (!c->st.viscous) throw ArgError("rdst_case_set_turbulence: the viscous terms are off (settings.viscous)");
            if (!(turb->sigma_nut > 0) || !(turb->kappa > 0) || !(turb->Prt > 0) || !(turb->Cv2 > 0))
                throw ArgError("rdst_case_set_turbulence: sigmaNut, kappa, Prt and Cv2 must be > 0");
            if (turb->relax > 1.0) throw ArgError("rdst_case_set_turbulence: relaxation factor > 1");
            if (turb->max_iter < 1) throw ArgError("rdst_case_set_turbulence: max_iter < 1");
            if (turb->convection != 0 && turb->convection != 1)
                throw ArgError("rdst_case_set_turbulence: convection is 0 (upwind) or 1 (linearUpwind)",
                               RDST_ERR_UNKNOWN_NAME);
        }
        std::vector<double> wd;
        if (turb->model != RDST_TURB_LAMINAR && !wall_dist) {
            wd.resize(c->mesh->nC);
            wall_distance(*c->mesh, c->stream, wd.data());
            wall_dist = wd.data();
        }
        sa_set(*c, turb, nu_tilda_bcs, wall_dist, nu_tilda);
    })
}

int rdst_turbulence_correct(rdst_case* cc, rdst_solver_perf* perf) {
    RDST_API_BODY({
        Case* c = reinterpret_cast<Case*>(cc);
        if (!c) throw ArgError("null case");
        sa_correct(*c, perf);
        RDST_HIP(hipStreamSynchronize(c->stream));
    })
}

static Case* adjoint_case(rdst_case* cc) {
    Case* c = reinterpret_cast<Case*>(cc);
    if (!c) throw ArgError("null case");
    if (!c->hasState) throw ArgError("adjoint: no state", RDST_ERR_STATE);
    adjoint_check_scheme(*c);
    return c;
}

int rdst_adjoint_linearise(rdst_case* cc) {
    RDST_API_BODY({
        Case* c = adjoint_case(cc);
        adjoint_linearise(*c);
        RDST_HIP(hipStreamSynchronize(c->stream));
    })
}

int rdst_adjoint_apply(rdst_case* cc, int transpose, const double* x, double* y) {
    RDST_API_BODY({
        Case* c = adjoint_case(cc);
        if (!x || !y) throw ArgError("rdst_adjoint_apply: null array");
        const size_t n5 = 5 * (size_t)c->mesh->nC;
        DBuf<double> dx, dy;
        dx.upload(x, n5);
        dy.alloc(n5);
        adjoint_apply(*c, transpose != 0, dx.p, dy.p);
        RDST_HIP(hipStreamSynchronize(c->stream));
        RDST_HIP(hipMemcpy(y, dy.p, n5 * sizeof(double), hipMemcpyDeviceToHost));
    })
}

int rdst_adjoint_residual(rdst_case* cc, double* R) {
    RDST_API_BODY({
        Case* c = adjoint_case(cc);
        if (!R) throw ArgError("rdst_adjoint_residual: null array");
        const size_t n5 = 5 * (size_t)c->mesh->nC;
        DBuf<double> d;
        d.alloc(n5);
        adjoint_residual(*c, d.p);
        RDST_HIP(hipStreamSynchronize(c->stream));
        RDST_HIP(hipMemcpy(R, d.p, n5 * sizeof(double), hipMemcpyDeviceToHost));
    })
}

static void check_objective(const Case& c, int patch, const double* dir, double q) {
    if (!dir) throw ArgError("adjoint: null direction");
    if (patch < 0 || patch >= c.mesh->nPatches) throw ArgError("adjoint: patch index out of range");
    if (!(q != 0.0) || !std::isfinite(q)) throw ArgError("adjoint: the reference dynamic pressure x area q must be finite and nonzero");
}

int rdst_adjoint_objective_gradient(rdst_case* cc, int patch, const double* dir, double q, double* g) {
    RDST_API_BODY({
        Case* c = adjoint_case(cc);
        check_objective(*c, patch, dir, q);
        if (!g) throw ArgError("rdst_adjoint_objective_gradient: null array");
        const size_t n5 = 5 * (size_t)c->mesh->nC;
        DBuf<double> d;
        d.alloc(n5);
        adjoint_objective_gradient(*c, patch, dir, q, d.p);
        RDST_HIP(hipStreamSynchronize(c->stream));
        RDST_HIP(hipMemcpy(g, d.p, n5 * sizeof(double), hipMemcpyDeviceToHost));
    })
}

int rdst_adjoint_objective_cells(rdst_case* cc, int patch, const double* dir, double q, double p_ref, double* out) {
    RDST_API_BODY({
        Case* c = adjoint_case(cc);
        check_objective(*c, patch, dir, q);
        if (!out) throw ArgError("rdst_adjoint_objective_cells: null array");
        const size_t n = (size_t)c->mesh->nC;
        DBuf<double> d;
        d.alloc(n);
        adjoint_objective_cells(*c, patch, dir, q, p_ref, d.p);
        RDST_HIP(hipStreamSynchronize(c->stream));
        RDST_HIP(hipMemcpy(out, d.p, n * sizeof(double), hipMemcpyDeviceToHost));
    })
}

int rdst_adjoint_solve(rdst_case* cc, int patch, const double* dir, double q, const rdst_adjoint_settings* s,
                       double* lambda, rdst_solver_perf* perf) {
    RDST_API_BODY({
        Case* c = adjoint_case(cc);
        check_objective(*c, patch, dir, q);
        if (!s || !lambda) throw ArgError("rdst_adjoint_solve: null argument");
        if (s->precision != RDST_PREC_FP64 && s->precision != RDST_PREC_MIXED)
            throw ArgError("rdst_adjoint_solve: unknown precision", RDST_ERR_UNKNOWN_NAME);
        if (s->n_directions < 1 || s->max_cycles < 1) throw ArgError("rdst_adjoint_solve: n_directions, max_cycles >= 1");
        if (s->inner_operator != 0 && s->inner_operator != 1)
            throw ArgError("rdst_adjoint_solve: inner_operator is 0 or 1", RDST_ERR_UNKNOWN_NAME);
        if (!(s->precond_shift >= 0.0) || !std::isfinite(s->precond_shift))
            throw ArgError("rdst_adjoint_solve: precond_shift must be finite and >= 0");
        adjoint_linearise(*c);  // at the current state
        const size_t n5 = 5 * (size_t)c->mesh->nC;
        DBuf<double> g, x;
        g.alloc(n5);
        x.upload(lambda, n5);
        adjoint_objective_gradient(*c, patch, dir, q, g.p);
        adjoint_solve(*c, g.p, x.p, *s, perf);
        RDST_HIP(hipMemcpy(lambda, x.p, n5 * sizeof(double), hipMemcpyDeviceToHost));
    })
}

int rdst_last_flux_ms(rdst_case* cc, double* ms) {
    RDST_API_BODY({
        Case* c = reinterpret_cast<Case*>(cc);
        if (!c || !ms) throw ArgError("null argument");
        *ms = c->lastFluxMs;
    })
}

int rdst_time_flux(rdst_case* cc, int n, double* ms_per_call) {
    RDST_API_BODY({
        Case* c = reinterpret_cast<Case*>(cc);
        if (!c || !ms_per_call || n < 1) throw ArgError("rdst_time_flux: bad argument");
        *ms_per_call = flux_time(*c, n);
    })
}

}  // extern "C"


