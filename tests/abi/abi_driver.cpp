// abi_driver.cpp -- a C++ caller of include/rhodst.h, compiled with g++ and
// linked against librhodst_amd.so, as a foam-extend adapter would be.
//
// Reads a polyMesh and a case (the C structs of the header, written as raw
// bytes by the test that runs this program) from <dir>, then through the
// C-ABI only: creates the mesh (from the points, and again from the first
// mesh's geometry read back), creates the case, takes one linear solve of
// the assembled system both ways (rdst_block_solve, and rdst_ldu_solve on
// the same coefficients rearranged into foam's LDU order), runs n steps of
// Single::iterate and integrates the force on one patch with rdst_forces and
// rdst_forces_noalloc.  Prints one JSON line; exit status 0 unless a call
// failed.
//
// usage: abi_driver <dir> <n_steps> <force_patch>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "rhodst.h"

namespace {

FILE* open_or_die(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        std::fprintf(stderr, "cannot open %s\n", path.c_str());
        std::exit(2);
    }
    return f;
}

template <class T>
void read_into(FILE* f, T* dst, size_t n) {
    if (n && std::fread(dst, sizeof(T), n, f) != n) {
        std::fprintf(stderr, "short read\n");
        std::exit(2);
    }
}

void check(int rc, const char* what) {
    if (rc != RDST_OK) {
        std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, rdst_last_error());
        std::exit(1);
    }
}

void print_array(const char* key, const double* a, int n, bool last = false) {
    std::printf("\"%s\": [", key);
    for (int i = 0; i < n; ++i) std::printf("%s%.17g", i ? ", " : "", a[i]);
    std::printf("]%s", last ? "" : ", ");
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s <dir> <n_steps> <force_patch>\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    const int nSteps = std::atoi(argv[2]);
    const int forcePatch = std::atoi(argv[3]);

    // ---- polyMesh primitives
    FILE* fm = open_or_die(dir + "/mesh.bin");
    int32_t hdr[5];
    read_into(fm, hdr, 5);
    const int nPoints = hdr[0], nFaces = hdr[1], nInt = hdr[2], nCells = hdr[3], nPatches = hdr[4];
    std::vector<double> points(3 * (size_t)nPoints);
    std::vector<int32_t> faceOffsets(nFaces + 1), owner(nFaces), neighbour(nInt);
    std::vector<rdst_patch> patches(nPatches);
    read_into(fm, points.data(), points.size());
    read_into(fm, faceOffsets.data(), faceOffsets.size());
    std::vector<int32_t> facePoints(faceOffsets[nFaces]);
    read_into(fm, facePoints.data(), facePoints.size());
    read_into(fm, owner.data(), owner.size());
    read_into(fm, neighbour.data(), neighbour.size());
    read_into(fm, patches.data(), patches.size());
    std::fclose(fm);

    // ---- thermo, per-patch BCs, settings and the initial state
    FILE* fc = open_or_die(dir + "/case.bin");
    rdst_thermo thermo;
    rdst_settings settings;
    std::vector<rdst_bc> bcs(nPatches);
    read_into(fc, &thermo, 1);
    read_into(fc, bcs.data(), bcs.size());
    read_into(fc, &settings, 1);
    std::vector<double> p(nCells), T(nCells), U(3 * (size_t)nCells);
    read_into(fc, p.data(), p.size());
    read_into(fc, T.data(), T.size());
    read_into(fc, U.data(), U.size());
    std::fclose(fc);

    check(rdst_init(0), "rdst_init");
    rdst_mesh_desc desc{nPoints, nFaces, nInt, nCells, nPatches, 0, points.data(), faceOffsets.data(),
                        facePoints.data(), owner.data(), neighbour.data(), patches.data()};
    rdst_mesh* mesh0 = nullptr;
    check(rdst_mesh_create(&desc, &mesh0), "rdst_mesh_create");

    // the host's own geometry: read back the computed one, in polyMesh face
    // order, and create the mesh the case runs on from it
    int64_t info[RDST_I_COUNT];
    check(rdst_mesh_info(mesh0, info), "rdst_mesh_info");
    const int nB = (int)info[RDST_I_NBFACES], nF = nInt + nB;
    const long nnzb = (long)info[RDST_I_NNZB];
    std::vector<double> Sf(3 * (size_t)nF), magSf(nF), Cf(3 * (size_t)nF), C(3 * (size_t)nCells), V(nCells), w(nF),
        dc(nF);
    check(rdst_mesh_get_geom(mesh0, RDST_G_SF, Sf.data()), "geom Sf");
    check(rdst_mesh_get_geom(mesh0, RDST_G_MAGSF, magSf.data()), "geom magSf");
    check(rdst_mesh_get_geom(mesh0, RDST_G_CF, Cf.data()), "geom Cf");
    check(rdst_mesh_get_geom(mesh0, RDST_G_C, C.data()), "geom C");
    check(rdst_mesh_get_geom(mesh0, RDST_G_V, V.data()), "geom V");
    check(rdst_mesh_get_geom(mesh0, RDST_G_WEIGHTS, w.data()), "geom weights");
    check(rdst_mesh_get_geom(mesh0, RDST_G_DELTACOEFFS, dc.data()), "geom deltaCoeffs");
    std::vector<double> pSf(3 * (size_t)nFaces, 0.0), pMag(nFaces, 1.0), pCf(3 * (size_t)nFaces, 0.0), pW(nFaces, 1.0),
        pDc(nFaces, 1.0);
    // surface fields hold the internal faces, then the faces of the non-empty
    // patches in patch order
    std::vector<int> polyFace(nF);
    for (int k = 0; k < nInt; ++k) polyFace[k] = k;
    for (int pi = 0, k = nInt; pi < nPatches; ++pi)
        if (patches[pi].type != RDST_PATCH_EMPTY)
            for (int i = 0; i < patches[pi].size; ++i) polyFace[k++] = patches[pi].start + i;
    for (int k = 0; k < nF; ++k) {
        const int f = polyFace[k];
        for (int x = 0; x < 3; ++x) {
            pSf[3 * (size_t)f + x] = Sf[3 * (size_t)k + x];
            pCf[3 * (size_t)f + x] = Cf[3 * (size_t)k + x];
        }
        pMag[f] = magSf[k];
        pW[f] = w[k];
        pDc[f] = dc[k];
    }
    rdst_mesh_geometry geom{pSf.data(), pMag.data(), pCf.data(), C.data(), V.data(), pW.data(), pDc.data()};
    rdst_mesh* mesh = nullptr;
    check(rdst_mesh_create_geometry(&desc, &geom, &mesh), "rdst_mesh_create_geometry");
    std::vector<double> V2(nCells);
    check(rdst_mesh_get_geom(mesh, RDST_G_V, V2.data()), "geom V (2)");
    int geomSame = 1;
    for (int c = 0; c < nCells; ++c) geomSame &= (V2[c] == V[c]);

    // an unknown fluxScheme is refused with the list of valid names
    rdst_settings bad = settings;
    bad.flux_scheme = 99;
    rdst_case* none = nullptr;
    const int unknownRc = rdst_case_create(mesh, &thermo, bcs.data(), &bad, &none);

    rdst_case* cs = nullptr;
    check(rdst_case_create(mesh, &thermo, bcs.data(), &settings, &cs), "rdst_case_create");
    check(rdst_case_set_state(cs, p.data(), T.data(), U.data()), "rdst_case_set_state");

    // one linear solve both ways on the same assembled system
    check(rdst_flux_residual(cs), "rdst_flux_residual");
    check(rdst_jacobian_assemble(cs), "rdst_jacobian_assemble");
    std::vector<double> bcsr(25 * (size_t)nnzb), src(5 * (size_t)nCells), W0(5 * (size_t)nCells);
    check(rdst_case_get(cs, RDST_F_BCSR_VALUES, bcsr.data()), "get BCSR");
    check(rdst_case_get(cs, RDST_F_SOURCE, src.data()), "get source");
    check(rdst_case_get(cs, RDST_F_W, W0.data()), "get W");
    std::vector<int32_t> slotDiag(nCells), slotUpper(nInt), slotLower(nInt);
    check(rdst_mesh_get_int(mesh, RDST_M_SLOT_DIAG, slotDiag.data()), "slot diag");
    check(rdst_mesh_get_int(mesh, RDST_M_SLOT_UPPER, slotUpper.data()), "slot upper");
    check(rdst_mesh_get_int(mesh, RDST_M_SLOT_LOWER, slotLower.data()), "slot lower");
    std::vector<double> diag(25 * (size_t)nCells), upper(25 * (size_t)nInt), lower(25 * (size_t)nInt);
    for (int c = 0; c < nCells; ++c)
        for (int e = 0; e < 25; ++e) diag[25 * (size_t)c + e] = bcsr[25 * (size_t)slotDiag[c] + e];
    for (int f = 0; f < nInt; ++f)
        for (int e = 0; e < 25; ++e) {
            upper[25 * (size_t)f + e] = bcsr[25 * (size_t)slotUpper[f] + e];
            lower[25 * (size_t)f + e] = bcsr[25 * (size_t)slotLower[f] + e];
        }
    std::vector<double> xLdu = W0;
    rdst_solver_perf pl{}, pb{};
    check(rdst_ldu_solve(cs, diag.data(), upper.data(), lower.data(), src.data(), xLdu.data(), RDST_LDU_W, &pl),
          "rdst_ldu_solve");
    check(rdst_jacobian_assemble(cs), "rdst_jacobian_assemble (2)");
    check(rdst_block_solve(cs, &pb), "rdst_block_solve");
    std::vector<double> W1(5 * (size_t)nCells);
    check(rdst_case_get(cs, RDST_F_W, W1.data()), "get W (2)");
    double dmax = 0, wmax = 0;
    for (size_t i = 0; i < W1.size(); ++i) {
        dmax = std::fmax(dmax, std::fabs(W1[i] - xLdu[i]));
        wmax = std::fmax(wmax, std::fabs(W1[i]));
    }

    // Single::iterate from the start state
    check(rdst_case_set_state(cs, p.data(), T.data(), U.data()), "rdst_case_set_state (2)");
    std::vector<double> res(nSteps);
    check(rdst_iterate(cs, nSteps, res.data(), nullptr), "rdst_iterate");
    const double drag[3] = {1, 0, 0}, lift[3] = {0, 1, 0};
    double fa[8], fb[8];
    check(rdst_forces(cs, forcePatch, drag, lift, 1.0e5, 1.2, 100.0, 1.0e-4, fa), "rdst_forces");
    check(rdst_forces_noalloc(cs, forcePatch, drag, lift, 1.0e5, 1.2, 100.0, 1.0e-4, fb), "rdst_forces_noalloc");

    std::printf("{");
    print_array("res", res.data(), nSteps);
    print_array("forces", fa, 8);
    print_array("forces_noalloc", fb, 8);
    std::printf("\"ldu_vs_block_solve\": %.17g, \"ldu_iterations\": %d, \"block_iterations\": %d, "
                "\"geometry_kept\": %d, \"unknown_scheme_rc\": %d, \"n_cells\": %d}\n",
                wmax > 0 ? dmax / wmax : dmax, pl.n_iterations, pb.n_iterations, geomSame, unknownRc, nCells);
    check(rdst_case_destroy(cs), "rdst_case_destroy");
    check(rdst_mesh_destroy(mesh), "rdst_mesh_destroy");
    check(rdst_mesh_destroy(mesh0), "rdst_mesh_destroy (0)");
    return 0;
}
