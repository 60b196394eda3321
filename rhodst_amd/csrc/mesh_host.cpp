// mesh_host.cpp -- fvMesh geometry and addressing for the device path.
//
// The reference reads Sf, magSf, Cf, C, V, weights and deltaCoeffs from
// foam-extend's fvMesh (e.g. Roe.C:79, :195-197; timeDST.C:237); this file
// computes the same quantities from the polyMesh primitives on the host
// (face centres/areas by triangle decomposition about the point average,
// cell centres/volumes by face pyramids about the face-centre average,
// linear weights from the face-normal distances, stabilised deltaCoeffs
// 1/max(n.d, 0.05|d|) as the comment at timeDST.C:237 states) and derives
// the device addressing: the ELL cell->face map the cell-centric face loop
// walks, the owner/neighbour-sorted block-CSR with its face->slot maps, a
// greedy face colouring, and the level sets of the natural-order ILU sweeps.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "rdst_internal.hpp"

namespace rdst {

namespace {
struct P3 { double x, y, z; };
inline P3 add(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline P3 sub(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline P3 scl(double s, P3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dotp(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline P3 crossp(P3 a, P3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(P3 a) { return std::sqrt(dotp(a, a)); }
}  // namespace

namespace {

// Recursive coordinate bisection: split the cell set at the median of its
// longest extent (ties broken by cell index) until a set holds at most
// maxCells cells.
void rcb(const std::vector<double>& C, std::vector<int>& idx, int lo, int hi, int maxCells, std::vector<int32_t>& part,
         int& nParts) {
    if (hi - lo <= maxCells) {
        for (int k = lo; k < hi; ++k) part[idx[k]] = nParts;
        ++nParts;
        return;
    }
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int k = lo; k < hi; ++k)
        for (int d = 0; d < 3; ++d) {
            mn[d] = std::min(mn[d], C[3 * (size_t)idx[k] + d]);
            mx[d] = std::max(mx[d], C[3 * (size_t)idx[k] + d]);
        }
    int ax = 0;
    for (int d = 1; d < 3; ++d) if (mx[d] - mn[d] > mx[ax] - mn[ax]) ax = d;
    const int mid = lo + (hi - lo) / 2;
    std::nth_element(idx.begin() + lo, idx.begin() + mid, idx.begin() + hi, [&](int a, int b) {
        double ca = C[3 * (size_t)a + ax], cb = C[3 * (size_t)b + ax];
        return ca < cb || (ca == cb && a < b);
    });
    rcb(C, idx, lo, mid, maxCells, part, nParts);
    rcb(C, idx, mid, hi, maxCells, part, nParts);
}

}  // namespace

// The forward (L) or backward (Dinv, Dinv U) stream of the partitioned
// ILU(0) apply, in the layout PiluStream describes.  Left empty (kmax = 0)
// when a row has more than kPiluKD blocks in the sweep.
static void build_pilu_stream(const Mesh& m, bool bwd, PiluStream& ps) {
    ps.kmax = 0;
    ps.nBlk = 0;
    std::vector<int32_t> subStart(m.nParts + 1, 0), sub, ord, lc, src, row, jn;
    int kmaxAll = 0;
    struct RowBlocks {
        int t, cnt;
        int32_t src[kPiluKD], lc[kPiluKD];
    };
    std::vector<RowBlocks> rows;
    for (int p = 0; p < m.nParts; ++p) {
        const int rs = m.partRowStart[p];
        const int l0 = m.partLevStart[p], nl = m.partLevStart[p + 1] - l0 - 1;
        for (int li = 0; li < nl; ++li) {
            const int L = bwd ? nl - 1 - li : li;
            rows.clear();
            for (int t = m.partLevPtr[l0 + L]; t < m.partLevPtr[l0 + L + 1]; ++t) {
                const int i = m.partRows[t];
                RowBlocks rb;
                rb.t = t - rs;
                rb.cnt = 0;
                auto add = [&](int32_t sv, int32_t lv) {
                    if (rb.cnt == kPiluKD) { rb.cnt = kPiluKD + 1; return; }
                    if (rb.cnt > kPiluKD) return;
                    rb.src[rb.cnt] = sv;
                    rb.lc[rb.cnt] = lv;
                    ++rb.cnt;
                };
                if (!bwd) {
                    for (int s = m.rowPtr[i]; s < m.slotDiag[i]; ++s)
                        if (m.localCol[s] >= 0) add(s, m.localCol[s]);
                    if (rb.cnt == 0) continue;  // y = r on this row
                } else {
                    add(-(i + 1), t - rs);
                    for (int s = m.slotDiag[i] + 1; s < m.rowPtr[i + 1]; ++s)
                        if (m.localCol[s] >= 0) add(s, m.localCol[s]);
                }
                if (rb.cnt > kPiluKD) return;  // stays unbuilt
                rows.push_back(rb);
            }
            for (size_t r0 = 0; r0 < rows.size(); r0 += kPiluSubRows) {
                const size_t r1 = std::min(rows.size(), r0 + kPiluSubRows);
                std::stable_sort(rows.begin() + r0, rows.begin() + r1,
                                 [](const RowBlocks& a, const RowBlocks& b) { return a.cnt > b.cnt; });
                const int nr = (int)(r1 - r0), km = rows[r0].cnt;
                kmaxAll = std::max(kmaxAll, km);
                int32_t d[kPiluSubW] = {0};
                d[0] = (int32_t)ord.size();
                d[1] = nr;
                d[2] = km;
                for (size_t q = r0; q < r1; ++q) ord.push_back(rows[q].t);
                for (int k = 0; k < km; ++k) {
                    int nk = 0;
                    while (nk < nr && rows[r0 + nk].cnt > k) ++nk;
                    d[4 + k] = (int32_t)src.size();
                    d[4 + kPiluKD + k] = nk;
                    for (int j = 0; j < nk; ++j) {
                        const RowBlocks& rb = rows[r0 + j];
                        src.push_back(rb.src[k]);
                        lc.push_back(rb.lc[k]);
                        row.push_back(m.partRows[rs + rb.t]);
                        jn.push_back(j | (nk << 8));
                    }
                }
                sub.insert(sub.end(), d, d + kPiluSubW);
            }
        }
        subStart[p + 1] = (int32_t)(sub.size() / kPiluSubW);
    }
    sub.insert(sub.end(), kPiluSubW, 0);
    ps.kmax = std::max(kmaxAll, 1);
    ps.nBlk = (long)src.size();
    ps.subStart.swap(subStart);
    ps.sub.swap(sub);
    ps.ord.swap(ord);
    ps.lc.swap(lc);
    ps.src.swap(src);
    ps.row.swap(row);
    ps.jn.swap(jn);
}

// Partition-local level of every cell (longest chain of in-partition lower
// neighbours, cells in increasing index); returns the deepest level count.
static int partition_levels(const Mesh& m, std::vector<int32_t>& lev) {
    int deepest = 0;
    for (int c = 0; c < m.nC; ++c) {
        int l = 0;
        for (int s = m.rowPtr[c]; s < m.slotDiag[c]; ++s)
            if (m.part[m.cols[s]] == m.part[c]) l = std::max(l, lev[m.cols[s]] + 1);
        lev[c] = l;
        deepest = std::max(deepest, l + 1);
    }
    return deepest;
}

// The apply of a partition is a chain of its levels; a partition deeper than
// maxLevels is cut at its middle level (the cells of the upper half form a
// new partition), until none is.
static void split_deep_partitions(Mesh& m, int maxLevels) {
    std::vector<int32_t> lev(m.nC, 0);
    for (int pass = 0; pass < 16; ++pass) {
        partition_levels(m, lev);
        std::vector<int> depth(m.nParts, 0);
        for (int c = 0; c < m.nC; ++c) depth[m.part[c]] = std::max(depth[m.part[c]], lev[c] + 1);
        std::vector<int> newId(m.nParts, -1);
        int added = 0;
        for (int p = 0; p < m.nParts; ++p)
            if (depth[p] > maxLevels) newId[p] = m.nParts + added++;
        if (!added) return;
        for (int c = 0; c < m.nC; ++c) {
            const int p = m.part[c];
            if (newId[p] >= 0 && lev[c] >= depth[p] / 2) m.part[c] = newId[p];
        }
        m.nParts += added;
    }
}

// Rows of each partition in (partition-local level, cell) order, the level
// offsets, and per block-CSR slot the partition-local column (-1 outside).
static void order_partitions(const Mesh& m, const std::vector<int32_t>& part, int nParts, PartSet& ps) {
    const int n = m.nC;
    std::vector<std::vector<int>> cellsOf(nParts);
    for (int c = 0; c < n; ++c) cellsOf[part[c]].push_back(c);
    std::vector<int32_t> lev(n, 0), local(n, -1);
    ps.nParts = nParts;
    ps.rowStart.assign(nParts + 1, 0);
    ps.levStart.assign(nParts + 1, 0);
    ps.rows.clear();
    ps.levPtr.clear();
    ps.maxRows = 0;
    ps.maxLevels = 0;
    for (int p = 0; p < nParts; ++p) {
        const std::vector<int>& cs = cellsOf[p];
        int nl = 0;
        for (int c : cs) {
            int l = 0;
            for (int s = m.rowPtr[c]; s < m.slotDiag[c]; ++s)
                if (part[m.cols[s]] == p) l = std::max(l, lev[m.cols[s]] + 1);
            lev[c] = l;
            nl = std::max(nl, l + 1);
        }
        std::vector<int> order(cs.begin(), cs.end());
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return lev[a] < lev[b]; });
        ps.rowStart[p + 1] = ps.rowStart[p] + (int)order.size();
        for (size_t k = 0; k < order.size(); ++k) {
            local[order[k]] = (int)k;
            ps.rows.push_back(order[k]);
        }
        // level offsets (absolute positions in rows)
        int pos = ps.rowStart[p];
        for (int l = 0; l < nl; ++l) {
            ps.levPtr.push_back(pos);
            while (pos < ps.rowStart[p + 1] && lev[ps.rows[pos]] == l) ++pos;
        }
        ps.levPtr.push_back(ps.rowStart[p + 1]);
        ps.levStart[p + 1] = (int)ps.levPtr.size();
        ps.maxRows = std::max(ps.maxRows, (int)order.size());
        ps.maxLevels = std::max(ps.maxLevels, nl);
    }
    ps.localCol.assign(m.nnzb, -1);
    for (int c = 0; c < n; ++c)
        for (int s = m.rowPtr[c]; s < m.rowPtr[c + 1]; ++s)
            if (part[m.cols[s]] == part[c]) ps.localCol[s] = local[m.cols[s]];
}

// Partitions for the block-Jacobi ILU(0): couplings between partitions are
// dropped, inside a partition the factor is the natural-order ILU(0) of the
// partition's diagonal block, scheduled by local levels.  The RCB boxes as
// they come (at most maxCells cells, no level cut) also drive the
// Spalart-Allmaras Gauss-Seidel sweep: a mesh of at most maxCells cells is
// one box, i.e. the serial natural-order sweep.
void build_partitions(Mesh& m, int maxCells) {
    const int n = m.nC;
    m.part.assign(n, 0);
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    m.nParts = 0;
    rcb(m.C, idx, 0, n, maxCells, m.part, m.nParts);
    m.saPart = m.part;
    order_partitions(m, m.saPart, m.nParts, m.sa);
    split_deep_partitions(m, kMaxPartLevels);
    PartSet ps;
    order_partitions(m, m.part, m.nParts, ps);
    m.partRowStart.swap(ps.rowStart);
    m.partRows.swap(ps.rows);
    m.partLevStart.swap(ps.levStart);
    m.partLevPtr.swap(ps.levPtr);
    m.localCol.swap(ps.localCol);
    m.maxPartRows = ps.maxRows;
    m.maxPartLevels = ps.maxLevels;
    build_pilu_stream(m, false, m.pf);
    build_pilu_stream(m, true, m.pb);
}

void mesh_build(Mesh& m, const rdst_mesh_desc* d, const rdst_mesh_geometry* g) {
    if (!d || d->n_cells <= 0 || d->n_faces <= 0 || d->n_internal_faces < 0 || d->n_internal_faces > d->n_faces)
        throw ArgError("rdst_mesh_create: bad mesh sizes");
    if (!d->points || !d->face_offsets || !d->face_points || !d->owner || (d->n_internal_faces && !d->neighbour))
        throw ArgError("rdst_mesh_create: null array");
    const int nFaces = d->n_faces, nInt = d->n_internal_faces, nCells = d->n_cells;
    for (int f = 0; f < nFaces; ++f) {
        if (d->owner[f] < 0 || d->owner[f] >= nCells) throw ArgError("rdst_mesh_create: owner out of range");
        if (d->face_offsets[f + 1] - d->face_offsets[f] < 3) throw ArgError("rdst_mesh_create: face with < 3 points");
    }
    for (int f = 0; f < nInt; ++f) {
        if (d->neighbour[f] < 0 || d->neighbour[f] >= nCells) throw ArgError("rdst_mesh_create: neighbour out of range");
        if (d->neighbour[f] <= d->owner[f]) throw ArgError("rdst_mesh_create: internal faces must have owner < neighbour");
    }
    m.patches.assign(d->patches, d->patches + d->n_patches);
    m.nPatches = d->n_patches;
    m.nC = nCells;
    m.nI = nInt;

    // --- face centres and areas
    std::vector<P3> fSf(nFaces), fCf(nFaces);
    std::vector<double> fmag(nFaces);
    auto pt = [&](int i) { return P3{d->points[3 * i], d->points[3 * i + 1], d->points[3 * i + 2]}; };
    for (int f = 0; f < nFaces; ++f) {
        const int s = d->face_offsets[f], n = d->face_offsets[f + 1] - s;
        const int32_t* fp = d->face_points + s;
        for (int k = 0; k < n; ++k)
            if (fp[k] < 0 || fp[k] >= d->n_points) throw ArgError("rdst_mesh_create: face point out of range");
        if (n == 3) {
            P3 p0 = pt(fp[0]), p1 = pt(fp[1]), p2 = pt(fp[2]);
            fCf[f] = scl(1.0 / 3.0, add(add(p0, p1), p2));
            fSf[f] = scl(0.5, crossp(sub(p1, p0), sub(p2, p0)));
        } else {
            P3 c = pt(fp[0]);
            for (int k = 1; k < n; ++k) c = add(c, pt(fp[k]));
            c = {c.x / n, c.y / n, c.z / n};
            P3 sN = {0, 0, 0}, sAc = {0, 0, 0};
            double sA = 0;
            for (int k = 0; k < n; ++k) {
                P3 a = pt(fp[k]), b = pt(fp[(k + 1) % n]);
                P3 cc = add(add(a, b), c);
                P3 nv = crossp(sub(b, a), sub(c, a));
                double ar = norm(nv);
                sN = add(sN, nv);
                sA += ar;
                sAc = add(sAc, scl(ar, cc));
            }
            if (sA < 1e-150) {
                fCf[f] = c;
                fSf[f] = {0, 0, 0};
            } else {
                P3 t = scl(1.0 / 3.0, sAc);  // (1/3)*sumAc/sumA, left to right
                fCf[f] = {t.x / sA, t.y / sA, t.z / sA};
                fSf[f] = scl(0.5, sN);
            }
        }
        fmag[f] = norm(fSf[f]);
    }
    // --- wall triangles (the surface the wall distance is measured to): each
    // wall face fanned from the mean of its points, one triangle per edge, so
    // a warped quad is followed on both of its diagonals (a fan from one
    // corner can pass a first-cell centre at many times its height)
    for (int pi = 0; pi < d->n_patches; ++pi) {
        const rdst_patch& pa = d->patches[pi];
        if (pa.type != RDST_PATCH_WALL) continue;
        if (pa.start < nInt || pa.size < 0 || pa.start + pa.size > nFaces) throw ArgError("rdst_mesh_create: patch range");
        for (int f = pa.start; f < pa.start + pa.size; ++f) {
            const int s = d->face_offsets[f], n = d->face_offsets[f + 1] - s;
            const int32_t* fp = d->face_points + s;
            double c[3] = {d->points[3 * fp[0]], d->points[3 * fp[0] + 1], d->points[3 * fp[0] + 2]};
            for (int k = 1; k < n; ++k)
                for (int x = 0; x < 3; ++x) c[x] = c[x] + d->points[3 * fp[k] + x];
            for (int x = 0; x < 3; ++x) c[x] = c[x] / n;
            for (int k = 0; k < n; ++k) {
                for (int x = 0; x < 3; ++x) m.wallTri.push_back(c[x]);
                for (int q : {fp[k], fp[(k + 1) % n]})
                    for (int x = 0; x < 3; ++x) m.wallTri.push_back(d->points[3 * q + x]);
            }
        }
    }
    // --- cell centres and volumes
    std::vector<P3> cEst(nCells, P3{0, 0, 0}), cC(nCells, P3{0, 0, 0});
    std::vector<int> nCF(nCells, 0);
    std::vector<double> vol(nCells, 0.0);
    for (int f = 0; f < nFaces; ++f) { cEst[d->owner[f]] = add(cEst[d->owner[f]], fCf[f]); nCF[d->owner[f]]++; }
    for (int f = 0; f < nInt; ++f) { cEst[d->neighbour[f]] = add(cEst[d->neighbour[f]], fCf[f]); nCF[d->neighbour[f]]++; }
    for (int c = 0; c < nCells; ++c) cEst[c] = {cEst[c].x / nCF[c], cEst[c].y / nCF[c], cEst[c].z / nCF[c]};
    for (int f = 0; f < nFaces; ++f) {
        int o = d->owner[f];
        double v3 = std::max(dotp(fSf[f], sub(fCf[f], cEst[o])), 1e-300);
        P3 pc = add(scl(0.75, fCf[f]), scl(0.25, cEst[o]));
        cC[o] = add(cC[o], scl(v3, pc));
        vol[o] += v3;
    }
    for (int f = 0; f < nInt; ++f) {
        int nb = d->neighbour[f];
        double v3 = std::max(dotp(fSf[f], sub(cEst[nb], fCf[f])), 1e-300);
        P3 pc = add(scl(0.75, fCf[f]), scl(0.25, cEst[nb]));
        cC[nb] = add(cC[nb], scl(v3, pc));
        vol[nb] += v3;
    }
    for (int c = 0; c < nCells; ++c) {
        cC[c] = {cC[c].x / vol[c], cC[c].y / vol[c], cC[c].z / vol[c]};
        vol[c] *= (1.0 / 3.0);
        if (!(vol[c] > 0)) throw ArgError("rdst_mesh_create: non-positive cell volume");
    }
    // --- compacted boundary faces (empty patches have fvPatch size 0)
    m.bPatch.clear();
    m.bPolyFace.clear();
    for (int p = 0; p < m.nPatches; ++p) {
        const rdst_patch& pp = m.patches[p];
        if (pp.start < nInt || pp.start + pp.size > nFaces) throw ArgError("rdst_mesh_create: patch range");
        if (pp.type == RDST_PATCH_EMPTY) continue;
        for (int i = 0; i < pp.size; ++i) { m.bPatch.push_back(p); m.bPolyFace.push_back(pp.start + i); }
    }
    m.nB = (int)m.bPatch.size();
    m.nF = m.nI + m.nB;
    const int nF = m.nF;
    m.Sf.resize(3 * (size_t)nF); m.Cf.resize(3 * (size_t)nF); m.magSf.resize(nF); m.w.assign(nF, 1.0); m.dc.resize(nF);
    m.owner.resize(nF); m.neigh.assign(d->neighbour, d->neighbour + nInt);
    for (int k = 0; k < nF; ++k) {
        int f = k < nInt ? k : m.bPolyFace[k - nInt];
        m.Sf[3 * k] = fSf[f].x; m.Sf[3 * k + 1] = fSf[f].y; m.Sf[3 * k + 2] = fSf[f].z;
        m.Cf[3 * k] = fCf[f].x; m.Cf[3 * k + 1] = fCf[f].y; m.Cf[3 * k + 2] = fCf[f].z;
        m.magSf[k] = fmag[f];
        m.owner[k] = d->owner[f];
    }
    m.C.resize(3 * (size_t)nCells); m.V = vol;
    for (int c = 0; c < nCells; ++c) { m.C[3 * c] = cC[c].x; m.C[3 * c + 1] = cC[c].y; m.C[3 * c + 2] = cC[c].z; }
    for (int f = 0; f < nInt; ++f) {
        P3 sf = fSf[f], cf = fCf[f], co = cC[d->owner[f]], cn = cC[d->neighbour[f]];
        double dOwn = std::fabs(dotp(sf, sub(cf, co)));
        double dNei = std::fabs(dotp(sf, sub(cn, cf)));
        m.w[f] = dNei / (dOwn + dNei);
        P3 delta = sub(cn, co);
        P3 ua = {sf.x / fmag[f], sf.y / fmag[f], sf.z / fmag[f]};
        m.dc[f] = 1.0 / std::max(dotp(ua, delta), 0.05 * norm(delta));
    }
    for (int b = 0; b < m.nB; ++b) {
        int f = m.bPolyFace[b];
        P3 delta = sub(fCf[f], cC[d->owner[f]]);
        m.dc[nInt + b] = 1.0 / norm(delta);
    }
    // the host fvMesh's own geometry, kept bit for bit (faces of empty
    // patches, which have no fvPatch faces, are dropped)
    if (g) {
        if (!g->Sf || !g->Cf || !g->C || !g->V || !g->weights || !g->delta_coeffs)
            throw ArgError("rdst_mesh_create_geometry: null geometry array");
        for (int k = 0; k < nF; ++k) {
            const int f = k < nInt ? k : m.bPolyFace[k - nInt];
            for (int x = 0; x < 3; ++x) {
                m.Sf[3 * k + x] = g->Sf[3 * (size_t)f + x];
                m.Cf[3 * k + x] = g->Cf[3 * (size_t)f + x];
            }
            m.magSf[k] = g->mag_sf ? g->mag_sf[f] : std::sqrt(m.Sf[3 * k] * m.Sf[3 * k] + m.Sf[3 * k + 1] * m.Sf[3 * k + 1] +
                                                             m.Sf[3 * k + 2] * m.Sf[3 * k + 2]);
            m.w[k] = k < nInt ? g->weights[f] : 1.0;
            m.dc[k] = g->delta_coeffs[f];
            if (!(m.magSf[k] > 0)) throw ArgError("rdst_mesh_create_geometry: face of zero area");
        }
        for (int c = 0; c < nCells; ++c) {
            for (int x = 0; x < 3; ++x) m.C[3 * (size_t)c + x] = g->C[3 * (size_t)c + x];
            m.V[c] = g->V[c];
            if (!(m.V[c] > 0)) throw ArgError("rdst_mesh_create_geometry: non-positive cell volume");
        }
    }
    // correction vectors n - d deltaCoeffs of the corrected snGrad
    m.corr.resize(3 * (size_t)nInt);
    for (int f = 0; f < nInt; ++f) {
        const int o = d->owner[f], nb = d->neighbour[f];
        P3 delta = {m.C[3 * (size_t)nb] - m.C[3 * (size_t)o], m.C[3 * (size_t)nb + 1] - m.C[3 * (size_t)o + 1],
                    m.C[3 * (size_t)nb + 2] - m.C[3 * (size_t)o + 2]};
        P3 ua = {m.Sf[3 * f] / m.magSf[f], m.Sf[3 * f + 1] / m.magSf[f], m.Sf[3 * f + 2] / m.magSf[f]};
        P3 cv = sub(ua, scl(m.dc[f], delta));
        m.corr[3 * f] = cv.x; m.corr[3 * f + 1] = cv.y; m.corr[3 * f + 2] = cv.z;
    }

    // --- ELL cell->face map, faces of a cell in surface-field order
    std::vector<int> deg(nCells, 0);
    for (int f = 0; f < nInt; ++f) { deg[m.owner[f]]++; deg[m.neigh[f]]++; }
    for (int b = 0; b < m.nB; ++b) deg[m.owner[nInt + b]]++;
    m.ellW = nCells ? *std::max_element(deg.begin(), deg.end()) : 0;
    const size_t E = (size_t)m.ellW * nCells;
    m.ellCode.assign(E, (ELL_NONE << 30));
    m.ellFace.assign(E, -1);
    std::vector<int> fill(nCells, 0);
    // faces in increasing surface-field index: internal faces then boundary
    for (int f = 0; f < nInt; ++f) {
        int o = m.owner[f], nb = m.neigh[f];
        size_t eo = (size_t)fill[o]++ * nCells + o;
        m.ellCode[eo] = (ELL_OWNER << 30) | (uint32_t)nb;
        m.ellFace[eo] = f;
        size_t en = (size_t)fill[nb]++ * nCells + nb;
        m.ellCode[en] = (ELL_NEIGH << 30) | (uint32_t)o;
        m.ellFace[en] = f;
    }
    for (int b = 0; b < m.nB; ++b) {
        int c = m.owner[nInt + b];
        size_t e = (size_t)fill[c]++ * nCells + c;
        m.ellCode[e] = (ELL_BOUNDARY << 30) | (uint32_t)b;
        m.ellFace[e] = nInt + b;
    }

    // --- block-CSR (owner/neighbour-sorted): row i holds the diagonal, the
    // upper coefficients of faces it owns (column = neighbour) and the lower
    // coefficients of faces it neighbours (column = owner), columns ascending
    std::vector<std::vector<std::pair<int, int>>> rows(nCells);  // (col, tag) tag: -1 diag, 2f upper, 2f+1 lower
    for (int c = 0; c < nCells; ++c) rows[c].push_back({c, -1});
    for (int f = 0; f < nInt; ++f) {
        rows[m.owner[f]].push_back({m.neigh[f], 2 * f});
        rows[m.neigh[f]].push_back({m.owner[f], 2 * f + 1});
    }
    m.rowPtr.assign(nCells + 1, 0);
    for (int c = 0; c < nCells; ++c) {
        std::sort(rows[c].begin(), rows[c].end());
        m.rowPtr[c + 1] = m.rowPtr[c] + (int)rows[c].size();
    }
    m.nnzb = m.rowPtr[nCells];
    m.cols.resize(m.nnzb);
    m.slotUpper.resize(nInt); m.slotLower.resize(nInt); m.slotDiag.resize(nCells);
    for (int c = 0; c < nCells; ++c)
        for (size_t k = 0; k < rows[c].size(); ++k) {
            int s = m.rowPtr[c] + (int)k;
            m.cols[s] = rows[c][k].first;
            int tag = rows[c][k].second;
            if (tag == -1) m.slotDiag[c] = s;
            else if ((tag & 1) == 0) m.slotUpper[tag >> 1] = s;
            else m.slotLower[tag >> 1] = s;
        }
    for (int c = 0; c < nCells; ++c)
        for (int s = m.rowPtr[c] + 1; s < m.rowPtr[c + 1]; ++s)
            if (m.cols[s] == m.cols[s - 1]) throw ArgError("rdst_mesh_create: duplicate internal face between two cells");

    // --- ELL slot of the off-diagonal block each cell writes: the owner
    // writes the lower coefficient (row neighbour, column owner), the
    // neighbour writes the upper one (row owner, column neighbour)
    m.ellSlot.assign(E, -1);
    for (size_t e = 0; e < E; ++e) {
        uint32_t kind = m.ellCode[e] >> 30;
        int f = m.ellFace[e];
        if (kind == ELL_OWNER) m.ellSlot[e] = m.slotLower[f];
        else if (kind == ELL_NEIGH) m.ellSlot[e] = m.slotUpper[f];
    }

    // --- greedy face colouring in face order (deterministic)
    m.faceColour.assign(nInt, -1);
    {
        std::vector<uint64_t> used(nCells, 0);
        int maxc = 0;
        for (int f = 0; f < nInt; ++f) {
            uint64_t u = used[m.owner[f]] | used[m.neigh[f]];
            int col = 0;
            while (col < 64 && (u >> col & 1ull)) ++col;
            if (col >= 64) throw ArgError("rdst_mesh_create: face colouring needs more than 64 colours");
            m.faceColour[f] = col;
            used[m.owner[f]] |= 1ull << col;
            used[m.neigh[f]] |= 1ull << col;
            maxc = std::max(maxc, col + 1);
        }
        m.nColours = maxc;
    }
    // --- greedy cell colouring (distance-1) in cell order
    m.cellColour.assign(nCells, -1);
    {
        int maxc = 0;
        for (int c = 0; c < nCells; ++c) {
            uint64_t u = 0;
            for (int s = m.rowPtr[c]; s < m.rowPtr[c + 1]; ++s) {
                int j = m.cols[s];
                if (j != c && m.cellColour[j] >= 0) u |= 1ull << m.cellColour[j];
            }
            int col = 0;
            while (col < 64 && (u >> col & 1ull)) ++col;
            if (col >= 64) throw ArgError("rdst_mesh_create: cell colouring needs more than 64 colours");
            m.cellColour[c] = col;
            maxc = std::max(maxc, col + 1);
        }
        m.nCellColours = maxc;
    }
    // --- level sets of the natural-order triangular sweeps
    m.levL.assign(nCells, 0);
    for (int c = 0; c < nCells; ++c) {
        int lv = 0;
        for (int s = m.rowPtr[c]; s < m.rowPtr[c + 1]; ++s)
            if (m.cols[s] < c) lv = std::max(lv, m.levL[m.cols[s]] + 1);
        m.levL[c] = lv;
    }
    m.levU.assign(nCells, 0);
    for (int c = nCells - 1; c >= 0; --c) {
        int lv = 0;
        for (int s = m.rowPtr[c]; s < m.rowPtr[c + 1]; ++s)
            if (m.cols[s] > c) lv = std::max(lv, m.levU[m.cols[s]] + 1);
        m.levU[c] = lv;
    }
    auto bucket = [&](const std::vector<int32_t>& lev, int& nLev, std::vector<int32_t>& ptr, std::vector<int32_t>& rowsOut) {
        nLev = nCells ? *std::max_element(lev.begin(), lev.end()) + 1 : 0;
        ptr.assign(nLev + 1, 0);
        for (int c = 0; c < nCells; ++c) ptr[lev[c] + 1]++;
        for (int l = 0; l < nLev; ++l) ptr[l + 1] += ptr[l];
        rowsOut.resize(nCells);
        std::vector<int> pos(ptr.begin(), ptr.end() - 1);
        for (int c = 0; c < nCells; ++c) rowsOut[pos[lev[c]]++] = c;
    };
    bucket(m.levL, m.nLevL, m.levLPtr, m.levLRows);
    bucket(m.levU, m.nLevU, m.levUPtr, m.levURows);
    auto runs = [&](const std::vector<int32_t>& ptr, int nLev, std::vector<LevelRun>& out) {
        out.clear();
        for (int L = 0; L < nLev;) {
            const bool small = ptr[L + 1] - ptr[L] <= kIluSegRows;
            int E = L + 1;
            while (E < nLev && ((ptr[E + 1] - ptr[E] <= kIluSegRows) == small)) ++E;
            out.push_back({L, E, small});
            L = E;
        }
    };
    runs(m.levLPtr, m.nLevL, m.runL);
    runs(m.levUPtr, m.nLevU, m.runU);
    build_partitions(m, kPartCells);
    // sliced ELL (64 rows per slice)
    m.nSlices = (nCells + 63) / 64;
    m.sellPtr.assign(m.nSlices + 1, 0);
    m.sellWidth.assign(m.nSlices, 0);
    for (int sl = 0; sl < m.nSlices; ++sl) {
        int w = 0;
        for (int r = 64 * sl; r < std::min(nCells, 64 * sl + 64); ++r) w = std::max(w, m.rowPtr[r + 1] - m.rowPtr[r]);
        m.sellWidth[sl] = w;
        m.sellPtr[sl + 1] = m.sellPtr[sl] + 64 * (int64_t)w;
    }
    m.sellBlocks = m.sellPtr[m.nSlices];
    m.sellCol.assign(m.sellBlocks, -1);
    m.sellOfSlot.assign(m.nnzb, 0);
    for (int r = 0; r < nCells; ++r) {
        const int sl = r / 64, lane = r % 64;
        for (int s = m.rowPtr[r]; s < m.rowPtr[r + 1]; ++s) {
            const int p = s - m.rowPtr[r];
            const int64_t pos = m.sellPtr[sl] + 64 * (int64_t)p + lane;
            m.sellCol[pos] = m.cols[s];
            m.sellOfSlot[s] = 25 * (m.sellPtr[sl] + 64 * (int64_t)p) + lane;
        }
        // padding blocks point at the row itself with zero values
        for (int p = m.rowPtr[r + 1] - m.rowPtr[r]; p < m.sellWidth[sl]; ++p) m.sellCol[m.sellPtr[sl] + 64 * (int64_t)p + lane] = r;
    }
    for (int64_t q = 0; q < m.sellBlocks; ++q)
        if (m.sellCol[q] < 0) m.sellCol[q] = 0;  // lanes past the last row of the last slice
}

void mesh_upload(Mesh& m) {
    const int nC = m.nC, nF = m.nF, W = m.ellW;
    m.d_owner.upload(m.owner.data(), m.nI);
    m.d_neigh.upload(m.neigh);
    std::vector<int32_t> bcell(m.nB), bwall(m.nB);
    for (int b = 0; b < m.nB; ++b) {
        bcell[b] = m.owner[m.nI + b];
        bwall[b] = m.patches[m.bPatch[b]].type == RDST_PATCH_WALL ? 1 : 0;
    }
    m.d_bcell.upload(bcell);
    m.d_bpatch.upload(m.bPatch);
    m.d_bWall.upload(bwall);
    std::vector<double> soa(3 * (size_t)nF);
    for (int k = 0; k < nF; ++k)
        for (int d = 0; d < 3; ++d) soa[(size_t)d * nF + k] = m.Sf[3 * (size_t)k + d];
    m.d_Sf.upload(soa);
    for (int k = 0; k < nF; ++k)
        for (int d = 0; d < 3; ++d) soa[(size_t)d * nF + k] = m.Cf[3 * (size_t)k + d];
    m.d_Cf.upload(soa);
    m.d_magSf.upload(m.magSf);
    m.d_w.upload(m.w);
    m.d_dc.upload(m.dc);
    m.d_V.upload(m.V);
    {
        std::vector<double> cs(3 * (size_t)m.nC);
        for (int c = 0; c < m.nC; ++c)
            for (int d = 0; d < 3; ++d) cs[(size_t)d * m.nC + c] = m.C[3 * (size_t)c + d];
        m.d_C.upload(cs);
    }
    m.d_ellCode.upload(m.ellCode);
    m.d_ellFace.upload(m.ellFace);
    m.d_ellSlot.upload(m.ellSlot);
    const size_t E = (size_t)W * nC;
    std::vector<double> esf(3 * E, 0.0), emag(E, 0.0), ew(E, 0.0), edc(E, 0.0), ecorr(3 * E, 0.0);
    for (size_t e = 0; e < E; ++e) {
        int k = m.ellFace[e];
        if (k < 0) continue;
        for (int d = 0; d < 3; ++d) esf[(size_t)d * E + e] = m.Sf[3 * (size_t)k + d];
        if (k < m.nI)
            for (int d = 0; d < 3; ++d) ecorr[(size_t)d * E + e] = m.corr[3 * (size_t)k + d];
        emag[e] = m.magSf[k];
        ew[e] = m.w[k];
        edc[e] = m.dc[k];
    }
    m.d_ellSf.upload(esf);
    m.d_ellMagSf.upload(emag);
    m.d_ellWt.upload(ew);
    m.d_ellDc.upload(edc);
    m.d_ellCorr.upload(ecorr);
    m.d_rowPtr.upload(m.rowPtr);
    m.d_cols.upload(m.cols);
    m.d_slotUpper.upload(m.slotUpper);
    m.d_slotLower.upload(m.slotLower);
    m.d_slotDiag.upload(m.slotDiag);
    m.d_levLRows.upload(m.levLRows);
    m.d_levLPtr.upload(m.levLPtr);
    m.d_levUPtr.upload(m.levUPtr);
    m.d_levURows.upload(m.levURows);
    m.d_partRowStart.upload(m.partRowStart);
    m.d_partRows.upload(m.partRows);
    m.d_partLevStart.upload(m.partLevStart);
    m.d_partLevPtr.upload(m.partLevPtr);
    m.d_localCol.upload(m.localCol);
    m.sa.d_rowStart.upload(m.sa.rowStart);
    m.sa.d_rows.upload(m.sa.rows);
    m.sa.d_levStart.upload(m.sa.levStart);
    m.sa.d_levPtr.upload(m.sa.levPtr);
    m.sa.d_localCol.upload(m.sa.localCol);
    for (PiluStream* ps : {&m.pf, &m.pb}) {
        if (ps->kmax == 0) continue;
        ps->d_subStart.upload(ps->subStart);
        ps->d_sub.upload(ps->sub);
        ps->d_ord.upload(ps->ord);
        ps->d_lc.upload(ps->lc);
        ps->d_src.upload(ps->src);
        ps->d_row.upload(ps->row);
        ps->d_jn.upload(ps->jn);
    }
    m.d_sellPtr.upload(m.sellPtr);
    m.d_sellOfSlot.upload(m.sellOfSlot);
    m.d_sellWidth.upload(m.sellWidth);
    m.d_sellCol.upload(m.sellCol);
}

}  // namespace rdst
