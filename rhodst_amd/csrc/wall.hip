// wall.hip -- exact distance from every cell centre to the wall patches: the
// wallDist d_ of the Spalart-Allmaras models (SAneg.C:160, SARC2.C:158 hold a
// wallDist; foam-extend computes it by meshWave propagation).  Here it is the
// exact Euclidean distance to the nearest wall triangle (each wall face fanned
// from the mean of its points), evaluated on the device: cells in blocks of 256
// threads, the wall triangles streamed through LDS in tiles, each triangle
// first rejected by its bounding sphere against the running minimum, then the
// closest point found by the region tests of Ericson, Real-Time Collision
// Detection (2004), section 5.1.5.
#include <hip/hip_runtime.h>

#include <vector>

#include "rdst_internal.hpp"

namespace rdst {

namespace {

constexpr int WD_THREADS = 256;
constexpr int WD_TILE = 256;  // triangles per LDS tile: 256 x 13 doubles = 26 KB

struct Tri {
    double ax, ay, az, bx, by, bz, cx, cy, cz;  // corners
    double ox, oy, oz, r;                       // bounding sphere (centroid, radius)
};

__device__ inline double dot3(double ax, double ay, double az, double bx, double by, double bz) {
    return ax * bx + ay * by + az * bz;
}

// squared distance from p to triangle t
__device__ inline double tri_dist2(double px, double py, double pz, const Tri& t) {
    const double abx = t.bx - t.ax, aby = t.by - t.ay, abz = t.bz - t.az;
    const double acx = t.cx - t.ax, acy = t.cy - t.ay, acz = t.cz - t.az;
    const double apx = px - t.ax, apy = py - t.ay, apz = pz - t.az;
    double qx, qy, qz;
    const double d1 = dot3(abx, aby, abz, apx, apy, apz), d2 = dot3(acx, acy, acz, apx, apy, apz);
    const double bpx = px - t.bx, bpy = py - t.by, bpz = pz - t.bz;
    const double d3 = dot3(abx, aby, abz, bpx, bpy, bpz), d4 = dot3(acx, acy, acz, bpx, bpy, bpz);
    const double cpx = px - t.cx, cpy = py - t.cy, cpz = pz - t.cz;
    const double d5 = dot3(abx, aby, abz, cpx, cpy, cpz), d6 = dot3(acx, acy, acz, cpx, cpy, cpz);
    const double vc = d1 * d4 - d3 * d2, vb = d5 * d2 - d1 * d6, va = d3 * d6 - d5 * d4;
    if (d1 <= 0.0 && d2 <= 0.0) {                       // vertex A
        qx = t.ax; qy = t.ay; qz = t.az;
    } else if (d3 >= 0.0 && d4 <= d3) {                 // vertex B
        qx = t.bx; qy = t.by; qz = t.bz;
    } else if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {   // edge AB
        const double v = d1 / (d1 - d3);
        qx = t.ax + v * abx; qy = t.ay + v * aby; qz = t.az + v * abz;
    } else if (d6 >= 0.0 && d5 <= d6) {                 // vertex C
        qx = t.cx; qy = t.cy; qz = t.cz;
    } else if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {   // edge AC
        const double w = d2 / (d2 - d6);
        qx = t.ax + w * acx; qy = t.ay + w * acy; qz = t.az + w * acz;
    } else if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {  // edge BC
        const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        qx = t.bx + w * (t.cx - t.bx); qy = t.by + w * (t.cy - t.by); qz = t.bz + w * (t.cz - t.bz);
    } else {                                            // face interior
        const double den = 1.0 / (va + vb + vc);
        const double v = vb * den, w = vc * den;
        qx = t.ax + abx * v + acx * w; qy = t.ay + aby * v + acy * w; qz = t.az + abz * v + acz * w;
    }
    const double dx = px - qx, dy = py - qy, dz = pz - qz;
    return dx * dx + dy * dy + dz * dz;
}

__global__ __launch_bounds__(WD_THREADS) void k_wall_distance(int nC, const double* __restrict__ C,
                                                              const Tri* __restrict__ tri, int nT,
                                                              double* __restrict__ d) {
    __shared__ Tri tile[WD_TILE];
    const int c = blockIdx.x * WD_THREADS + threadIdx.x;
    const bool live = c < nC;
    const double px = live ? C[3 * (size_t)c] : 0.0;
    const double py = live ? C[3 * (size_t)c + 1] : 0.0;
    const double pz = live ? C[3 * (size_t)c + 2] : 0.0;
    double best = 1.0e300, sbest = 1.0e150;
    for (int t0 = 0; t0 < nT; t0 += WD_TILE) {
        const int nt = min(WD_TILE, nT - t0);
        __syncthreads();
        for (int k = threadIdx.x; k < nt; k += WD_THREADS) tile[k] = tri[t0 + k];
        __syncthreads();
        if (!live) continue;
        for (int k = 0; k < nt; ++k) {
            const Tri& t = tile[k];
            const double ox = px - t.ox, oy = py - t.oy, oz = pz - t.oz;
            const double lim = sbest + t.r;
            if (ox * ox + oy * oy + oz * oz >= lim * lim) continue;  // the sphere cannot beat the minimum
            const double d2 = tri_dist2(px, py, pz, t);
            if (d2 < best) {
                best = d2;
                sbest = sqrt(d2);
            }
        }
    }
    if (live) d[c] = sqrt(best);
}

}  // namespace

void wall_distance(const Mesh& m, hipStream_t stream, double* d_host) {
    const int nT = (int)(m.wallTri.size() / 9);
    if (nT == 0) throw ArgError("wall distance: the mesh has no wall-patch faces");
    std::vector<Tri> tri(nT);
    for (int t = 0; t < nT; ++t) {
        const double* q = &m.wallTri[9 * (size_t)t];
        Tri& o = tri[t];
        o.ax = q[0]; o.ay = q[1]; o.az = q[2];
        o.bx = q[3]; o.by = q[4]; o.bz = q[5];
        o.cx = q[6]; o.cy = q[7]; o.cz = q[8];
        o.ox = (q[0] + q[3] + q[6]) / 3.0;
        o.oy = (q[1] + q[4] + q[7]) / 3.0;
        o.oz = (q[2] + q[5] + q[8]) / 3.0;
        double r2 = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double dx = q[3 * k] - o.ox, dy = q[3 * k + 1] - o.oy, dz = q[3 * k + 2] - o.oz;
            r2 = std::max(r2, dx * dx + dy * dy + dz * dz);
        }
        // a hair above the exact radius, so the rejection test never drops the nearest triangle
        o.r = std::sqrt(r2) * (1.0 + 1e-12) + 1e-300;
    }
    DBuf<Tri> dTri;
    DBuf<double> dC, dD;
    dTri.upload(tri.data(), tri.size());
    dC.upload(m.C.data(), m.C.size());
    dD.alloc(m.nC);
    hipLaunchKernelGGL(k_wall_distance, dim3(blocks_for(m.nC, WD_THREADS)), dim3(WD_THREADS), 0, stream, m.nC, dC.p,
                       dTri.p, nT, dD.p);
    RDST_HIP(hipGetLastError());
    RDST_HIP(hipMemcpyAsync(d_host, dD.p, sizeof(double) * m.nC, hipMemcpyDeviceToHost, stream));
    RDST_HIP(hipStreamSynchronize(stream));
}

}  // namespace rdst
