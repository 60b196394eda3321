// solve.hip -- the block linear solve A.x = b behind fvBlockMatrixDST::solve
// (rhoDST/fvBlockMatrixDST/fvBlockMatrixDST.C:98-155), i.e. the foam-extend
// BlockLduSolver<vector5> GMRES / BiCGStab with the ILUC0 preconditioner
// selected from fvSolution.solvers.W (Vector5typedefs.C:1596-1711).
//
// Matrix: owner/neighbour-sorted block-CSR of 5x5 row-major blocks.
// Preconditioner: block ILU(0) in the natural cell order, computed and
// applied level by level (rows of one level are independent), which gives
// the same factors as a sequential natural-order ILU(0) -- ILU(0) factors
// are unique for a given ordering and sparsity.
// GMRES: restarted, left-preconditioned, modified Gram-Schmidt; each MGS
// step is one fused kernel (w -= h_j V_j and the partial dot with the next
// vector), reductions are deterministic two-level (wavefront shuffles, then
// every block re-reduces the previous kernel's block partials in the same
// order), and the Hessenberg/Givens/back-substitution run on the device in
// FP64, so a restart cycle needs no host round trip.
// Mixed precision: FP32 matrix, FP32 ILU factors and FP32 Krylov basis
// inside an FP64 defect-correction loop (r = b - A x, norms, Hessenberg and
// Givens in FP64).
#include <algorithm>
#include <cmath>
#include <vector>

#include "rdst_internal.hpp"
#include "reduce.hpp"

namespace rdst {

struct SolverWork {
    int n = 0, n5 = 0, nDirs = 0, G = 0;
    bool mixed = false;
    DBuf<double> LU, Dinv;     // FP64 factors
    DBuf<float> LU32, Dinv32;  // FP32 factors
    DBuf<double> b, r, wA, tmp, x0;
    DBuf<double> dW;           // the correction of the deltaW / mixed forms
    DBuf<double> V;            // (nDirs+1) x n5 FP64 basis
    DBuf<float> V32, w32, r32, t32;
    DBuf<double> scal;         // device scalars
    DBuf<double> partA, partB; // reduction partials
    DBuf<double> multi;        // classical Gram-Schmidt partials [nDirs+1][G]
    DBuf<double> sf, sb;       // FP64 forward / backward ILU(0) apply streams
    DBuf<float> sf32, sb32;    // FP32 ones
    std::vector<double> hscal;
    // BiCGStab
    DBuf<double> p, ph, v, t, th, rw;
};

// scalar slots in SolverWork::scal
enum {
    S_BETA = 0, S_NORM = 1, S_INIT = 2 /*5*/, S_FIN = 7 /*5*/, S_XREF = 12 /*5*/, S_TMP = 17 /*8*/,
    S_H = 32  // then H[nDirs*nDirs], cs[nDirs], sn[nDirs], bh[nDirs+1], yh[nDirs]
};

namespace {

// ---------------------------------------------------------------- reductions
template <class TA, class TB>
__global__ __launch_bounds__(RED_BS) void k_dot_partial(long n, const TA* a, const TB* b, double* part) {
    __shared__ double sh[16];
    double s = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        s += (double)a[i] * (double)b[i];
    double t = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// out = sum(part) (op 0) or sqrt(sum) (op 1); single block
__global__ __launch_bounds__(RED_BS) void k_finalize(const double* part, int G, double* out, int op) {
    __shared__ double sh[16];
    double t = reduce_partials(part, G, sh);
    if (threadIdx.x == 0) out[0] = op == 1 ? sqrt(t) : t;
}

// per-component sum of |r| over an AoS [n][5] vector: partials [5][G]
__global__ __launch_bounds__(RED_BS) void k_cmptmag_partial(int n, const double* r, double* part) {
    __shared__ double sh[16];
    double s[5] = {0, 0, 0, 0, 0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        for (int k = 0; k < 5; ++k) s[k] += fabs(r[5 * i + k]);
    for (int k = 0; k < 5; ++k) {
        double t = block_sum(s[k], sh);
        if (threadIdx.x == 0) part[k * gridDim.x + blockIdx.x] = t;
    }
}
// out[k] = sum_k / norm
__global__ __launch_bounds__(RED_BS) void k_cmptmag_final(const double* part, int G, const double* norm, double* out) {
    __shared__ double sh[16];
    for (int k = 0; k < 5; ++k) {
        double t = reduce_partials(part + k * G, G, sh);
        if (threadIdx.x == 0) out[k] = t / norm[0];
    }
}
// per-component sums of x (for xRef): partials [5][G]
__global__ __launch_bounds__(RED_BS) void k_cmptsum_partial(int n, const double* x, double* part) {
    __shared__ double sh[16];
    double s[5] = {0, 0, 0, 0, 0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        for (int k = 0; k < 5; ++k) s[k] += x[5 * i + k];
    for (int k = 0; k < 5; ++k) {
        double t = block_sum(s[k], sh);
        if (threadIdx.x == 0) part[k * gridDim.x + blockIdx.x] = t;
    }
}
__global__ __launch_bounds__(RED_BS) void k_xref_final(const double* part, int G, int n, double* xref) {
    __shared__ double sh[16];
    for (int k = 0; k < 5; ++k) {
        double t = reduce_partials(part + k * G, G, sh);
        if (threadIdx.x == 0) xref[k] = t / n;
    }
}
// normFactor partials: sum |wA - pA| + |b - pA| (5-vector magnitudes)
__global__ __launch_bounds__(RED_BS) void k_norm_partial(int n, const double* wA, const double* pA, const double* b,
                                                          double* part) {
    __shared__ double sh[16];
    double s = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double a = 0, c = 0;
        for (int k = 0; k < 5; ++k) {
            double d1 = wA[5 * i + k] - pA[5 * i + k], d2 = b[5 * i + k] - pA[5 * i + k];
            a += d1 * d1;
            c += d2 * d2;
        }
        s += sqrt(a) + sqrt(c);
    }
    double t = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}
__global__ __launch_bounds__(RED_BS) void k_norm_final(const double* part, int G, double* norm) {
    __shared__ double sh[16];
    double t = reduce_partials(part, G, sh);
    if (threadIdx.x == 0) norm[0] = t + 1.0e-20;  // + BlockLduMatrix::small_
}

// ---------------------------------------------------------------- vectors
__global__ void k_fill_xref(int n, const double* xref, double* x) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) for (int k = 0; k < 5; ++k) x[5 * i + k] = xref[k];
}
template <class TY, class TX>
__global__ void k_copy(long n, TY* y, const TX* x) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] = (TY)x[i];
}
__global__ void k_sub(long n, double* r, const double* b, const double* a) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) r[i] = b[i] - a[i];
}
template <class T>
__global__ void k_acc(long n, T* y, const T* a) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] += a[i];
}
__global__ void k_add(long n, double* y, const double* a) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] += a[i];
}
// y = x / s[0]
template <class T>
__global__ void k_scale_div(long n, T* y, const T* x, const double* s) {
    const double d = s[0];
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = (T)((double)x[i] / d);
}
// fused MGS step: h = reduce(partIn); w -= h Vj; partial dot(w, Vnext) -> partOut; block 0 stores h
template <class T>
__global__ __launch_bounds__(RED_BS) void k_mgs_step(long n, T* w, const T* Vj, const T* Vnext, const double* partIn,
                                                      int G, double* partOut, double* hOut) {
    __shared__ double sh[16];
    const double h = reduce_partials(partIn, G, sh);
    if (blockIdx.x == 0 && threadIdx.x == 0) hOut[0] = h;
    double s = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        T wi = (T)((double)w[i] - h * (double)Vj[i]);
        w[i] = wi;
        s += (double)wi * (double)Vnext[i];
    }
    double t = block_sum(s, sh);
    if (threadIdx.x == 0) partOut[blockIdx.x] = t;
}
// x += sum_i yh[i] V_i
template <class TV, class TX>
__global__ void k_update_x(long n, TX* x, const TV* V, const double* yh, int m) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double xi = (double)x[i];
        for (int k = 0; k < m; ++k) xi += yh[k] * (double)V[(size_t)k * n + i];
        x[i] = (TX)xi;
    }
}

// ---------------------------------------------------------------- GMRES scalar work (one thread, FP64)
__device__ inline void givens(double h, double beta, double& c, double& s) {
    if (beta == 0) { c = 1; s = 0; }
    else if (fabs(beta) > fabs(h)) { double tau = -h / beta; s = 1.0 / sqrt(1.0 + tau * tau); c = s * tau; }
    else { double tau = -beta / h; c = 1.0 / sqrt(1.0 + tau * tau); s = c * tau; }
}
// beta = sqrt(reduce(part)); rotate column i, new rotation, update bh
__global__ __launch_bounds__(RED_BS) void k_gmres_column(double* sc, const double* part, int G, int i, int m) {
    __shared__ double sh[16];
    const double beta = sqrt(reduce_partials(part, G, sh));
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double* H = sc + S_H;
    double* cs = H + m * m;
    double* sn = cs + m;
    double* bh = sn + m;
    for (int j = 0; j < i; ++j) {
        const double Hji = H[j * m + i];
        H[j * m + i] = cs[j] * Hji - sn[j] * H[(j + 1) * m + i];
        H[(j + 1) * m + i] = sn[j] * Hji + cs[j] * H[(j + 1) * m + i];
    }
    givens(H[i * m + i], beta, cs[i], sn[i]);
    const double bhi = bh[i];
    bh[i] = cs[i] * bhi - sn[i] * bh[i + 1];
    bh[i + 1] = sn[i] * bhi + cs[i] * bh[i + 1];
    H[i * m + i] = cs[i] * H[i * m + i] - sn[i] * beta;
}
__global__ void k_gmres_start(double* sc, int m) {
    double* H = sc + S_H;
    double* bh = H + m * m + 2 * m;
    for (int k = 0; k <= m; ++k) bh[k] = 0.0;
    bh[0] = sc[S_BETA];
}
__global__ void k_gmres_backsub(double* sc, int m) {
    double* H = sc + S_H;
    double* bh = H + m * m + 2 * m;
    double* yh = bh + m + 1;
    for (int i = m - 1; i >= 0; --i) {
        double s = bh[i];
        for (int j = i + 1; j < m; ++j) s -= H[i * m + j] * yh[j];
        yh[i] = s / H[i * m + i];
    }
}

// ---------------------------------------------------------------- block SpMV
template <class TA, class TX, class TY>
__global__ __launch_bounds__(256) void k_bspmv(int n, const int32_t* rowPtr, const int32_t* cols, const TA* A, const TX* x,
                                               TY* y) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc[5] = {0, 0, 0, 0, 0};
    for (int p = rowPtr[i]; p < rowPtr[i + 1]; ++p) {
        const TA* B = A + 25 * (size_t)p;
        const TX* xv = x + 5 * (size_t)cols[p];
        double x0 = xv[0], x1 = xv[1], x2 = xv[2], x3 = xv[3], x4 = xv[4];
        for (int r = 0; r < 5; ++r)
            acc[r] += (double)B[5 * r] * x0 + (double)B[5 * r + 1] * x1 + (double)B[5 * r + 2] * x2 +
                      (double)B[5 * r + 3] * x3 + (double)B[5 * r + 4] * x4;
    }
    for (int r = 0; r < 5; ++r) y[5 * (size_t)i + r] = (TY)acc[r];
}

// ---------------------------------------------------------------- sliced-ELL SpMV
// block-CSR -> sliced ELL (64 rows per slice, entry-major inside a block column)
template <class T>
__global__ void k_to_sell(long nnzb, const int64_t* sellOfSlot, const T* A, T* S) {
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < nnzb; j += (long)gridDim.x * blockDim.x) {
        const int64_t base = sellOfSlot[j];
        const T* a = A + 25 * (size_t)j;
        for (int e = 0; e < 25; ++e) S[base + 64 * e] = a[e];
    }
}
// one 64-lane wavefront per slice, lane = row: every value load is 64
// consecutive entries (fully coalesced), x is gathered per block column
template <class TA, class TX, class TY>
__global__ __launch_bounds__(64) void k_sell_spmv(int n, const int64_t* sellPtr, const int32_t* sellWidth,
                                                  const int32_t* sellCol, const TA* S, const TX* x, TY* y) {
    const int sl = blockIdx.x;
    const int lane = threadIdx.x;
    const int row = 64 * sl + lane;
    const int64_t b0 = sellPtr[sl];
    const int w = sellWidth[sl];
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0;
    for (int p = 0; p < w; ++p) {
        const int64_t blk = b0 + 64 * (int64_t)p;
        const int col = sellCol[blk + lane];
        const TX* xv = x + 5 * (size_t)col;
        const double x0 = xv[0], x1 = xv[1], x2 = xv[2], x3 = xv[3], x4 = xv[4];
        const TA* v = S + 25 * blk + lane;
        a0 += (double)v[0] * x0 + (double)v[64] * x1 + (double)v[128] * x2 + (double)v[192] * x3 + (double)v[256] * x4;
        a1 += (double)v[320] * x0 + (double)v[384] * x1 + (double)v[448] * x2 + (double)v[512] * x3 + (double)v[576] * x4;
        a2 += (double)v[640] * x0 + (double)v[704] * x1 + (double)v[768] * x2 + (double)v[832] * x3 + (double)v[896] * x4;
        a3 += (double)v[960] * x0 + (double)v[1024] * x1 + (double)v[1088] * x2 + (double)v[1152] * x3 + (double)v[1216] * x4;
        a4 += (double)v[1280] * x0 + (double)v[1344] * x1 + (double)v[1408] * x2 + (double)v[1472] * x3 + (double)v[1536] * x4;
    }
    if (row < n) {
        TY* yr = y + 5 * (size_t)row;
        yr[0] = (TY)a0; yr[1] = (TY)a1; yr[2] = (TY)a2; yr[3] = (TY)a3; yr[4] = (TY)a4;
    }
}

// ---------------------------------------------------------------- classical Gram-Schmidt
// partial dots of w with NACC consecutive basis vectors V_j0 .. V_j0+NACC-1
// (cnt of them valid).  Each thread walks its grid-stride elements four at a
// time so that 4 (NACC+1) loads are in flight; the accumulation order per
// thread is the plain grid-stride order, so the partials do not depend on it.
template <class T, int NACC>
__global__ __launch_bounds__(RED_BS) void k_multidot_partial(long n, const T* w, const T* V, int j0, int cnt,
                                                             double* part) {
    __shared__ double sh[16];
    double acc[NACC];
    for (int q = 0; q < NACC; ++q) acc[q] = 0.0;
    const long st = (long)gridDim.x * blockDim.x;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const T* Vb = V + (size_t)j0 * n;
    for (; i + 3 * st < n; i += 4 * st) {
        double wi[4];
        T v[4][NACC];
#pragma unroll
        for (int u = 0; u < 4; ++u) wi[u] = (double)w[i + u * st];
#pragma unroll
        for (int q = 0; q < NACC; ++q)
            if (q < cnt)
#pragma unroll
                for (int u = 0; u < 4; ++u) v[u][q] = Vb[(size_t)q * n + i + u * st];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int q = 0; q < NACC; ++q)
                if (q < cnt) acc[q] += wi[u] * (double)v[u][q];
    }
    for (; i < n; i += st) {
        const double wi = (double)w[i];
#pragma unroll
        for (int q = 0; q < NACC; ++q)
            if (q < cnt) acc[q] += wi * (double)Vb[(size_t)q * n + i];
    }
#pragma unroll
    for (int q = 0; q < NACC; ++q) {
        if (q >= cnt) break;
        const double t = block_sum(acc[q], sh);
        if (threadIdx.x == 0) part[(size_t)(j0 + q) * gridDim.x + blockIdx.x] = t;
    }
}
// the same for an FP32 basis whose length is a multiple of four: each thread
// reads four consecutive entries per 16-byte load (twice the bytes in flight
// of the scalar form); per thread the entries are summed in grid-stride order
// of their float4 groups
template <int NACC>
__global__ __launch_bounds__(RED_BS) void k_multidot_partial4(long n, const float* w, const float* V, int j0, int cnt,
                                                              double* part) {
    __shared__ double sh[16];
    double acc[NACC];
    for (int q = 0; q < NACC; ++q) acc[q] = 0.0;
    const long n4 = n / 4;
    const long st = (long)gridDim.x * blockDim.x;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const float4* w4 = reinterpret_cast<const float4*>(w);
    const float4* Vb = reinterpret_cast<const float4*>(V + (size_t)j0 * n);
    for (; i + st < n4; i += 2 * st) {
        float4 wi[2], v[2][NACC];
#pragma unroll
        for (int u = 0; u < 2; ++u) wi[u] = w4[i + u * st];
#pragma unroll
        for (int q = 0; q < NACC; ++q)
            if (q < cnt)
#pragma unroll
                for (int u = 0; u < 2; ++u) v[u][q] = Vb[(size_t)q * n4 + i + u * st];
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int q = 0; q < NACC; ++q)
                if (q < cnt)
                    acc[q] += (double)wi[u].x * (double)v[u][q].x + (double)wi[u].y * (double)v[u][q].y +
                              (double)wi[u].z * (double)v[u][q].z + (double)wi[u].w * (double)v[u][q].w;
    }
    for (; i < n4; i += st) {
        const float4 wi = w4[i];
#pragma unroll
        for (int q = 0; q < NACC; ++q)
            if (q < cnt) {
                const float4 v = Vb[(size_t)q * n4 + i];
                acc[q] += (double)wi.x * (double)v.x + (double)wi.y * (double)v.y + (double)wi.z * (double)v.z +
                          (double)wi.w * (double)v.w;
            }
    }
#pragma unroll
    for (int q = 0; q < NACC; ++q) {
        if (q >= cnt) break;
        const double t = block_sum(acc[q], sh);
        if (threadIdx.x == 0) part[(size_t)(j0 + q) * gridDim.x + blockIdx.x] = t;
    }
}

// H[j][col] = sum of the G partials of dot j; block j reduces dot j
__global__ __launch_bounds__(RED_BS) void k_multidot_final(const double* part, int G, int cnt, double* H, int m, int col) {
    __shared__ double sh[16];
    const int j = blockIdx.x;
    if (j >= cnt) return;
    const double t = reduce_partials(part + (size_t)j * G, G, sh);
    if (threadIdx.x == 0) H[j * m + col] = t;
}
// w -= sum_j H[j][col] V_j, and the partial sums of |w|^2; four grid-stride
// elements per pass (same per-element arithmetic and partial order as one)
template <class T>
__global__ __launch_bounds__(RED_BS) void k_cgs_update(long n, T* w, const T* V, const double* H, int m, int col, int cnt,
                                                       double* part) {
    constexpr int HC = RED_BS;
    __shared__ double sh[16];
    __shared__ double hc[HC];
    for (int t = threadIdx.x; t < cnt && t < HC; t += blockDim.x) hc[t] = H[t * m + col];
    __syncthreads();
    double s = 0.0;
    const long st = (long)gridDim.x * blockDim.x;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * st < n; i += 4 * st) {
        double wi[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) wi[u] = (double)w[i + u * st];
        for (int j = 0; j < cnt; ++j) {
            const double h = j < HC ? hc[j] : H[j * m + col];
            T v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = V[(size_t)j * n + i + u * st];
#pragma unroll
            for (int u = 0; u < 4; ++u) wi[u] -= h * (double)v[u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const T wt = (T)wi[u];
            w[i + u * st] = wt;
            s += (double)wt * (double)wt;
        }
    }
    for (; i < n; i += st) {
        double wi = (double)w[i];
        for (int j = 0; j < cnt; ++j) wi -= (j < HC ? hc[j] : H[j * m + col]) * (double)V[(size_t)j * n + i];
        const T wt = (T)wi;
        w[i] = wt;
        s += (double)wt * (double)wt;
    }
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// the same for an FP32 basis whose length is a multiple of four: 16-byte
// loads, two float4 groups per thread and two basis vectors per pass; each
// entry is updated in the same order (j ascending) as above, so w is
// unchanged; the |w|^2 partials are summed per float4 group
__global__ __launch_bounds__(RED_BS) void k_cgs_update4(long n, float* w, const float* V, const double* H, int m, int col,
                                                        int cnt, double* part) {
    constexpr int HC = RED_BS;
    __shared__ double sh[16];
    __shared__ double hc[HC];
    for (int t = threadIdx.x; t < cnt && t < HC; t += blockDim.x) hc[t] = H[t * m + col];
    __syncthreads();
    double s = 0.0;
    const long n4 = n / 4;
    const long st = (long)gridDim.x * blockDim.x;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float4* w4 = reinterpret_cast<float4*>(w);
    const float4* V4 = reinterpret_cast<const float4*>(V);
    auto upd = [](double* a, double h, const float4& v) {
        a[0] -= h * (double)v.x; a[1] -= h * (double)v.y; a[2] -= h * (double)v.z; a[3] -= h * (double)v.w;
    };
    for (; i < n4; i += 2 * st) {
        const bool two = i + st < n4;
        double a[2][4];
        {
            const float4 w0 = w4[i];
            a[0][0] = w0.x; a[0][1] = w0.y; a[0][2] = w0.z; a[0][3] = w0.w;
            const float4 w1 = two ? w4[i + st] : make_float4(0.f, 0.f, 0.f, 0.f);
            a[1][0] = w1.x; a[1][1] = w1.y; a[1][2] = w1.z; a[1][3] = w1.w;
        }
        int j = 0;
        for (; j + 1 < cnt; j += 2) {
            const double h0 = j < HC ? hc[j] : H[j * m + col];
            const double h1 = j + 1 < HC ? hc[j + 1] : H[(j + 1) * m + col];
            const float4* Vj = V4 + (size_t)j * n4;
            const float4 v00 = Vj[i], v10 = Vj[n4 + i];
            const float4 v01 = two ? Vj[i + st] : make_float4(0.f, 0.f, 0.f, 0.f);
            const float4 v11 = two ? Vj[n4 + i + st] : make_float4(0.f, 0.f, 0.f, 0.f);
            upd(a[0], h0, v00); upd(a[0], h1, v10);
            upd(a[1], h0, v01); upd(a[1], h1, v11);
        }
        if (j < cnt) {
            const double h0 = j < HC ? hc[j] : H[j * m + col];
            const float4* Vj = V4 + (size_t)j * n4;
            upd(a[0], h0, Vj[i]);
            if (two) upd(a[1], h0, Vj[i + st]);
        }
        for (int u = 0; u < (two ? 2 : 1); ++u) {
            const float4 o = make_float4((float)a[u][0], (float)a[u][1], (float)a[u][2], (float)a[u][3]);
            w4[i + u * st] = o;
            s += (double)o.x * (double)o.x + (double)o.y * (double)o.y + (double)o.z * (double)o.z +
                 (double)o.w * (double)o.w;
        }
    }
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// ---------------------------------------------------------------- block ILU(0)
template <class T>
__device__ inline void mm5(const T* A, const T* B, T* C) {
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) {
            T s = 0;
            for (int k = 0; k < 5; ++k) s += A[5 * i + k] * B[5 * k + j];
            C[5 * i + j] = s;
        }
}
// Gauss-Jordan inverse with partial pivoting
template <class T>
__device__ inline bool inv5(const T* A, T* Ainv) {
    T a[5][10];
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) { a[i][j] = A[5 * i + j]; a[i][5 + j] = (i == j) ? T(1) : T(0); }
    for (int col = 0; col < 5; ++col) {
        int piv = col;
        for (int r = col + 1; r < 5; ++r) if (fabs((double)a[r][col]) > fabs((double)a[piv][col])) piv = r;
        if (a[piv][col] == T(0)) return false;
        if (piv != col) for (int j = 0; j < 10; ++j) { T t = a[piv][j]; a[piv][j] = a[col][j]; a[col][j] = t; }
        T d = T(1) / a[col][col];
        for (int j = 0; j < 10; ++j) a[col][j] *= d;
        for (int r = 0; r < 5; ++r) {
            if (r == col) continue;
            T f = a[r][col];
            if (f == T(0)) continue;
            for (int j = 0; j < 10; ++j) a[r][j] -= f * a[col][j];
        }
    }
    for (int i = 0; i < 5; ++i) for (int j = 0; j < 5; ++j) Ainv[5 * i + j] = a[i][5 + j];
    return true;
}

// factor row i (IKJ, row i reads only rows of earlier levels)
template <class T>
__device__ inline void ilu_factor_row(int i, const int32_t* rowPtr, const int32_t* cols, const int32_t* diagPos, T* LU,
                                      T* Dinv, int* fail) {
    const int rb = rowPtr[i], re = rowPtr[i + 1], di = diagPos[i];
    for (int p = rb; p < di; ++p) {
        const int k = cols[p];
        T L[25];
        mm5(LU + 25 * (size_t)p, Dinv + 25 * (size_t)k, L);
        for (int e = 0; e < 25; ++e) LU[25 * (size_t)p + e] = L[e];
        // A_ij -= L_ik U_kj for j > k present in row i (both rows sorted)
        int qi = p + 1;
        for (int q = diagPos[k] + 1; q < rowPtr[k + 1]; ++q) {
            const int j = cols[q];
            while (qi < re && cols[qi] < j) ++qi;
            if (qi >= re) break;
            if (cols[qi] != j) continue;
            T P[25];
            mm5(L, LU + 25 * (size_t)q, P);
            for (int e = 0; e < 25; ++e) LU[25 * (size_t)qi + e] -= P[e];
        }
    }
    if (!inv5(LU + 25 * (size_t)di, Dinv + 25 * (size_t)i)) atomicExch(fail, 1);
}

// factor the rows of one level
template <class T>
__global__ __launch_bounds__(64) void k_ilu_factor(int cnt, const int32_t* rows, const int32_t* rowPtr, const int32_t* cols,
                                                    const int32_t* diagPos, T* LU, T* Dinv, int* fail) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    ilu_factor_row(rows[t], rowPtr, cols, diagPos, LU, Dinv, fail);
}

// forward sweep of row i: y_i = r_i - sum_{k<i} L_ik y_k
template <class TM, class TV>
__device__ inline void ilu_fwd_row(int i, const int32_t* rowPtr, const int32_t* cols, const int32_t* diagPos,
                                   const TM* LU, const TV* r, TV* y) {
    TV s[5];
    for (int k = 0; k < 5; ++k) s[k] = r[5 * (size_t)i + k];
    for (int p = rowPtr[i]; p < diagPos[i]; ++p) {
        const TM* B = LU + 25 * (size_t)p;
        const TV* yv = y + 5 * (size_t)cols[p];
        TV y0 = yv[0], y1 = yv[1], y2 = yv[2], y3 = yv[3], y4 = yv[4];
        for (int rr = 0; rr < 5; ++rr)
            s[rr] -= (TV)(B[5 * rr] * y0 + B[5 * rr + 1] * y1 + B[5 * rr + 2] * y2 + B[5 * rr + 3] * y3 + B[5 * rr + 4] * y4);
    }
    for (int k = 0; k < 5; ++k) y[5 * (size_t)i + k] = s[k];
}
template <class TM, class TV>
__global__ __launch_bounds__(64) void k_ilu_fwd(int cnt, const int32_t* rows, const int32_t* rowPtr, const int32_t* cols,
                                                const int32_t* diagPos, const TM* LU, const TV* r, TV* y) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    ilu_fwd_row(rows[t], rowPtr, cols, diagPos, LU, r, y);
}
// backward sweep of row i: x_i = Dinv_i (y_i - sum_{j>i} U_ij x_j)
template <class TM, class TV>
__device__ inline void ilu_bwd_row(int i, const int32_t* rowPtr, const int32_t* cols, const int32_t* diagPos,
                                   const TM* LU, const TM* Dinv, const TV* y, TV* x) {
    TV s[5];
    for (int k = 0; k < 5; ++k) s[k] = y[5 * (size_t)i + k];
    for (int p = diagPos[i] + 1; p < rowPtr[i + 1]; ++p) {
        const TM* B = LU + 25 * (size_t)p;
        const TV* xv = x + 5 * (size_t)cols[p];
        TV x0 = xv[0], x1 = xv[1], x2 = xv[2], x3 = xv[3], x4 = xv[4];
        for (int rr = 0; rr < 5; ++rr)
            s[rr] -= (TV)(B[5 * rr] * x0 + B[5 * rr + 1] * x1 + B[5 * rr + 2] * x2 + B[5 * rr + 3] * x3 + B[5 * rr + 4] * x4);
    }
    const TM* D = Dinv + 25 * (size_t)i;
    for (int rr = 0; rr < 5; ++rr)
        x[5 * (size_t)i + rr] = (TV)(D[5 * rr] * s[0] + D[5 * rr + 1] * s[1] + D[5 * rr + 2] * s[2] + D[5 * rr + 3] * s[3] +
                                     D[5 * rr + 4] * s[4]);
}
template <class TM, class TV>
__global__ __launch_bounds__(64) void k_ilu_bwd(int cnt, const int32_t* rows, const int32_t* rowPtr, const int32_t* cols,
                                                const int32_t* diagPos, const TM* LU, const TM* Dinv, const TV* y, TV* x) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    ilu_bwd_row(rows[t], rowPtr, cols, diagPos, LU, Dinv, y, x);
}

// A run of consecutive small levels [L0, L1) in one launch of one workgroup:
// the levels one after the other, a barrier between them (the rows a level
// reads were written by this workgroup before the barrier; the workgroup-
// scope release/acquire of the barrier makes them visible).  Level launches
// cost several microseconds each, far more than a small level's work.
constexpr int ILU_SEG_THREADS = 64;
constexpr int ILU_SEG_FACTOR_THREADS = 64;
template <class T>
__global__ __launch_bounds__(ILU_SEG_FACTOR_THREADS) void k_ilu_factor_seg(int L0, int L1, const int32_t* levPtr,
                                                                     const int32_t* rows, const int32_t* rowPtr,
                                                                     const int32_t* cols, const int32_t* diagPos, T* LU,
                                                                     T* Dinv, int* fail) {
    for (int L = L0; L < L1; ++L) {
        const int b = levPtr[L], e = levPtr[L + 1];
        for (int t = b + (int)threadIdx.x; t < e; t += ILU_SEG_FACTOR_THREADS)
            ilu_factor_row(rows[t], rowPtr, cols, diagPos, LU, Dinv, fail);
        __syncthreads();
    }
}
template <class TM, class TV>
__global__ __launch_bounds__(ILU_SEG_THREADS) void k_ilu_fwd_seg(int L0, int L1, const int32_t* levPtr,
                                                                  const int32_t* rows, const int32_t* rowPtr,
                                                                  const int32_t* cols, const int32_t* diagPos,
                                                                  const TM* LU, const TV* r, TV* y) {
    for (int L = L0; L < L1; ++L) {
        const int b = levPtr[L], e = levPtr[L + 1];
        for (int t = b + (int)threadIdx.x; t < e; t += ILU_SEG_THREADS)
            ilu_fwd_row(rows[t], rowPtr, cols, diagPos, LU, r, y);
        __syncthreads();
    }
}
template <class TM, class TV>
__global__ __launch_bounds__(ILU_SEG_THREADS) void k_ilu_bwd_seg(int L0, int L1, const int32_t* levPtr,
                                                                  const int32_t* rows, const int32_t* rowPtr,
                                                                  const int32_t* cols, const int32_t* diagPos,
                                                                  const TM* LU, const TM* Dinv, const TV* y, TV* x) {
    for (int L = L0; L < L1; ++L) {
        const int b = levPtr[L], e = levPtr[L + 1];
        for (int t = b + (int)threadIdx.x; t < e; t += ILU_SEG_THREADS)
            ilu_bwd_row(rows[t], rowPtr, cols, diagPos, LU, Dinv, y, x);
        __syncthreads();
    }
}
// ---------------------------------------------------------------- partitioned block ILU(0)
// One workgroup per partition walks the partition's local levels; entries
// whose column lies in another partition are dropped (block-Jacobi).
// The factor: one thread per block row of a level (the rows of a level only
// read rows of earlier levels); each thread forms L_ik = A_ik Dinv_k, the
// updates A_ij -= L_ik U_kj and the inverse of its diagonal block.  (A
// five-lanes-per-row form of this kernel measured 1.7x slower on MI355X: the
// diagonal inversion then runs on one lane in five.)
constexpr int PILU_THREADS = 320;  // the apply: 5 wavefronts x 12 rows
constexpr int PILU_FACTOR_THREADS = 256;
template <class T>
__global__ __launch_bounds__(PILU_FACTOR_THREADS) void k_pilu_factor(const int32_t* partLevStart,
                                                                     const int32_t* partLevPtr, const int32_t* partRows,
                                                                     const int32_t* rowPtr, const int32_t* cols,
                                                                     const int32_t* diagPos, const int32_t* localCol,
                                                                     T* LU, T* Dinv, int* fail) {
    const int pt = blockIdx.x;
    const int l0 = partLevStart[pt], l1 = partLevStart[pt + 1];
    for (int L = l0; L + 1 < l1; ++L) {
        const int s = partLevPtr[L], e = partLevPtr[L + 1];
        for (int t = s + (int)threadIdx.x; t < e; t += blockDim.x) {
            const int i = partRows[t];
            const int re = rowPtr[i + 1], di = diagPos[i];
            for (int p = rowPtr[i]; p < di; ++p) {
                if (localCol[p] < 0) continue;
                const int k = cols[p];
                T Lb[25];
                mm5(LU + 25 * (size_t)p, Dinv + 25 * (size_t)k, Lb);
                for (int q = 0; q < 25; ++q) LU[25 * (size_t)p + q] = Lb[q];
                int qi = p + 1;
                for (int q = diagPos[k] + 1; q < rowPtr[k + 1]; ++q) {
                    if (localCol[q] < 0) continue;
                    const int j = cols[q];
                    while (qi < re && cols[qi] < j) ++qi;
                    if (qi >= re) break;
                    if (cols[qi] != j) continue;
                    T P[25];
                    mm5(Lb, LU + 25 * (size_t)q, P);
                    for (int q2 = 0; q2 < 25; ++q2) LU[25 * (size_t)qi + q2] -= P[q2];
                }
            }
            if (!inv5(LU + 25 * (size_t)di, Dinv + 25 * (size_t)i)) atomicExch(fail, 1);
        }
        __syncthreads();
    }
}

// Scatter the factor into the apply streams (layout of PiluStream): one
// thread per (stream block, block row r).  Backward blocks other than the
// row's Dinv are stored as Dinv U, formed in FP64.
template <class T, bool BWD>
__global__ __launch_bounds__(256) void k_pilu_pack(long nBlk, const int32_t* src, const int32_t* rowc,
                                                   const int32_t* jn, const T* LU, const T* Dinv, T* val) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 5 * nBlk) return;
    const long b = e / 5;
    const int r = (int)(e - 5 * b);
    const int code = jn[b];
    const int j = code & 255, nk = code >> 8;
    T* dst = val + 25 * (size_t)(b - j) + 5 * j + r;
    const int sl = src[b];
    if (!BWD || sl < 0) {
        const T* B = BWD ? Dinv + 25 * (size_t)(-sl - 1) + 5 * r : LU + 25 * (size_t)sl + 5 * r;
        for (int c = 0; c < 5; ++c) dst[(size_t)c * 5 * nk] = B[c];
    } else {
        const T* D = Dinv + 25 * (size_t)rowc[b] + 5 * r;
        const T* U = LU + 25 * (size_t)sl;
        for (int c = 0; c < 5; ++c) {
            double a = 0.0;
            for (int q = 0; q < 5; ++q) a += (double)D[q] * (double)U[5 * q + c];
            dst[(size_t)c * 5 * nk] = (T)a;
        }
    }
}

// One sweep over the sub-levels [s0, s1) of a partition (y in LDS).  Lane
// layout: 12 rows of five lanes per wavefront (lanes 60..63 idle), so the
// lanes of a row share a wavefront and the backward row's read of its own y
// (for Dinv y) precedes its write in every lane.  All loads of a sub-level
// depend only on its descriptor, so they are in flight together.
template <class TM, class TV, int KDU, bool BWD>
__device__ inline void pilu_sweep(int s0, int s1, const int32_t* __restrict__ sub, const int32_t* __restrict__ ord,
                                  const int32_t* __restrict__ lcs, const TM* __restrict__ val, TV* y) {
    const int l = threadIdx.x & 63, g = l / 5, r = l - 5 * g;
    const int j = 12 * (int)(threadIdx.x >> 6) + g;
    const bool laneOn = g < 12;
    for (int s = s0; s < s1; ++s) {
        const int32_t* d = sub + (size_t)kPiluSubW * s;
        const int rowOff = d[0], nRows = d[1], kmax = d[2];
        if (laneOn && j < nRows) {
            const int t = ord[rowOff + j];
            TM v[KDU][5];
            int lc[KDU];
#pragma unroll
            for (int k = 0; k < KDU; ++k) {
                if (k < kmax && j < d[4 + kPiluKD + k]) {
                    const int bk = d[4 + k], nk = d[4 + kPiluKD + k];
                    lc[k] = lcs[bk + j];
                    const TM* vb = val + 25 * (size_t)bk + 5 * j + r;
#pragma unroll
                    for (int c = 0; c < 5; ++c) v[k][c] = vb[(size_t)c * 5 * nk];
                }
            }
            TV acc = BWD ? TV(0) : y[5 * t + r];
#pragma unroll
            for (int k = 0; k < KDU; ++k) {
                if (k < kmax && j < d[4 + kPiluKD + k]) {
                    const TV* yv = y + 5 * lc[k];
                    const TV dot = (TV)(v[k][0] * yv[0] + v[k][1] * yv[1] + v[k][2] * yv[2] + v[k][3] * yv[3] +
                                        v[k][4] * yv[4]);
                    if (BWD && k == 0) acc += dot;
                    else acc -= dot;
                }
            }
            y[5 * t + r] = acc;
        }
        __syncthreads();
    }
}

// x = M^-1 r for the partitioned factor from its streams; one workgroup per
// partition, the partition's vector in LDS for both sweeps
template <class TM, class TV, int KDU>
__global__ __launch_bounds__(256) void k_pilu_stream(const int32_t* __restrict__ partRowStart,
                                                     const int32_t* __restrict__ partRows,
                                                     const int32_t* __restrict__ fStart, const int32_t* __restrict__ fSub,
                                                     const int32_t* __restrict__ fOrd, const int32_t* __restrict__ fLc,
                                                     const TM* __restrict__ fVal, const int32_t* __restrict__ bStart,
                                                     const int32_t* __restrict__ bSub, const int32_t* __restrict__ bOrd,
                                                     const int32_t* __restrict__ bLc, const TM* __restrict__ bVal,
                                                     const TV* __restrict__ r, TV* __restrict__ x) {
    extern __shared__ unsigned char smem_raw[];
    TV* y = reinterpret_cast<TV*>(smem_raw);
    const int pt = blockIdx.x;
    const int rs = partRowStart[pt], re = partRowStart[pt + 1];
    const int nl5 = (re - rs) * 5;
    for (int t = threadIdx.x; t < nl5; t += blockDim.x) y[t] = r[5 * (size_t)partRows[rs + t / 5] + t % 5];
    __syncthreads();
    pilu_sweep<TM, TV, KDU, false>(fStart[pt], fStart[pt + 1], fSub, fOrd, fLc, fVal, y);
    pilu_sweep<TM, TV, KDU, true>(bStart[pt], bStart[pt + 1], bSub, bOrd, bLc, bVal, y);
    for (int t = threadIdx.x; t < nl5; t += blockDim.x) x[5 * (size_t)partRows[rs + t / 5] + t % 5] = y[t];
}

// The same apply with one thread per block row: measured faster than the
// five-lane form for the FP32 factors (the five-lane form wins in FP64,
// where a row's 25 block entries span four cache lines).
template <class TM, class TV>
__global__ __launch_bounds__(256) void k_pilu_apply_rows(const int32_t* partRowStart, const int32_t* partLevStart,
                                                    const int32_t* partLevPtr, const int32_t* partRows,
                                                    const int32_t* rowPtr, const int32_t* cols, const int32_t* diagPos,
                                                    const int32_t* localCol, const TM* LU, const TM* Dinv, const TV* r,
                                                    TV* x) {
    extern __shared__ unsigned char smem_raw[];
    TV* y = reinterpret_cast<TV*>(smem_raw);
    const int pt = blockIdx.x;
    const int rs = partRowStart[pt], re = partRowStart[pt + 1];
    const int nl5 = (re - rs) * 5;
    for (int t = threadIdx.x; t < nl5; t += blockDim.x) y[t] = r[5 * (size_t)partRows[rs + t / 5] + t % 5];
    __syncthreads();
    const int l0 = partLevStart[pt], l1 = partLevStart[pt + 1];
    for (int L = l0; L + 1 < l1; ++L) {
        const int s = partLevPtr[L], e = partLevPtr[L + 1];
        for (int t = s + (int)threadIdx.x; t < e; t += blockDim.x) {
            const int i = partRows[t];
            TV* yi = y + 5 * (t - rs);
            TV a0 = yi[0], a1 = yi[1], a2 = yi[2], a3 = yi[3], a4 = yi[4];
            for (int p = rowPtr[i]; p < diagPos[i]; ++p) {
                const int lc = localCol[p];
                if (lc < 0) continue;
                const TM* B = LU + 25 * (size_t)p;
                const TV* yv = y + 5 * lc;
                const TV y0 = yv[0], y1 = yv[1], y2 = yv[2], y3 = yv[3], y4 = yv[4];
                a0 -= (TV)(B[0] * y0 + B[1] * y1 + B[2] * y2 + B[3] * y3 + B[4] * y4);
                a1 -= (TV)(B[5] * y0 + B[6] * y1 + B[7] * y2 + B[8] * y3 + B[9] * y4);
                a2 -= (TV)(B[10] * y0 + B[11] * y1 + B[12] * y2 + B[13] * y3 + B[14] * y4);
                a3 -= (TV)(B[15] * y0 + B[16] * y1 + B[17] * y2 + B[18] * y3 + B[19] * y4);
                a4 -= (TV)(B[20] * y0 + B[21] * y1 + B[22] * y2 + B[23] * y3 + B[24] * y4);
            }
            yi[0] = a0; yi[1] = a1; yi[2] = a2; yi[3] = a3; yi[4] = a4;
        }
        __syncthreads();
    }
    for (int L = l1 - 2; L >= l0; --L) {
        const int s = partLevPtr[L], e = partLevPtr[L + 1];
        for (int t = s + (int)threadIdx.x; t < e; t += blockDim.x) {
            const int i = partRows[t];
            TV* yi = y + 5 * (t - rs);
            TV a0 = yi[0], a1 = yi[1], a2 = yi[2], a3 = yi[3], a4 = yi[4];
            for (int p = diagPos[i] + 1; p < rowPtr[i + 1]; ++p) {
                const int lc = localCol[p];
                if (lc < 0) continue;
                const TM* B = LU + 25 * (size_t)p;
                const TV* yv = y + 5 * lc;
                const TV y0 = yv[0], y1 = yv[1], y2 = yv[2], y3 = yv[3], y4 = yv[4];
                a0 -= (TV)(B[0] * y0 + B[1] * y1 + B[2] * y2 + B[3] * y3 + B[4] * y4);
                a1 -= (TV)(B[5] * y0 + B[6] * y1 + B[7] * y2 + B[8] * y3 + B[9] * y4);
                a2 -= (TV)(B[10] * y0 + B[11] * y1 + B[12] * y2 + B[13] * y3 + B[14] * y4);
                a3 -= (TV)(B[15] * y0 + B[16] * y1 + B[17] * y2 + B[18] * y3 + B[19] * y4);
                a4 -= (TV)(B[20] * y0 + B[21] * y1 + B[22] * y2 + B[23] * y3 + B[24] * y4);
            }
            const TM* D = Dinv + 25 * (size_t)i;
            yi[0] = (TV)(D[0] * a0 + D[1] * a1 + D[2] * a2 + D[3] * a3 + D[4] * a4);
            yi[1] = (TV)(D[5] * a0 + D[6] * a1 + D[7] * a2 + D[8] * a3 + D[9] * a4);
            yi[2] = (TV)(D[10] * a0 + D[11] * a1 + D[12] * a2 + D[13] * a3 + D[14] * a4);
            yi[3] = (TV)(D[15] * a0 + D[16] * a1 + D[17] * a2 + D[18] * a3 + D[19] * a4);
            yi[4] = (TV)(D[20] * a0 + D[21] * a1 + D[22] * a2 + D[23] * a3 + D[24] * a4);
        }
        __syncthreads();
    }
    for (int t = threadIdx.x; t < nl5; t += blockDim.x) x[5 * (size_t)partRows[rs + t / 5] + t % 5] = y[t];
}


// x = M^-1 r for the partitioned factor; the partition's vector stays in LDS
// for the forward and the backward sweep (reverse local levels).  Five lanes
// per block row, lane = component r: each lane reads row r of every 5x5 block
// (the five lanes of a row read the block's 25 contiguous entries together)
// and a 64-lane wavefront carries 12 rows; the Dinv product gathers the five
// components of a row with wavefront shuffles.  Every lane of the block runs
// the same number of passes, so the shuffles and barriers stay uniform.
template <class TM, class TV>
__global__ __launch_bounds__(PILU_THREADS) void k_pilu_apply(const int32_t* partRowStart, const int32_t* partLevStart,
                                                             const int32_t* partLevPtr, const int32_t* partRows,
                                                             const int32_t* rowPtr, const int32_t* cols,
                                                             const int32_t* diagPos, const int32_t* localCol,
                                                             const TM* LU, const TM* Dinv, const TV* r, TV* x) {
    extern __shared__ unsigned char smem_raw[];
    TV* y = reinterpret_cast<TV*>(smem_raw);
    const int pt = blockIdx.x;
    const int rs = partRowStart[pt], re = partRowStart[pt + 1];
    const int nl5 = (re - rs) * 5;
    for (int t = threadIdx.x; t < nl5; t += blockDim.x) y[t] = r[5 * (size_t)partRows[rs + t / 5] + t % 5];
    __syncthreads();
    const int wl = threadIdx.x & 63;                 // lane in the wavefront
    const int comp = wl % 5;                         // block row handled by this lane
    const int rowInWave = wl / 5;                    // 0..11 (12 = idle lanes 60..63)
    const bool laneOn = wl < 60;
    const int slot = (threadIdx.x >> 6) * 12 + rowInWave;
    const int rowsPerPass = (blockDim.x >> 6) * 12;
    const int base5 = wl - comp;                     // first lane of this row in the wavefront
    const int l0 = partLevStart[pt], l1 = partLevStart[pt + 1];
    for (int L = l0; L + 1 < l1; ++L) {
        const int s = partLevPtr[L], e = partLevPtr[L + 1];
        for (int b = s; b < e; b += rowsPerPass) {
            const int t = b + slot;
            if (laneOn && t < e) {
                const int i = partRows[t];
                TV a = y[5 * (t - rs) + comp];
                for (int p = rowPtr[i]; p < diagPos[i]; ++p) {
                    const int lc = localCol[p];
                    if (lc < 0) continue;
                    const TM* B = LU + 25 * (size_t)p + 5 * comp;
                    const TV* yv = y + 5 * lc;
                    a -= (TV)(B[0] * yv[0] + B[1] * yv[1] + B[2] * yv[2] + B[3] * yv[3] + B[4] * yv[4]);
                }
                y[5 * (t - rs) + comp] = a;
            }
        }
        __syncthreads();
    }
    for (int L = l1 - 2; L >= l0; --L) {
        const int s = partLevPtr[L], e = partLevPtr[L + 1];
        for (int b = s; b < e; b += rowsPerPass) {
            const int t = b + slot;
            const bool on = laneOn && t < e;
            TV a = 0;
            int i = 0;
            if (on) {
                i = partRows[t];
                a = y[5 * (t - rs) + comp];
                for (int p = diagPos[i] + 1; p < rowPtr[i + 1]; ++p) {
                    const int lc = localCol[p];
                    if (lc < 0) continue;
                    const TM* B = LU + 25 * (size_t)p + 5 * comp;
                    const TV* yv = y + 5 * lc;
                    a -= (TV)(B[0] * yv[0] + B[1] * yv[1] + B[2] * yv[2] + B[3] * yv[3] + B[4] * yv[4]);
                }
            }
            // the five components of the row, from the row's lanes
            const TV a0 = __shfl(a, base5 + 0, 64), a1 = __shfl(a, base5 + 1, 64), a2 = __shfl(a, base5 + 2, 64),
                     a3 = __shfl(a, base5 + 3, 64), a4 = __shfl(a, base5 + 4, 64);
            if (on) {
                const TM* D = Dinv + 25 * (size_t)i + 5 * comp;
                y[5 * (t - rs) + comp] = (TV)(D[0] * a0 + D[1] * a1 + D[2] * a2 + D[3] * a3 + D[4] * a4);
            }
        }
        __syncthreads();
    }
    for (int t = threadIdx.x; t < nl5; t += blockDim.x) x[5 * (size_t)partRows[rs + t / 5] + t % 5] = y[t];
}

// diagonal (block-Jacobi) preconditioner
template <class TM, class TV>
__global__ void k_diag_apply(int n, const TM* Dinv, const TV* r, TV* x) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const TM* D = Dinv + 25 * (size_t)i;
    for (int rr = 0; rr < 5; ++rr) {
        double s = 0;
        for (int k = 0; k < 5; ++k) s += (double)D[5 * rr + k] * (double)r[5 * (size_t)i + k];
        x[5 * (size_t)i + rr] = (TV)s;
    }
}
template <class T>
__global__ void k_diag_inv(int n, const int32_t* diagPos, const T* A, T* Dinv, int* fail) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!inv5(A + 25 * (size_t)diagPos[i], Dinv + 25 * (size_t)i)) atomicExch(fail, 1);
}

// BiCGStab vector updates
__global__ void k_bicg_p(long n, double* p, const double* r, const double* v, const double* sc) {
    const double beta = sc[0], omega = sc[1];
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        p[i] = r[i] + beta * (p[i] - omega * v[i]);
}
__global__ void k_axpy_s(long n, double* y, const double* x, const double* s, double sign) {
    const double a = sign * s[0];
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] += a * x[i];
}
__global__ void k_bicg_x(long n, double* x, const double* ph, const double* th, const double* sc) {
    const double alpha = sc[0], omega = sc[1];
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        x[i] += alpha * ph[i] + omega * th[i];
}
// BiCGStab scalar updates, single thread: mode 0: rho/beta, 1: alpha, 2: omega
__global__ void k_bicg_scalars(double* sc, int mode) {
    // sc: [0]=rho [1]=rhoOld [2]=alpha [3]=omega [4]=beta [5]=dot1 [6]=dot2 [8]=(beta,omega) [10]=(alpha,omega)
    if (mode == 0) {
        sc[1] = sc[0];
        sc[0] = sc[5];
        sc[4] = sc[0] / sc[1] * (sc[2] / sc[3]);
        sc[8] = sc[4]; sc[9] = sc[3];
    } else if (mode == 1) {
        sc[2] = sc[0] / sc[5];
        sc[12] = sc[2];
    } else {
        sc[3] = sc[5] / sc[6];
        sc[10] = sc[2]; sc[11] = sc[3];
        sc[13] = sc[3];
    }
}

}  // namespace

// ======================================================================= host side

static int grid_for(long n) { return std::max(1, std::min(RED_MAXG, blocks_for(n, RED_BS))); }

static void ensure_work(Case& c) {
    Mesh& m = *c.mesh;
    const bool mixed = c.st.precision == RDST_PREC_MIXED;
    SolverWork* w = c.work;
    if (w && w->n == m.nC && w->nDirs == c.st.n_directions && w->mixed == mixed) return;
    if (w) delete w;
    w = new SolverWork();
    c.work = w;
    w->n = m.nC;
    w->n5 = 5 * m.nC;
    w->nDirs = std::max(1, c.st.n_directions);
    w->mixed = mixed;
    w->G = grid_for(w->n5);
    const size_t n5 = w->n5, nn = (size_t)m.nnzb * 25;
    if (mixed) {
        w->LU32.alloc(nn);
        w->Dinv32.alloc(25 * (size_t)m.nC);
        w->V32.alloc((size_t)(w->nDirs + 1) * n5);
        w->w32.alloc(n5);
        w->r32.alloc(n5);
        w->t32.alloc(n5);
    } else {
        w->LU.alloc(nn);
        w->Dinv.alloc(25 * (size_t)m.nC);
        w->V.alloc((size_t)(w->nDirs + 1) * n5);
    }
    w->b.alloc(n5);
    w->r.alloc(n5);
    w->wA.alloc(n5);
    w->tmp.alloc(n5);
    w->x0.alloc(n5);
    w->dW.alloc(n5);
    const int m2 = w->nDirs;
    w->scal.alloc(S_H + m2 * m2 + 4 * m2 + 8);
    w->hscal.assign(w->scal.n, 0.0);
    w->partA.alloc(5 * RED_MAXG);
    w->partB.alloc(5 * RED_MAXG);
    w->multi.alloc((size_t)(w->nDirs + 1) * RED_MAXG);
}

void solver_free(Case& c) {
    delete c.work;
    c.work = nullptr;
}

namespace {

struct Ctx {
    Case& c;
    Mesh& m;
    SolverWork& w;
    hipStream_t s;
    int G;
    bool right = false;  // GMRES preconditioned from the right (minimises the true residual)
};

// y = A x: the sliced-ELL copy of A when one was built for this step
template <class TA, class TX, class TY>
void spmv(Ctx& k, const TA* A, const TX* x, TY* y) {
    const void* Av = (const void*)A;
    if (Av && (Av == (const void*)k.c.adjW.p || Av == (const void*)k.c.adjWT.p || Av == (const void*)k.c.adjWT32.p)) {
        // the viscous adjoint's operators on the second-neighbour pattern
        hipLaunchKernelGGL((k_bspmv<TA, TX, TY>), dim3(blocks_for(k.m.nC, 256)), dim3(256), 0, k.s, k.m.nC,
                           k.m.d_wideRowPtr.p, k.m.d_wideCols.p, A, x, y);
        return;
    }
    const TA* S = nullptr;
    if ((const void*)A == (const void*)k.c.A64.p) S = (const TA*)k.c.S64.p;
    if ((const void*)A == (const void*)k.c.A32.p) S = (const TA*)k.c.S32.p;
    if (S) {
        hipLaunchKernelGGL((k_sell_spmv<TA, TX, TY>), dim3(k.m.nSlices), dim3(64), 0, k.s, k.m.nC, k.m.d_sellPtr.p,
                           k.m.d_sellWidth.p, k.m.d_sellCol.p, S, x, y);
        return;
    }
    hipLaunchKernelGGL((k_bspmv<TA, TX, TY>), dim3(blocks_for(k.m.nC, 256)), dim3(256), 0, k.s, k.m.nC, k.m.d_rowPtr.p,
                       k.m.d_cols.p, A, x, y);
}

// refresh the sliced-ELL copy of the assembled matrix
template <class T>
void build_sell(Ctx& k, const T* A, DBuf<T>& S) {
    const size_t nv = 25 * (size_t)k.m.sellBlocks;
    if (S.n != nv) {
        S.alloc(nv);
        RDST_HIP(hipMemsetAsync(S.p, 0, sizeof(T) * nv, k.s));  // padding blocks stay zero
    }
    hipLaunchKernelGGL((k_to_sell<T>), dim3(std::min(4096, blocks_for(k.m.nnzb, 256))), dim3(256), 0, k.s,
                       (long)k.m.nnzb, k.m.d_sellOfSlot.p, A, S.p);
}

// the apply-stream buffers of the factor type
static DBuf<double>& stream_buf(SolverWork& w, bool bwd, const double*) { return bwd ? w.sb : w.sf; }
static DBuf<float>& stream_buf(SolverWork& w, bool bwd, const float*) { return bwd ? w.sb32 : w.sf32; }
static bool streams_built(const Mesh& m) { return m.pf.kmax > 0 && m.pb.kmax > 0; }

template <class T>
void pilu_pack(Ctx& k, const T* LU, const T* Dinv) {
    Mesh& m = k.m;
    for (int bwd = 0; bwd < 2; ++bwd) {
        PiluStream& ps = bwd ? m.pb : m.pf;
        DBuf<T>& buf = stream_buf(k.w, bwd != 0, LU);
        const size_t nv = 25 * (size_t)std::max(1L, ps.nBlk);
        if (buf.n != nv) buf.alloc(nv);
        if (ps.nBlk == 0) continue;
        const int nb = (int)((5 * ps.nBlk + 255) / 256);
        if (bwd)
            hipLaunchKernelGGL((k_pilu_pack<T, true>), dim3(nb), dim3(256), 0, k.s, ps.nBlk, ps.d_src.p, ps.d_row.p,
                               ps.d_jn.p, LU, Dinv, buf.p);
        else
            hipLaunchKernelGGL((k_pilu_pack<T, false>), dim3(nb), dim3(256), 0, k.s, ps.nBlk, ps.d_src.p, ps.d_row.p,
                               ps.d_jn.p, LU, Dinv, buf.p);
    }
}

template <class T>
void ilu_factor(Ctx& k, const T* A, T* LU, T* Dinv) {
    Mesh& m = k.m;
    RDST_HIP(hipMemcpyAsync(LU, A, sizeof(T) * 25 * (size_t)m.nnzb, hipMemcpyDeviceToDevice, k.s));
    int* fail = (int*)(k.w.partB.p);  // scratch int
    RDST_HIP(hipMemsetAsync(fail, 0, sizeof(int), k.s));
    if (k.c.st.precond == RDST_PRECOND_ILU0 && k.c.st.ilu_order == RDST_ILU_PARTITIONED) {
        hipLaunchKernelGGL((k_pilu_factor<T>), dim3(m.nParts), dim3(PILU_FACTOR_THREADS), 0, k.s, m.d_partLevStart.p, m.d_partLevPtr.p,
                           m.d_partRows.p, m.d_rowPtr.p, m.d_cols.p, m.d_slotDiag.p, m.d_localCol.p, LU, Dinv, fail);
        if (streams_built(m)) pilu_pack<T>(k, LU, Dinv);
    } else if (k.c.st.precond == RDST_PRECOND_ILU0) {
        for (const LevelRun& g : m.runL) {
            if (g.fused) {
                hipLaunchKernelGGL((k_ilu_factor_seg<T>), dim3(1), dim3(ILU_SEG_FACTOR_THREADS), 0, k.s, g.L0, g.L1,
                                   m.d_levLPtr.p, m.d_levLRows.p, m.d_rowPtr.p, m.d_cols.p, m.d_slotDiag.p, LU, Dinv,
                                   fail);
                continue;
            }
            for (int L = g.L0; L < g.L1; ++L) {
                int cnt = m.levLPtr[L + 1] - m.levLPtr[L];
                hipLaunchKernelGGL((k_ilu_factor<T>), dim3(blocks_for(cnt, 64)), dim3(64), 0, k.s, cnt,
                                   m.d_levLRows.p + m.levLPtr[L], m.d_rowPtr.p, m.d_cols.p, m.d_slotDiag.p, LU, Dinv,
                                   fail);
            }
        }
    } else if (k.c.st.precond == RDST_PRECOND_DIAG) {
        hipLaunchKernelGGL((k_diag_inv<T>), dim3(blocks_for(m.nC, 256)), dim3(256), 0, k.s, m.nC, m.d_slotDiag.p, A, Dinv,
                           fail);
    }
    RDST_HIP(hipGetLastError());
    int hf = 0;
    RDST_HIP(hipMemcpyAsync(&hf, fail, sizeof(int), hipMemcpyDeviceToHost, k.s));
    RDST_HIP(hipStreamSynchronize(k.s));
    if (hf) throw ArgError("block ILU(0): singular pivot block", RDST_ERR_STATE);
}

// x = M^{-1} r
template <class TM, class TV>
void precondition(Ctx& k, const TM* LU, const TM* Dinv, const TV* r, TV* x, TV* ytmp) {
    Mesh& m = k.m;
    const int pc = k.c.st.precond;
    if (pc == RDST_PRECOND_NONE) {
        RDST_HIP(hipMemcpyAsync(x, r, sizeof(TV) * 5 * (size_t)m.nC, hipMemcpyDeviceToDevice, k.s));
        return;
    }
    if (pc == RDST_PRECOND_DIAG) {
        hipLaunchKernelGGL((k_diag_apply<TM, TV>), dim3(blocks_for(m.nC, 256)), dim3(256), 0, k.s, m.nC, Dinv, r, x);
        return;
    }
    if (k.c.st.ilu_order == RDST_ILU_PARTITIONED) {
        const size_t shmem = sizeof(TV) * 5 * (size_t)m.maxPartRows;
        const DBuf<TM>& sfb = stream_buf(k.w, false, LU);
        const DBuf<TM>& sbb = stream_buf(k.w, true, LU);
        if (streams_built(m) && sfb.n > 0 && sbb.n > 0) {
            const PiluStream &f = m.pf, &b = m.pb;
            if (std::max(f.kmax, b.kmax) <= 4)
                hipLaunchKernelGGL((k_pilu_stream<TM, TV, 4>), dim3(m.nParts), dim3(256), shmem, k.s,
                                   m.d_partRowStart.p, m.d_partRows.p, f.d_subStart.p, f.d_sub.p, f.d_ord.p, f.d_lc.p,
                                   sfb.p, b.d_subStart.p, b.d_sub.p, b.d_ord.p, b.d_lc.p, sbb.p, r, x);
            else
                hipLaunchKernelGGL((k_pilu_stream<TM, TV, kPiluKD>), dim3(m.nParts), dim3(256), shmem, k.s,
                                   m.d_partRowStart.p, m.d_partRows.p, f.d_subStart.p, f.d_sub.p, f.d_ord.p, f.d_lc.p,
                                   sfb.p, b.d_subStart.p, b.d_sub.p, b.d_ord.p, b.d_lc.p, sbb.p, r, x);
            return;
        }
        if (sizeof(TM) == 4)
            hipLaunchKernelGGL((k_pilu_apply_rows<TM, TV>), dim3(m.nParts), dim3(256), shmem, k.s, m.d_partRowStart.p,
                               m.d_partLevStart.p, m.d_partLevPtr.p, m.d_partRows.p, m.d_rowPtr.p, m.d_cols.p,
                               m.d_slotDiag.p, m.d_localCol.p, LU, Dinv, r, x);
        else
            hipLaunchKernelGGL((k_pilu_apply<TM, TV>), dim3(m.nParts), dim3(PILU_THREADS), shmem, k.s,
                               m.d_partRowStart.p, m.d_partLevStart.p, m.d_partLevPtr.p, m.d_partRows.p, m.d_rowPtr.p,
                               m.d_cols.p, m.d_slotDiag.p, m.d_localCol.p, LU, Dinv, r, x);
        return;
    }
    for (const LevelRun& g : m.runL) {
        if (g.fused) {
            hipLaunchKernelGGL((k_ilu_fwd_seg<TM, TV>), dim3(1), dim3(ILU_SEG_THREADS), 0, k.s, g.L0, g.L1,
                               m.d_levLPtr.p, m.d_levLRows.p, m.d_rowPtr.p, m.d_cols.p, m.d_slotDiag.p, LU, r, ytmp);
            continue;
        }
        for (int L = g.L0; L < g.L1; ++L) {
            int cnt = m.levLPtr[L + 1] - m.levLPtr[L];
            hipLaunchKernelGGL((k_ilu_fwd<TM, TV>), dim3(blocks_for(cnt, 64)), dim3(64), 0, k.s, cnt,
                               m.d_levLRows.p + m.levLPtr[L], m.d_rowPtr.p, m.d_cols.p, m.d_slotDiag.p, LU, r, ytmp);
        }
    }
    for (const LevelRun& g : m.runU) {
        if (g.fused) {
            hipLaunchKernelGGL((k_ilu_bwd_seg<TM, TV>), dim3(1), dim3(ILU_SEG_THREADS), 0, k.s, g.L0, g.L1,
                               m.d_levUPtr.p, m.d_levURows.p, m.d_rowPtr.p, m.d_cols.p, m.d_slotDiag.p, LU, Dinv, ytmp,
                               x);
            continue;
        }
        for (int L = g.L0; L < g.L1; ++L) {
            int cnt = m.levUPtr[L + 1] - m.levUPtr[L];
            hipLaunchKernelGGL((k_ilu_bwd<TM, TV>), dim3(blocks_for(cnt, 64)), dim3(64), 0, k.s, cnt,
                               m.d_levURows.p + m.levUPtr[L], m.d_rowPtr.p, m.d_cols.p, m.d_slotDiag.p, LU, Dinv, ytmp,
                               x);
        }
    }
}

template <class TA, class TB>
void dot(Ctx& k, long n, const TA* a, const TB* b, double* out, int op) {
    hipLaunchKernelGGL((k_dot_partial<TA, TB>), dim3(k.G), dim3(RED_BS), 0, k.s, n, a, b, k.w.partA.p);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(RED_BS), 0, k.s, k.w.partA.p, k.G, out, op);
}

void cmptmag(Ctx& k, const double* r, double* out /*5, device*/, const double* norm) {
    const int g = std::max(1, std::min(RED_MAXG, blocks_for(k.m.nC, RED_BS)));
    hipLaunchKernelGGL(k_cmptmag_partial, dim3(g), dim3(RED_BS), 0, k.s, k.m.nC, r, k.w.partA.p);
    hipLaunchKernelGGL(k_cmptmag_final, dim3(1), dim3(RED_BS), 0, k.s, k.w.partA.p, g, norm, out);
}

// BlockLduSolver::normFactor: sum(|A x - A xRef| + |b - A xRef|) + small
template <class TA>
void norm_factor(Ctx& k, const TA* A, const double* x, const double* b) {
    const int n = k.m.nC;
    const int g = std::max(1, std::min(RED_MAXG, blocks_for(n, RED_BS)));
    double* sc = k.w.scal.p;
    hipLaunchKernelGGL(k_cmptsum_partial, dim3(g), dim3(RED_BS), 0, k.s, n, x, k.w.partA.p);
    hipLaunchKernelGGL(k_xref_final, dim3(1), dim3(RED_BS), 0, k.s, k.w.partA.p, g, n, sc + S_XREF);
    spmv(k, A, x, k.w.wA.p);
    hipLaunchKernelGGL(k_fill_xref, dim3(blocks_for(n, 256)), dim3(256), 0, k.s, n, sc + S_XREF, k.w.tmp.p);
    spmv(k, A, k.w.tmp.p, k.w.r.p);  // pA in r
    hipLaunchKernelGGL(k_norm_partial, dim3(g), dim3(RED_BS), 0, k.s, n, k.w.wA.p, k.w.r.p, b, k.w.partA.p);
    hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(RED_BS), 0, k.s, k.w.partA.p, g, sc + S_NORM);
}

bool converged(const rdst_settings& st, rdst_solver_perf& p) {
    auto mag5 = [](const double* v) {
        double s = 0;
        for (int k = 0; k < 5; ++k) s += v[k] * v[k];
        return std::sqrt(s);
    };
    bool conv = false;
    if (p.n_iterations >= st.min_iter) {
        double fm = mag5(p.final_residual), im = mag5(p.initial_residual);
        bool chk = (fm < st.tolerance) || (st.rel_tol > 1e-15 && fm <= st.rel_tol * im);
        p.converged = chk;
        conv = p.n_iterations >= st.max_iter || chk;
    }
    return conv;
}

void read_scalars(Ctx& k, int from, int count, double* dst) {
    RDST_HIP(hipMemcpyAsync(dst, k.w.scal.p + from, sizeof(double) * count, hipMemcpyDeviceToHost, k.s));
    RDST_HIP(hipStreamSynchronize(k.s));
}

// one left-preconditioned GMRES(m) restart cycle on A z = rhs from z
// (z updated in place); TV/TM = basis/matrix type
template <class TM, class TV>
void gmres_cycle(Ctx& k, const TM* A, const TM* LU, const TM* Dinv, TV* z, const TV* rres, TV* wv, TV* rv, TV* V, TV* ytmp) {
    const long n5 = k.w.n5;
    const int m = k.w.nDirs;
    const int G = k.G;
    double* sc = k.w.scal.p;
    double* H = sc + S_H;
    double* yh = H + m * m + 3 * m + 1;
    // wA = M^-1 rA (right: rA); beta = |wA|
    if (k.right) hipLaunchKernelGGL((k_copy<TV, TV>), dim3(G), dim3(RED_BS), 0, k.s, n5, wv, rres);
    else precondition(k, LU, Dinv, rres, wv, ytmp);
    dot(k, n5, wv, wv, sc + S_BETA, 1);
    hipLaunchKernelGGL(k_gmres_start, dim3(1), dim3(1), 0, k.s, sc, m);
    for (int i = 0; i < m; ++i) {
        TV* Vi = V + (size_t)i * n5;
        hipLaunchKernelGGL((k_scale_div<TV>), dim3(G), dim3(RED_BS), 0, k.s, n5, Vi, wv, sc + S_BETA);
        if (k.right) {
            precondition(k, LU, Dinv, Vi, rv, ytmp);
            spmv(k, A, rv, wv);
        } else {
            spmv(k, A, Vi, rv);
            precondition(k, LU, Dinv, rv, wv, ytmp);
        }
        double* pin = k.w.partA.p;
        double* pout = k.w.partB.p;
        if (k.c.st.gram_schmidt == RDST_GS_CLASSICAL) {
            // all projections from one multi-dot, then one fused update (+ |w|^2 partials)
            for (int j0 = 0; j0 <= i; j0 += 8) {
                if constexpr (sizeof(TV) == 4) {
                    if (n5 % 4 == 0) {
                        hipLaunchKernelGGL((k_multidot_partial4<8>), dim3(G), dim3(RED_BS), 0, k.s, n5, (const float*)wv,
                                           (const float*)V, j0, std::min(8, i + 1 - j0), k.w.multi.p);
                        continue;
                    }
                }
                hipLaunchKernelGGL((k_multidot_partial<TV, 8>), dim3(G), dim3(RED_BS), 0, k.s, n5, wv, V, j0,
                                   std::min(8, i + 1 - j0), k.w.multi.p);
            }
            hipLaunchKernelGGL(k_multidot_final, dim3(i + 1), dim3(RED_BS), 0, k.s, k.w.multi.p, G, i + 1, H, m, i);
            bool done = false;
            if constexpr (sizeof(TV) == 4) {
                if (n5 % 4 == 0) {
                    hipLaunchKernelGGL(k_cgs_update4, dim3(G), dim3(RED_BS), 0, k.s, n5, (float*)wv, (const float*)V, H,
                                       m, i, i + 1, pin);
                    done = true;
                }
            }
            if (!done)
                hipLaunchKernelGGL((k_cgs_update<TV>), dim3(G), dim3(RED_BS), 0, k.s, n5, wv, V, H, m, i, i + 1, pin);
        } else {
            // partial dot(w, V0), then the fused MGS steps
            hipLaunchKernelGGL((k_dot_partial<TV, TV>), dim3(G), dim3(RED_BS), 0, k.s, n5, wv, V, pin);
            for (int j = 0; j <= i; ++j) {
                const TV* Vj = V + (size_t)j * n5;
                const TV* Vn = (j < i) ? V + (size_t)(j + 1) * n5 : wv;
                hipLaunchKernelGGL((k_mgs_step<TV>), dim3(G), dim3(RED_BS), 0, k.s, n5, wv, Vj, Vn, pin, G, pout,
                                   H + j * m + i);
                std::swap(pin, pout);
            }
        }
        // pin holds partials of |w|^2; rotations, and beta = |w| for the next V
        hipLaunchKernelGGL(k_gmres_column, dim3(1), dim3(RED_BS), 0, k.s, sc, pin, G, i, m);
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(RED_BS), 0, k.s, pin, G, sc + S_BETA, 1);
    }
    hipLaunchKernelGGL(k_gmres_backsub, dim3(1), dim3(1), 0, k.s, sc, m);
    if (k.right) {
        // z += M^-1 (V y)
        RDST_HIP(hipMemsetAsync(wv, 0, sizeof(TV) * n5, k.s));
        hipLaunchKernelGGL((k_update_x<TV, TV>), dim3(G), dim3(RED_BS), 0, k.s, n5, wv, V, yh, m);
        precondition(k, LU, Dinv, wv, rv, ytmp);
        hipLaunchKernelGGL((k_acc<TV>), dim3(G), dim3(RED_BS), 0, k.s, n5, z, rv);
    } else {
        hipLaunchKernelGGL((k_update_x<TV, TV>), dim3(G), dim3(RED_BS), 0, k.s, n5, z, V, yh, m);
    }
}

rdst_solver_perf gmres_fp64(Ctx& k, const double* A, double* x, const double* b) {
    SolverWork& w = k.w;
    const long n5 = w.n5;
    rdst_solver_perf perf{};
    norm_factor(k, A, x, b);
    spmv(k, A, x, w.wA.p);
    hipLaunchKernelGGL(k_sub, dim3(k.G), dim3(RED_BS), 0, k.s, n5, w.r.p, b, w.wA.p);
    cmptmag(k, w.r.p, w.scal.p + S_INIT, w.scal.p + S_NORM);
    read_scalars(k, S_INIT, 5, perf.initial_residual);
    for (int q = 0; q < 5; ++q) perf.final_residual[q] = perf.initial_residual[q];
    if (converged(k.c.st, perf)) return perf;
    do {
        gmres_cycle<double, double>(k, A, w.LU.p, w.Dinv.p, x, w.r.p, w.wA.p, w.tmp.p, w.V.p, w.x0.p);
        spmv(k, A, x, w.wA.p);
        hipLaunchKernelGGL(k_sub, dim3(k.G), dim3(RED_BS), 0, k.s, n5, w.r.p, b, w.wA.p);
        cmptmag(k, w.r.p, w.scal.p + S_FIN, w.scal.p + S_NORM);
        read_scalars(k, S_FIN, 5, perf.final_residual);
        perf.n_iterations++;
    } while (!converged(k.c.st, perf));
    return perf;
}

// mixed precision: FP64 outer defect correction r = b - Aout x (Aout FP32
// promoted for the implicit step, FP64 for the adjoint), inner FP32 GMRES
// cycle on A e = r with the FP32 matrix, FP32 ILU and FP32 basis
template <class TO>
rdst_solver_perf gmres_mixed(Ctx& k, const TO* Aout, const float* A, double* x, const double* b) {
    SolverWork& w = k.w;
    const long n5 = w.n5;
    rdst_solver_perf perf{};
    norm_factor(k, Aout, x, b);
    spmv(k, Aout, x, w.wA.p);
    hipLaunchKernelGGL(k_sub, dim3(k.G), dim3(RED_BS), 0, k.s, n5, w.r.p, b, w.wA.p);
    cmptmag(k, w.r.p, w.scal.p + S_INIT, w.scal.p + S_NORM);
    read_scalars(k, S_INIT, 5, perf.initial_residual);
    for (int q = 0; q < 5; ++q) perf.final_residual[q] = perf.initial_residual[q];
    if (converged(k.c.st, perf)) return perf;
    do {
        hipLaunchKernelGGL((k_copy<float, double>), dim3(k.G), dim3(RED_BS), 0, k.s, n5, w.r32.p, w.r.p);
        RDST_HIP(hipMemsetAsync(w.t32.p, 0, sizeof(float) * n5, k.s));
        // correction e (t32) from zero
        gmres_cycle<float, float>(k, A, w.LU32.p, w.Dinv32.p, w.t32.p, w.r32.p, w.w32.p, w.r32.p, w.V32.p,
                                  (float*)w.tmp.p);
        // x += e in FP64
        hipLaunchKernelGGL((k_copy<double, float>), dim3(k.G), dim3(RED_BS), 0, k.s, n5, w.wA.p, w.t32.p);
        hipLaunchKernelGGL(k_add, dim3(k.G), dim3(RED_BS), 0, k.s, n5, x, w.wA.p);
        spmv(k, Aout, x, w.wA.p);
        hipLaunchKernelGGL(k_sub, dim3(k.G), dim3(RED_BS), 0, k.s, n5, w.r.p, b, w.wA.p);
        cmptmag(k, w.r.p, w.scal.p + S_FIN, w.scal.p + S_NORM);
        read_scalars(k, S_FIN, 5, perf.final_residual);
        perf.n_iterations++;
    } while (!converged(k.c.st, perf));
    return perf;
}

template <class TA, class TB>
__global__ void k_cast(long n, const TA* a, TB* b) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = (TB)a[i];
}

// foam-extend BlockBiCGStabSolver (right-preconditioned), FP64
rdst_solver_perf bicgstab_fp64(Ctx& k, double* x, const double* b) {
    SolverWork& w = k.w;
    const long n5 = w.n5;
    const double* A = k.c.A64.p;
    if (!w.p.p) { w.p.alloc(n5); w.ph.alloc(n5); w.v.alloc(n5); w.t.alloc(n5); w.th.alloc(n5); w.rw.alloc(n5); }
    rdst_solver_perf perf{};
    norm_factor(k, A, x, b);
    spmv(k, A, x, w.p.p);
    hipLaunchKernelGGL(k_sub, dim3(k.G), dim3(RED_BS), 0, k.s, n5, w.r.p, b, w.p.p);
    cmptmag(k, w.r.p, w.scal.p + S_INIT, w.scal.p + S_NORM);
    read_scalars(k, S_INIT, 5, perf.initial_residual);
    for (int q = 0; q < 5; ++q) perf.final_residual[q] = perf.initial_residual[q];
    if (converged(k.c.st, perf)) return perf;
    double* bs = w.scal.p + S_TMP;  // rho, rhoOld, alpha, omega, beta, dot1, dot2, -, (beta,omega), (alpha,omega), alpha, omega
    double init[14] = {1e20, 1e20, 0.0, 1e20, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    RDST_HIP(hipMemcpyAsync(bs, init, sizeof(init), hipMemcpyHostToDevice, k.s));
    RDST_HIP(hipMemsetAsync(w.p.p, 0, sizeof(double) * n5, k.s));
    RDST_HIP(hipMemsetAsync(w.v.p, 0, sizeof(double) * n5, k.s));
    RDST_HIP(hipMemcpyAsync(w.rw.p, w.r.p, sizeof(double) * n5, hipMemcpyDeviceToDevice, k.s));
    do {
        dot(k, n5, w.rw.p, w.r.p, bs + 5, 0);
        hipLaunchKernelGGL(k_bicg_scalars, dim3(1), dim3(1), 0, k.s, bs, 0);
        hipLaunchKernelGGL(k_bicg_p, dim3(k.G), dim3(RED_BS), 0, k.s, n5, w.p.p, w.r.p, w.v.p, bs + 8);
        precondition(k, (const double*)w.LU.p, (const double*)w.Dinv.p, (const double*)w.p.p, w.ph.p, w.x0.p);
        spmv(k, A, w.ph.p, w.v.p);
        dot(k, n5, w.rw.p, w.v.p, bs + 5, 0);
        hipLaunchKernelGGL(k_bicg_scalars, dim3(1), dim3(1), 0, k.s, bs, 1);
        hipLaunchKernelGGL(k_axpy_s, dim3(k.G), dim3(RED_BS), 0, k.s, n5, w.r.p, w.v.p, bs + 12, -1.0);
        precondition(k, (const double*)w.LU.p, (const double*)w.Dinv.p, (const double*)w.r.p, w.th.p, w.x0.p);
        spmv(k, A, w.th.p, w.t.p);
        dot(k, n5, w.t.p, w.r.p, bs + 5, 0);
        dot(k, n5, w.t.p, w.t.p, bs + 6, 0);
        hipLaunchKernelGGL(k_bicg_scalars, dim3(1), dim3(1), 0, k.s, bs, 2);
        hipLaunchKernelGGL(k_bicg_x, dim3(k.G), dim3(RED_BS), 0, k.s, n5, x, w.ph.p, w.th.p, bs + 10);
        hipLaunchKernelGGL(k_axpy_s, dim3(k.G), dim3(RED_BS), 0, k.s, n5, w.r.p, w.t.p, bs + 13, -1.0);
        cmptmag(k, w.r.p, w.scal.p + S_FIN, w.scal.p + S_NORM);
        read_scalars(k, S_FIN, 5, perf.final_residual);
        perf.n_iterations++;
    } while (!converged(k.c.st, perf));
    return perf;
}

}  // namespace

void block_solve(Case& c, rdst_solver_perf* perfOut) {
    if (!c.hasMatrix) throw ArgError("rdst_block_solve: no assembled matrix (call rdst_jacobian_assemble)", RDST_ERR_STATE);
    ensure_work(c);
    Mesh& m = *c.mesh;
    SolverWork& w = *c.work;
    Ctx k{c, m, w, c.stream, w.G};
    const long n5 = w.n5;
    const bool mixed = c.st.precision == RDST_PREC_MIXED;
    if (mixed && c.st.solver != RDST_SOLVER_GMRES)
        throw ArgError("mixed precision is provided for the GMRES solver", RDST_ERR_UNKNOWN_NAME);
    // preconditioner factors and the sliced-ELL copy for the SpMV; with the
    // exact Jacobian the ILU(0) is taken of the Roe operator P (diagonally
    // dominant through its spectral-radius terms) and preconditions the
    // exact operator A (Newton-Krylov)
    const bool newton = c.st.jacobian == RDST_JACOBIAN_EXACT;
    if (mixed) {
        ilu_factor<float>(k, newton ? c.P32.p : c.A32.p, w.LU32.p, w.Dinv32.p);
        if (!c.sellFresh) build_sell<float>(k, c.A32.p, c.S32);
    } else {
        ilu_factor<double>(k, newton ? c.P64.p : c.A64.p, w.LU.p, w.Dinv.p);
        if (!c.sellFresh) build_sell<double>(k, c.A64.p, c.S64);
    }
    rdst_solver_perf perf{};
    double* W = c.f.W;
    if (!mixed && c.st.dependent == RDST_DEP_W) {
        // sourceW = A.W + s; solve A.W = sourceW from W (fvBlockMatrixDST.C:112-123)
        spmv(k, (const double*)c.A64.p, (const double*)W, w.b.p);
        hipLaunchKernelGGL(k_add, dim3(k.G), dim3(RED_BS), 0, k.s, n5, w.b.p, c.src.p);
        perf = (c.st.solver == RDST_SOLVER_GMRES) ? gmres_fp64(k, c.A64.p, W, w.b.p) : bicgstab_fp64(k, W, w.b.p);
    } else {
        // A.dW = s from 0, W += dW (deltaW mode; the mixed form always solves for the correction)
        double* dW = w.dW.p;
        RDST_HIP(hipMemsetAsync(dW, 0, sizeof(double) * n5, k.s));
        if (mixed) perf = gmres_mixed<float>(k, c.A32.p, c.A32.p, dW, c.src.p);
        else perf = (c.st.solver == RDST_SOLVER_GMRES) ? gmres_fp64(k, c.A64.p, dW, c.src.p) : bicgstab_fp64(k, dW, c.src.p);
        hipLaunchKernelGGL(k_add, dim3(k.G), dim3(RED_BS), 0, k.s, n5, W, dW);
        RDST_HIP(hipStreamSynchronize(k.s));
    }
    RDST_HIP(hipGetLastError());
    if (perfOut) *perfOut = perf;
}

// BlockLduSolver<vector5>::solve(x, b) on a caller-assembled LDU system
// (fvBlockMatrixDST.C:98-155): diag [nC][25], upper/lower [nI][25] in
// polyMesh internal-face order (owner row x upper x neighbour value,
// neighbour row x lower x owner value), b and x [nC][5] on the host.  The
// coefficients are scattered into the block-CSR slots on the host and
// uploaded; the case's solver settings select GMRES/BiCGStab, ILU(0) and the
// precision.  mode PLAIN: A x = b from x; W: A x' = A x + b from x;
// DELTAW: A d = b from 0, x += d.  The mixed path solves for the correction
// in FP32 inside the FP64 defect correction on the caller's FP64 matrix.
void ldu_solve(Case& c, const double* diag, const double* upper, const double* lower, const double* b, double* x,
               int mode, rdst_solver_perf* perfOut) {
    Mesh& m = *c.mesh;
    const size_t nn = 25 * (size_t)m.nnzb, n5 = 5 * (size_t)m.nC;
    const bool mixed = c.st.precision == RDST_PREC_MIXED;
    if (mixed && c.st.solver != RDST_SOLVER_GMRES)
        throw ArgError("mixed precision is provided for the GMRES solver", RDST_ERR_UNKNOWN_NAME);
    std::vector<double> vals(nn);
    for (int i = 0; i < m.nC; ++i)
        std::copy(diag + 25 * (size_t)i, diag + 25 * (size_t)i + 25, vals.begin() + 25 * (size_t)m.slotDiag[i]);
    for (int f = 0; f < m.nI; ++f) {
        std::copy(upper + 25 * (size_t)f, upper + 25 * (size_t)f + 25, vals.begin() + 25 * (size_t)m.slotUpper[f]);
        std::copy(lower + 25 * (size_t)f, lower + 25 * (size_t)f + 25, vals.begin() + 25 * (size_t)m.slotLower[f]);
    }
    c.A64.upload(vals);
    ensure_work(c);
    SolverWork& w = *c.work;
    Ctx k{c, m, w, c.stream, w.G};
    if (c.lduX.n != n5) c.lduX.alloc(n5);
    if (c.lduB.n != n5) c.lduB.alloc(n5);
    RDST_HIP(hipMemcpyAsync(c.lduX.p, x, sizeof(double) * n5, hipMemcpyHostToDevice, c.stream));
    RDST_HIP(hipMemcpyAsync(c.lduB.p, b, sizeof(double) * n5, hipMemcpyHostToDevice, c.stream));
    c.sellFresh = false;
    c.csrStale = false;
    c.hasMatrix = true;
    if (mixed) {
        if (c.A32.n != nn) c.A32.alloc(nn);
        hipLaunchKernelGGL((k_cast<double, float>), dim3(std::min(4096, blocks_for((long)nn, 256))), dim3(256), 0, c.stream,
                           (long)nn, c.A64.p, c.A32.p);
        ilu_factor<float>(k, c.A32.p, w.LU32.p, w.Dinv32.p);
        build_sell<float>(k, c.A32.p, c.S32);
        c.S64.release();  // the FP64 outer products read the block-CSR A64, never a stale sliced copy
    } else {
        ilu_factor<double>(k, c.A64.p, w.LU.p, w.Dinv.p);
        build_sell<double>(k, c.A64.p, c.S64);
    }
    rdst_solver_perf perf{};
    double* X = c.lduX.p;
    if (mode == RDST_LDU_PLAIN && !mixed) {
        perf = (c.st.solver == RDST_SOLVER_GMRES) ? gmres_fp64(k, c.A64.p, X, c.lduB.p) : bicgstab_fp64(k, X, c.lduB.p);
    } else if (mode == RDST_LDU_PLAIN) {
        perf = gmres_mixed<double>(k, c.A64.p, c.A32.p, X, c.lduB.p);
    } else if (mode == RDST_LDU_W && !mixed) {
        spmv(k, (const double*)c.A64.p, (const double*)X, w.b.p);
        hipLaunchKernelGGL(k_add, dim3(k.G), dim3(RED_BS), 0, k.s, (long)n5, w.b.p, c.lduB.p);
        perf = (c.st.solver == RDST_SOLVER_GMRES) ? gmres_fp64(k, c.A64.p, X, w.b.p) : bicgstab_fp64(k, X, w.b.p);
    } else {
        // deltaW, and the W form in mixed precision (A x' = A x + b is A (x' - x) = b)
        double* dW = w.dW.p;
        RDST_HIP(hipMemsetAsync(dW, 0, sizeof(double) * n5, k.s));
        if (mixed) perf = gmres_mixed<double>(k, c.A64.p, c.A32.p, dW, c.lduB.p);
        else perf = (c.st.solver == RDST_SOLVER_GMRES) ? gmres_fp64(k, c.A64.p, dW, c.lduB.p) : bicgstab_fp64(k, dW, c.lduB.p);
        hipLaunchKernelGGL(k_add, dim3(k.G), dim3(RED_BS), 0, k.s, (long)n5, X, dW);
    }
    RDST_HIP(hipGetLastError());
    RDST_HIP(hipMemcpyAsync(x, X, sizeof(double) * n5, hipMemcpyDeviceToHost, c.stream));
    RDST_HIP(hipStreamSynchronize(c.stream));
    if (perfOut) *perfOut = perf;
}

// ---------------------------------------------------------------- adjoint
void adjoint_apply(Case& c, bool transpose, const double* x, double* y) {
    if (!c.hasAdjJ) throw ArgError("rdst_adjoint_apply: no linearisation (call rdst_adjoint_linearise)", RDST_ERR_STATE);
    ensure_work(c);
    Ctx k{c, *c.mesh, *c.work, c.stream, c.work->G};
    if (c.adjWide) spmv(k, (const double*)(transpose ? c.adjWT.p : c.adjW.p), x, y);
    else spmv(k, (const double*)(transpose ? c.adjJT.p : c.adjJ.p), x, y);
    RDST_HIP(hipGetLastError());
}

// P = A + shift rDeltaT on the diagonal of every block-CSR row
__global__ void k_shift_diag(int n, const int32_t* slotDiag, const double* A, const double* rDeltaT, double shift,
                             double* P) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const size_t d = 25 * (size_t)slotDiag[c];
    for (int q = 0; q < 25; q += 6) P[d + q] = A[d + q] + shift * rDeltaT[c];
}

// J^T lambda = g with the primal solver machinery on the transposed
// matrix: block ILU(0) of J^T (natural or partitioned order as the case's
// settings) and GMRES(m); mixed: FP32 factors and cycles, FP64 outer residual.
// Viscous: the operator is J^T on the second-neighbour pattern and the ILU(0)
// is taken of its block-CSR part (the face-neighbour blocks)
void adjoint_solve(Case& c, const double* g, double* lambda, const rdst_adjoint_settings& as, rdst_solver_perf* perf) {
    if (!c.hasAdjJ) throw ArgError("rdst_adjoint_solve: no linearisation", RDST_ERR_STATE);
    const rdst_settings saved = c.st;
    c.st.precision = as.precision;
    c.st.n_directions = as.n_directions;
    c.st.max_iter = as.max_cycles;
    c.st.min_iter = 1;
    c.st.tolerance = as.tolerance;
    c.st.rel_tol = as.rel_tol;
    c.st.solver = RDST_SOLVER_GMRES;
    if (c.st.precond == RDST_PRECOND_NONE || c.st.precond == RDST_PRECOND_DIAG) c.st.precond = RDST_PRECOND_ILU0;
    try {
        ensure_work(c);
        Mesh& m = *c.mesh;
        Ctx k{c, m, *c.work, c.stream, c.work->G};
        // the viscous operator differs from the matrix the ILU factors: GMRES
        // then minimises the true residual (right preconditioning)
        k.right = c.adjWide;
        rdst_solver_perf p{};
        // the preconditioner's matrix: J^T's block-CSR part, shifted on request
        const size_t nn = 25 * (size_t)m.nnzb;
        const double* PT = c.adjJT.p;
        if (as.precond_shift > 0.0) {
            if (c.adjPT.n != nn) c.adjPT.alloc(nn);
            RDST_HIP(hipMemcpyAsync(c.adjPT.p, c.adjJT.p, nn * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
            hipLaunchKernelGGL(k_shift_diag, dim3(blocks_for(m.nC, 256)), dim3(256), 0, c.stream, m.nC, m.d_slotDiag.p,
                               c.adjJT.p, c.f.rDeltaT, as.precond_shift, c.adjPT.p);
            PT = c.adjPT.p;
        }
        if (as.precision == RDST_PREC_MIXED) {
            // FP32 factors of the preconditioner's matrix; the inner FP32
            // operator is J^T itself (cast after the factor when shifted) or,
            // with inner_operator 1, that matrix (a defect correction by the
            // pseudo-time operator)
            if (c.adjJT32.n != nn) c.adjJT32.alloc(nn);
            const int gb = std::min(4096, blocks_for((long)nn, 256));
            hipLaunchKernelGGL((k_cast<double, float>), dim3(gb), dim3(256), 0, c.stream, (long)nn, PT, c.adjJT32.p);
            ilu_factor<float>(k, c.adjJT32.p, k.w.LU32.p, k.w.Dinv32.p);
            const double* outer = c.adjWide ? c.adjWT.p : c.adjJT.p;
            if (as.inner_operator == 1) {
                p = gmres_mixed<double>(k, outer, c.adjJT32.p, lambda, g);
            } else if (c.adjWide) {
                const size_t nw = c.adjWT.n;
                if (c.adjWT32.n != nw) c.adjWT32.alloc(nw);
                hipLaunchKernelGGL((k_cast<double, float>), dim3(std::min(4096, blocks_for((long)nw, 256))), dim3(256),
                                   0, c.stream, (long)nw, c.adjWT.p, c.adjWT32.p);
                p = gmres_mixed<double>(k, outer, c.adjWT32.p, lambda, g);
            } else {
                if (PT != c.adjJT.p)
                    hipLaunchKernelGGL((k_cast<double, float>), dim3(gb), dim3(256), 0, c.stream, (long)nn, c.adjJT.p,
                                       c.adjJT32.p);
                p = gmres_mixed<double>(k, outer, c.adjJT32.p, lambda, g);
            }
        } else {
            ilu_factor<double>(k, PT, k.w.LU.p, k.w.Dinv.p);
            p = gmres_fp64(k, c.adjWide ? c.adjWT.p : c.adjJT.p, lambda, g);
        }
        RDST_HIP(hipStreamSynchronize(c.stream));
        if (perf) *perf = p;
    } catch (...) {
        c.st = saved;
        throw;
    }
    c.st = saved;
}

}  // namespace rdst
