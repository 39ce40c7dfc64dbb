// h8_hex20.h -- the 20-node serendipity hexahedron (Hex20): element stiffness
// for general geometry, and the matrix-free operator, Jacobi diagonal and PCG
// kernels on a structured box of congruent bricks.
//
// Node order of an element (the reference's, fe/elements/solids.py:331-406):
// corners 0-7 as the Hex8 element, then the mid-edge nodes
//   8 (0,1)   9 (1,2)  10 (2,3)  11 (3,0)    edges of the bottom face,
//   12 (4,5) 13 (5,6)  14 (6,7)  15 (7,4)    edges of the top face,
//   16 (0,4) 17 (1,5)  18 (2,6)  19 (3,7)    vertical edges;
// natural coordinates h20_xi, h20_eta, h20_zeta.
//
// Box numbering (nx x ny x nz bricks; NI = nx+1, NJ = ny+1, NK = nz+1):
//   corner (i, j, k)                    : (k NI + i) NJ + j                    NI NJ NK nodes
//   x-edge (i, j, k), between i and i+1 : Nc + (k nx + i) NJ + j               nx NJ NK
//   y-edge (i, j, k), between j and j+1 : Nc + Nx + (k NI + i) ny + j          NI ny NK
//   z-edge (i, j, k), between k and k+1 : Nc + Nx + Ny + (k NI + i) NJ + j     NI NJ nz
// (the corners keep the Hex8 numbering); DOF 3 node + c; element e(i, j, k) =
// (k nx + i) ny + j as the Hex8 box.
//
// The operator y = sum_e s_e K0 u_e is node-centric: one launch, each row
// gathered from the elements around its node in a fixed order (no atomics,
// deterministic), with the K0 rows of a slot uniform across a wave so they are
// read through the scalar cache.
#pragma once
#include "h8_device.h"

namespace h8 {

// The two corners a <= b whose midpoint is node n (a = b = n for a corner):
// 8..11 the bottom edges (0,1) (1,2) (2,3) (3,0), 12..15 the top edges (4,5)
// (5,6) (6,7) (7,4), 16..19 the vertical edges (0,4) .. (3,7).  Arithmetic,
// so a node index known only at run time needs no table.
__host__ __device__ constexpr int h20_end_a(int n) {
    return n < 8 ? n : (n < 12 ? (n == 11 ? 0 : n - 8) : (n < 16 ? (n == 15 ? 4 : n - 8) : n - 16));
}
__host__ __device__ constexpr int h20_end_b(int n) {
    return n < 8 ? n : (n < 12 ? (n == 11 ? 3 : n - 7) : (n < 16 ? (n == 15 ? 7 : n - 7) : n - 12));
}
// natural coordinates (-1, 0 or 1) of node n: the mean of its two corners'
__host__ __device__ constexpr int h20_xi(int n) { return OX(h20_end_a(n)) + OX(h20_end_b(n)) - 1; }
__host__ __device__ constexpr int h20_eta(int n) { return OY(h20_end_a(n)) + OY(h20_end_b(n)) - 1; }
__host__ __device__ constexpr int h20_zeta(int n) { return OZ(h20_end_a(n)) + OZ(h20_end_b(n)) - 1; }
static_assert(h20_xi(8) == 0 && h20_eta(8) == -1 && h20_zeta(8) == -1, "node 8 on edge (0,1)");
static_assert(h20_xi(9) == 1 && h20_eta(9) == 0 && h20_zeta(9) == -1, "node 9 on edge (1,2)");
static_assert(h20_xi(11) == -1 && h20_eta(11) == 0 && h20_zeta(11) == -1, "node 11 on edge (3,0)");
static_assert(h20_xi(14) == 0 && h20_eta(14) == 1 && h20_zeta(14) == 1, "node 14 on edge (6,7)");
static_assert(h20_xi(15) == -1 && h20_eta(15) == 0 && h20_zeta(15) == 1, "node 15 on edge (7,4)");
static_assert(h20_xi(18) == 1 && h20_eta(18) == 1 && h20_zeta(18) == 0, "node 18 on edge (2,6)");
static_assert(h20_xi(6) == 1 && h20_eta(6) == 1 && h20_zeta(6) == 1, "corner 6");

struct H20Grid {
    int nx, ny, nz;
    int NI, NJ, NK;
    long long nc, nxe, nye, nze;     // corner, x-, y-, z-edge node counts
    long long nnodes, ndof, m;
};

__host__ __device__ inline H20Grid h20_grid(int nx, int ny, int nz) {
    H20Grid g;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.NI = nx + 1; g.NJ = ny + 1; g.NK = nz + 1;
    g.nc = (long long)g.NI * g.NJ * g.NK;
    g.nxe = (long long)nx * g.NJ * g.NK;
    g.nye = (long long)g.NI * ny * g.NK;
    g.nze = (long long)g.NI * g.NJ * nz;
    g.nnodes = g.nc + g.nxe + g.nye + g.nze;
    g.ndof = 3 * g.nnodes;
    g.m = (long long)nx * ny * nz;
    return g;
}

// Serendipity shape-function derivatives dN_n/d(xi, eta, zeta) at a point.
__device__ __forceinline__ void h20_dshape(double xi, double et, double ze, int n, double& dx, double& dy,
                                           double& dz) {
    const double xn = h20_xi(n), yn = h20_eta(n), zn = h20_zeta(n);
    if (xn != 0 && yn != 0 && zn != 0) {
        const double a = xi * xn, b = et * yn, c = ze * zn;
        dx = 0.125 * xn * (1 + b) * (1 + c) * (2 * a + b + c - 1);
        dy = 0.125 * yn * (1 + a) * (1 + c) * (a + 2 * b + c - 1);
        dz = 0.125 * zn * (1 + a) * (1 + b) * (a + b + 2 * c - 1);
    } else if (xn == 0) {
        dx = -0.5 * xi * (1 + et * yn) * (1 + ze * zn);
        dy = 0.25 * (1 - xi * xi) * yn * (1 + ze * zn);
        dz = 0.25 * (1 - xi * xi) * (1 + et * yn) * zn;
    } else if (yn == 0) {
        dx = 0.25 * xn * (1 - et * et) * (1 + ze * zn);
        dy = -0.5 * et * (1 + xi * xn) * (1 + ze * zn);
        dz = 0.25 * (1 + xi * xn) * (1 - et * et) * zn;
    } else {
        dx = 0.25 * xn * (1 + et * yn) * (1 - ze * ze);
        dy = 0.25 * (1 + xi * xn) * yn * (1 - ze * ze);
        dz = -0.5 * ze * (1 + xi * xn) * (1 + et * yn);
    }
}

// General-geometry Hex20 stiffness, one 256-thread block per element: threads
// 0..26 each take one point of the 3x3x3 Gauss rule (xi outer, eta middle,
// zeta inner; weights 5/9, 8/9, 5/9 per axis) and form J = dN X, detJ w and the
// physical gradients J^-1 dN in LDS; then each thread owns entries of the 60 x
// 60 sum over the points of B^T D B detJ w (Voigt rows xx, yy, zz, xy, yz, zx).
__global__ void __launch_bounds__(256)
k_ke20(long long m, const double* __restrict__ X, const double* __restrict__ D, long long d_stride,
       double* __restrict__ ke) {
    __shared__ double sG[27][3][20];
    __shared__ double sW[27];
    __shared__ double sD[36];
    const long long e = blockIdx.x;
    if (e >= m) return;
    if (threadIdx.x < 36) sD[threadIdx.x] = D[e * d_stride + threadIdx.x];
    if (threadIdx.x < 27) {
        const int q = threadIdx.x;
        const double g = 0.77459666924148337704;    // sqrt(3/5)
        const int ix = q / 9, ie = (q / 3) % 3, iz = q % 3;
        const double xi = (ix - 1) * g, et = (ie - 1) * g, ze = (iz - 1) * g;
        const double w = (ix == 1 ? 8.0 / 9.0 : 5.0 / 9.0) * (ie == 1 ? 8.0 / 9.0 : 5.0 / 9.0) *
                         (iz == 1 ? 8.0 / 9.0 : 5.0 / 9.0);
        double dN[3][20];
#pragma unroll
        for (int n = 0; n < 20; ++n) h20_dshape(xi, et, ze, n, dN[0][n], dN[1][n], dN[2][n]);
        double J[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double acc = 0.0;
#pragma unroll
                for (int n = 0; n < 20; ++n) acc += dN[r][n] * X[e * 60 + 3 * n + c];
                J[r][c] = acc;
            }
        const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
        const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
        const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
        const double id = 1.0 / det;
        double Ji[3][3];
        Ji[0][0] = c00 * id;
        Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
        Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
        Ji[1][0] = c01 * id;
        Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
        Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
        Ji[2][0] = c02 * id;
        Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
        Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int n = 0; n < 20; ++n)
                sG[q][r][n] = Ji[r][0] * dN[0][n] + Ji[r][1] * dN[1][n] + Ji[r][2] * dN[2][n];
        sW[q] = det * w;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < 3600; idx += blockDim.x) {
        const int I = idx / 60, Jc = idx % 60;
        const int a = I / 3, c = I % 3, b = Jc / 3, d = Jc % 3;
        double acc = 0.0;
        for (int q = 0; q < 27; ++q) {
            const double ga0 = sG[q][0][a], ga1 = sG[q][1][a], ga2 = sG[q][2][a];
            const double gb0 = sG[q][0][b], gb1 = sG[q][1][b], gb2 = sG[q][2][b];
            double Bi[6] = {0, 0, 0, 0, 0, 0}, Bj[6] = {0, 0, 0, 0, 0, 0};
            if (c == 0) { Bi[0] = ga0; Bi[3] = ga1; Bi[5] = ga2; }
            else if (c == 1) { Bi[1] = ga1; Bi[3] = ga0; Bi[4] = ga2; }
            else { Bi[2] = ga2; Bi[4] = ga1; Bi[5] = ga0; }
            if (d == 0) { Bj[0] = gb0; Bj[3] = gb1; Bj[5] = gb2; }
            else if (d == 1) { Bj[1] = gb1; Bj[3] = gb0; Bj[4] = gb2; }
            else { Bj[2] = gb2; Bj[4] = gb1; Bj[5] = gb0; }
            double sum = 0.0;
#pragma unroll
            for (int s = 0; s < 6; ++s) {
                double dbj = 0.0;
#pragma unroll
                for (int t = 0; t < 6; ++t) dbj += sD[6 * s + t] * Bj[t];
                sum += Bi[s] * dbj;
            }
            acc += sum * sW[q];
        }
        ke[e * 3600 + idx] = acc;
    }
}

// ---- the node-centric operator ---------------------------------------------
// Row 3n + c of K = sum_e s_e K0 gathers, over the elements e around node n,
// s_e K0[3a + c][.] u_e with a the node's local index in e.  A node of family
// FAM (0 corner, 1 x-edge, 2 y-edge, 3 z-edge) at box position (i, j, k) has
// up to 8 (corner) or 4 (edge) elements around it; slot s names one of them:
// its offset (di, dj, dk) from (i, j, k) and the local index a of the node in
// it, the same for every node of the family -- so a K0 row address depends on
// s only, which is uniform across a wave (the rows are read through the
// scalar cache).
struct H20Slot {
    int di, dj, dk, a;
};
__host__ __device__ constexpr int h20_corner_of(int ox, int oy, int oz) {
    return (oy == 0 ? ox : 3 - ox) + 4 * oz;
}
__host__ __device__ constexpr H20Slot h20_slot(int fam, int s) {
    // offset -1 or 0 along each axis where the node is an element corner, 0
    // along an edge node's own axis
    return fam == 0 ? H20Slot{(s & 1) - 1, ((s >> 1) & 1) - 1, (s >> 2) - 1,
                              h20_corner_of(1 - (s & 1), 1 - ((s >> 1) & 1), 1 - (s >> 2))}
         : fam == 1 ? H20Slot{0, (s & 1) - 1, (s >> 1) - 1, 8 + 2 * (1 - (s & 1)) + 4 * (1 - (s >> 1))}
         : fam == 2 ? H20Slot{(s & 1) - 1, 0, (s >> 1) - 1, ((s & 1) ? 11 : 9) + 4 * (1 - (s >> 1))}
                    : H20Slot{(s & 1) - 1, (s >> 1) - 1, 0, 16 + h20_corner_of(1 - (s & 1), 1 - (s >> 1), 0)};
}
// The local node a of the element at offset (di, dj, dk) sits at the node:
// natural coordinate +1 where the offset is -1, -1 where it is 0, and 0
// along an edge node's own axis.
__host__ __device__ constexpr bool h20_slot_ok(int fam, int s) {
    return h20_xi(h20_slot(fam, s).a) == (fam == 1 ? 0 : (h20_slot(fam, s).di < 0 ? 1 : -1)) &&
           h20_eta(h20_slot(fam, s).a) == (fam == 2 ? 0 : (h20_slot(fam, s).dj < 0 ? 1 : -1)) &&
           h20_zeta(h20_slot(fam, s).a) == (fam == 3 ? 0 : (h20_slot(fam, s).dk < 0 ? 1 : -1));
}
static_assert(h20_slot_ok(0, 0) && h20_slot_ok(0, 1) && h20_slot_ok(0, 2) && h20_slot_ok(0, 3) &&
              h20_slot_ok(0, 4) && h20_slot_ok(0, 5) && h20_slot_ok(0, 6) && h20_slot_ok(0, 7),
              "corner slots");
static_assert(h20_slot_ok(1, 0) && h20_slot_ok(1, 1) && h20_slot_ok(1, 2) && h20_slot_ok(1, 3), "x-edge slots");
static_assert(h20_slot_ok(2, 0) && h20_slot_ok(2, 1) && h20_slot_ok(2, 2) && h20_slot_ok(2, 3), "y-edge slots");
static_assert(h20_slot_ok(3, 0) && h20_slot_ok(3, 1) && h20_slot_ok(3, 2) && h20_slot_ok(3, 3), "z-edge slots");

// global node (32-bit: the host caps the box at 2^31 DOFs) of local node n of element (i, j, k)
__device__ __forceinline__ int h20_node32(const H20Grid& g, int i, int j, int k, int n) {
    const int x = h20_xi(n), y = h20_eta(n), z = h20_zeta(n);
    const int gi = i + (x > 0), gj = j + (y > 0), gk = k + (z > 0);
    if (x != 0 && y != 0 && z != 0) return (gk * g.NI + gi) * g.NJ + gj;
    if (x == 0) return int(g.nc) + (gk * g.nx + i) * g.NJ + gj;
    if (y == 0) return int(g.nc + g.nxe) + (gk * g.NI + gi) * g.ny + j;
    return int(g.nc + g.nxe + g.nye) + (k * g.NI + gi) * g.NJ + gj;
}

// Rows 3n .. 3n+2 of node n (family FAM at (i, j, k)): the sum over the
// elements around the node of s_e K0[3a + c][.] u_e; DIAG: of s_e
// K0[3a + c][3a + c].  The element's 60 values are loaded once for the three
// rows (all loads in flight before the products); elements in a fixed order,
// each one's 60 products in a fixed order: the result does not depend on the
// launch.
template <int FAM, bool DIAG>
__device__ __forceinline__ void h20_rows(const H20Grid& g, int i, int j, int k,
                                         const double* __restrict__ s_e, const double* __restrict__ K0,
                                         const double* __restrict__ u, double (&acc)[3]) {
    constexpr int NS = FAM == 0 ? 8 : 4;
#pragma unroll 1
    for (int s = 0; s < NS; ++s) {
        const H20Slot sl = h20_slot(FAM, s);
        const int ei = i + sl.di, ej = j + sl.dj, ek = k + sl.dk;
        if (ei < 0 || ej < 0 || ek < 0 || ei >= g.nx || ej >= g.ny || ek >= g.nz) continue;
        const double se = s_e[(ek * g.nx + ei) * g.ny + ej];
        const double* __restrict__ Kr = K0 + 3 * sl.a * 60;
        if constexpr (DIAG) {
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[c] += se * Kr[c * 60 + 3 * sl.a + c];
        } else {
            double ue[60];
#pragma unroll
            for (int b = 0; b < 20; ++b) {
                const double* __restrict__ ub = u + 3 * h20_node32(g, ei, ej, ek, b);
                ue[3 * b] = ub[0];
                ue[3 * b + 1] = ub[1];
                ue[3 * b + 2] = ub[2];
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double e = 0.0;
#pragma unroll
                for (int t = 0; t < 60; ++t) e += Kr[c * 60 + t] * ue[t];
                acc[c] += se * e;
            }
        }
    }
}

template <bool DIAG>
__device__ __forceinline__ void h20_node_rows(const H20Grid& g, int n, const double* __restrict__ s_e,
                                              const double* __restrict__ K0, const double* __restrict__ u,
                                              double (&acc)[3]) {
    acc[0] = acc[1] = acc[2] = 0.0;
    if (n < g.nc) {
        const int j = n % g.NJ, t = n / g.NJ;
        h20_rows<0, DIAG>(g, t % g.NI, j, t / g.NI, s_e, K0, u, acc);
        return;
    }
    n -= int(g.nc);
    if (n < g.nxe) {
        const int j = n % g.NJ, t = n / g.NJ;
        h20_rows<1, DIAG>(g, t % g.nx, j, t / g.nx, s_e, K0, u, acc);
        return;
    }
    n -= int(g.nxe);
    if (n < g.nye) {
        const int j = n % g.ny, t = n / g.ny;
        h20_rows<2, DIAG>(g, t % g.NI, j, t / g.NI, s_e, K0, u, acc);
        return;
    }
    n -= int(g.nye);
    const int j = n % g.NJ, t = n / g.NJ;
    h20_rows<3, DIAG>(g, t % g.NI, j, t / g.NI, s_e, K0, u, acc);
}

// One thread per node; at most H20_MAX_BLOCKS blocks, striding over the nodes.
constexpr int H20_THREADS = 256;
constexpr int H20_MAX_BLOCKS = 2048;

// y = K u (DIAG: y = the diagonal of K), every DOF.
template <bool DIAG>
__global__ void __launch_bounds__(H20_THREADS)
k_h20_node(H20Grid g, const double* __restrict__ s_e, const double* __restrict__ K0,
           const double* __restrict__ u, double* __restrict__ y) {
    for (long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x; n < g.nnodes;
         n += (long long)gridDim.x * blockDim.x) {
        double acc[3];
        h20_node_rows<DIAG>(g, int(n), s_e, K0, u, acc);
#pragma unroll
        for (int c = 0; c < 3; ++c) y[3 * n + c] = acc[c];
    }
}

// Inside the PCG iterations: q = K p with p the direction of the state's
// iteration (p[k & 1]), and p.q summed over the blocks; the last block forms
// alpha, or stops the solve with NOT_SPD when p.q <= 0.
__global__ void __launch_bounds__(H20_THREADS)
k_h20_node_pcg(H20Grid g, const double* __restrict__ s_e, const double* __restrict__ K0,
               const double* __restrict__ p0, const double* __restrict__ p1, double* __restrict__ q,
               PcgState* st, double* partials) {
    __shared__ double sRed[16];
    if (st->status != ST_RUNNING) return;
    const double* __restrict__ p = (st->k & 1) ? p1 : p0;
    double v[1] = {0.0};
    for (long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x; n < g.nnodes;
         n += (long long)gridDim.x * blockDim.x) {
        double acc[3];
        h20_node_rows<false>(g, int(n), s_e, K0, p, acc);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            q[3 * n + c] = acc[c];
            v[0] += p[3 * n + c] * acc[c];
        }
    }
    double tot[1];
    if (reduce_last_block<1>(v, partials, &st->cnt[0], tot, sRed) && threadIdx.x == 0) {
        st->pq = tot[0];
        if (!(tot[0] > 0.0)) {
            st->bad_pq = tot[0];
            st->iters = st->k;
            st->status = ST_NOT_SPD;
            st->pending = 0;
        } else {
            st->alpha = st->rz_old / tot[0];
            st->pending = 1;
        }
    }
}

// dinv = 1/d at free DOFs (0 at fixed ones); *flag |= 1 where a free d is 0.
__global__ void __launch_bounds__(256)
k_h20_dinv(long long n, const double* __restrict__ d, const uint8_t* __restrict__ free_mask,
           double* __restrict__ dinv, int* flag) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const bool fr = free_mask[i] != 0;
        if (fr && d[i] == 0.0) atomicOr(flag, 1);
        dinv[i] = (fr && d[i] != 0.0) ? 1.0 / d[i] : 0.0;
    }
}

// PCG direction of iteration k: p[k&1] = dinv r (+ beta p[(k-1)&1] after the
// first iteration).
__global__ void __launch_bounds__(256)
k_h20_dir(long long n, const double* __restrict__ r, const double* __restrict__ dinv, double* __restrict__ p0,
          double* __restrict__ p1, const PcgState* st) {
    if (st->status != ST_RUNNING) return;
    const bool odd = (st->k & 1) != 0;
    double* __restrict__ pn = odd ? p1 : p0;
    const double* __restrict__ po = odd ? p0 : p1;
    const double beta = st->first ? 0.0 : st->beta;
    const bool first = st->first != 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const double z = dinv[i] * r[i];
        pn[i] = first ? z : z + beta * po[i];
    }
}

}  // namespace h8
