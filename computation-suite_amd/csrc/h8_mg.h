// h8_mg.h -- geometric multigrid V-cycle used as the preconditioner of the
// FP32 iterations of the refined solve (M^-1 r instead of dinv * r).
//
// Hierarchy.  Level 0 is the solve's own matrix-free operator K = sum_e s_e K0
// on the box grid.  Level l+1 halves every axis of level l whose element count
// is even (the others are kept), and its operator is the Galerkin product
// A_{l+1} = P_l^T A_l P_l with P_l the trilinear interpolation from the
// coarse nodes, restricted to the free DOFs of level l (P_l = M_l P, M_l the
// free mask).  A coarse DOF none of whose fine neighbours is free has a zero
// row and column and is masked (dinv = 0) on its level.  Coarse operators are
// stored assembled, as a 27-point stencil of 3x3 blocks per node:
//     S[(nb * 9 + 3 c + d) * nnodes + n] = A[3 n + c][3 m + d],
//     m = n + (dk * NI + di) * NJ + dj, nb = ((dk + 1) * 3 + di + 1) * 3 + dj + 1,
// FP32 (the preconditioner's precision; the Galerkin sums are formed in FP64).
// The coarsest level (a few hundred DOFs) is inverted densely once per setup.
//
// Level 1 from the elements.  On a nested brick grid the trilinear P is
// element-local: a coarse element E (r_x r_y r_z fine elements, children c)
// maps its 8 corners to the corners of each child by P_c (24 x 24), so
//     A_1 = sum_E sum_c s_{E,c} P_c^T M_e K0 M_e P_c,
// and when all 24 DOFs of child e are free this is s_e M_c with the 8 fixed
// matrices M_c = P_c^T K0 P_c (computed once per setup from K0).
//
// Smoother: Chebyshev iteration of degree 2 on D^-1 A over [lam/30, lam],
// lam = the Gershgorin bound max_i sum_j |A_ij| / A_ii of each level (no
// power iterations), the same polynomial before and after the coarse
// correction, so the V-cycle is a symmetric positive definite operator (the
// FP32 PCG stays a CG method).  Step s of a smoothing pass:
//     d = ca_s d + cb_s dinv (b - A x),  x = x + d,
// (ca_0, cb_0) = (0, 1/theta), (ca_1, cb_1) = (rho1 rho0, 2 rho1 / delta),
// theta = (lam + lam/30) / 2, delta = (lam - lam/30) / 2, rho0 = delta/theta,
// rho1 = 1 / (2 theta/delta - rho0) (Saad, Iterative Methods for Sparse
// Linear Systems, 2nd ed., Alg. 12.1).
//
// Every kernel of a cycle returns at once unless the solve state is RUNNING,
// so a captured chunk of iterations costs nothing after convergence.
#pragma once
#include "h8_device.h"
#include "h8_node.h"

namespace h8 {

constexpr int MG_MAX_LEVELS = 12;
constexpr int MG_STENCIL = 243;            // 27 neighbours x 3 x 3
constexpr long long MG_COARSE_DOF = 128;   // stop coarsening at or below this many DOFs
constexpr int MG_LDS_GJ = 128;             // coarsest levels up to this many DOFs are inverted in registers
constexpr long long MG_DENSE_MAX = 512;    // the coarsest level must be this small (dense inverse)
#ifndef H8_MG_CHEB_RATIO
#define H8_MG_CHEB_RATIO 30.0f
#endif
constexpr float MG_CHEB_RATIO = H8_MG_CHEB_RATIO;   // smoothing interval [lam / ratio, lam]
// Degree of the level-0 smoother: every smoothing pass of the finest level is
// this many Chebyshev steps (stencil levels always take 2).  Each step is one
// operator application, so a higher degree buys a stronger cycle (fewer
// iterations) for more level-0 launches per cycle.
constexpr int MG_CHEB_MAXDEG = 6;
// The library's degree: 4 on grids up to 2^20 DOFs, where every kernel of the
// cycle is short and the launches set the pace (fewer iterations win), 3 above,
// where the level-0 steps move most of the cycle's bytes (measured: 128x32x32
// 8.1 against 8.8 ms at degree 3; 192x48x48 even; 256x64x64 1-5% faster at 3).
constexpr long long MG_DEG0_SPLIT = 1LL << 20;
__host__ __device__ constexpr int mg_deg0_default(long long ndof) { return ndof <= MG_DEG0_SPLIT ? 4 : 3; }
#ifndef H8_MG_BOTTOM_NODES
#define H8_MG_BOTTOM_NODES 256
#endif
#ifndef H8_MG_BOTTOM_THREADS
#define H8_MG_BOTTOM_THREADS 512
#endif
constexpr long long MG_BOTTOM_NODES = H8_MG_BOTTOM_NODES;   // levels this small run in one workgroup (k_mg_bottom)
constexpr int MG_BOTTOM_THREADS = H8_MG_BOTTOM_THREADS;

// One level of the hierarchy (by value in kernel arguments).
struct MgLevel {
    int nx, ny, nz;        // elements
    int NJ, NI, NK;        // nodes per axis
    long long nnodes, ndof;
    int rx, ry, rz;        // ratios to the next coarser level (1 or 2; 0 on the coarsest)
    float* S;              // stencil (levels >= 1)
    float* dinv;           // inverse diagonal, 0 at masked DOFs (level 0: the solve's)
    float* x0;             // iterates (levels >= 1: x1 holds a cycle's result)
    float* x1;
    float* b;              // right-hand side (levels >= 1)
    float* d;              // Chebyshev direction
    float* y;              // masked residual M (b - A x1) after the pre-smoothing (the restriction's input)
};

// Per-level smoother constants, in device memory (written by the setup kernels).
struct MgConst {
    unsigned lam_bits;     // Gershgorin bound (bits of a positive float; atomicMax)
    // Chebyshev steps d = ca_k d + cb_k dinv (b - A x), x += d: step 0 has
    // ca = 0, cb = c0; step k >= 1 reads (ca_k, cb_k) = (&c0)[2k - 1], (&c0)[2k]
    // (c1a, c1b, then ck in pairs), so a kernel indexes one float array
    float c0, c1a, c1b;
    float ck[2 * (MG_CHEB_MAXDEG - 2)];
};

__device__ __forceinline__ bool mg_live(const PcgState* st) { return !st || st->status == ST_RUNNING; }

// (j, i, k) of node n.  32-bit arithmetic: the plan keeps every index of a
// stencil level below 2^31 (243 x its nodes), and 64-bit division is a long
// software sequence on the GPU.
__device__ __forceinline__ void mg_decode(const MgLevel& L, long long n, int& j, int& i, int& k) {
    const unsigned nn = unsigned(n), NJ = unsigned(L.NJ), NI = unsigned(L.NI);
    const unsigned q = nn / NJ;
    j = int(nn - q * NJ);
    k = int(q / NI);
    i = int(q - unsigned(k) * NI);
}

// Interpolation weight of fine position f (in fine element units, within one
// coarse element of ratio r) from coarse corner C (0 or 1).
__host__ __device__ constexpr double mg_w(int f, int C, int r) {
    return r == 1 ? (f == C ? 1.0 : 0.0) : (f == 2 * C ? 1.0 : (f == 1 ? 0.5 : 0.0));
}
// The same weight as a power of two: mg_w = 2^-mg_wexp, or 0 when mg_wexp is 16
// (a sum of three is then >= 16).
__host__ __device__ constexpr unsigned mg_wexp(int f, int C, int r) {
    return r == 1 ? (f == C ? 0u : 16u) : (f == 2 * C ? 0u : (f == 1 ? 1u : 16u));
}

// M_c = P_c^T K0 P_c for the (up to 8) children c of a coarse element with
// ratios (rx, ry, rz); child c = cx + 2 cy + 4 cz, entries [c][24][24].
__global__ void __launch_bounds__(256) k_mg_mc(const double* __restrict__ K0, int rx, int ry, int rz,
                                               double* __restrict__ M) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < 8 * 576; t += gridDim.x * blockDim.x) {
        const int c = t / 576, row = (t / 24) % 24, col = t % 24;
        const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
        if (cx >= rx || cy >= ry || cz >= rz) {
            M[t] = 0.0;
            continue;
        }
        const int A = row / 3, i = row % 3, B = col / 3, j = col % 3;
        // all 64 terms, the 8 of a corner a unrolled so that their K0 loads are
        // in flight together; a zero weight adds an exact zero, so the sum
        // equals that of the nonzero terms alone
        double acc = 0.0;
#pragma unroll 1
        for (int a = 0; a < 8; ++a) {
            const double pa = mg_w(cx + OX(a), OX(A), rx) * mg_w(cy + OY(a), OY(A), ry) * mg_w(cz + OZ(a), OZ(A), rz);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const double pb =
                    mg_w(cx + OX(b), OX(B), rx) * mg_w(cy + OY(b), OY(B), ry) * mg_w(cz + OZ(b), OZ(B), rz);
                acc += pa * K0[(3 * a + i) * 24 + 3 * b + j] * pb;
            }
        }
        M[t] = acc;
    }
}

// Block-wide max of a non-negative float, then one atomicMax of its bits per
// block (per-thread atomics on the one word serialise: an 81-node level's
// kernel took 25 us).  Every thread of the block must call it.
__device__ __forceinline__ void block_max_to(unsigned* word, float v) {
    __shared__ float s_max[16];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if (lane == 0) s_max[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = 0.0f;
        for (int w = 0; w < nw; ++w) m = fmaxf(m, s_max[w]);
        if (m > 0.0f) atomicMax(word, __float_as_uint(m));
    }
}

// Gershgorin bound of D^-1 K on level 0 from the elements: row (n, c) has
// sum_j |K_ij| <= sum_e s_e sum_(b,d) |K0[3a+c][3b+d]| (a = the node's corner
// in e).  Free rows only; the bound's bits go to cst->lam_bits by atomicMax.
__global__ void __launch_bounds__(256) k_mg_gersh0(GridDev g, const float* __restrict__ s_e,
                                                   const double* __restrict__ K0, const float* __restrict__ dinv,
                                                   MgConst* cst) {
    __shared__ float rs[24];
    if (threadIdx.x < 24) {
        double acc = 0.0;
        for (int j = 0; j < 24; ++j) acc += fabs(K0[threadIdx.x * 24 + j]);
        rs[threadIdx.x] = float(acc);
    }
    __syncthreads();
    float best = 0.0f;
    for (long long nd = (long long)blockIdx.x * blockDim.x + threadIdx.x; nd < g.nnodes;
         nd += (long long)gridDim.x * blockDim.x) {
        const unsigned q = unsigned(nd) / unsigned(g.NJ);
        const int j = int(unsigned(nd) - q * unsigned(g.NJ)), k = int(q / unsigned(g.NI));
        const int i = int(q - unsigned(k) * unsigned(g.NI));
        float sum[3] = {0.f, 0.f, 0.f};
        for (int a = 0; a < 8; ++a) {
            const int je = j - OY(a), ie = i - OX(a), ke = k - OZ(a);
            if (je < 0 || je >= g.ny || ie < 0 || ie >= g.nx || ke < 0 || ke >= g.nz) continue;
            const float s = s_e[((long long)ke * g.nx + ie) * g.ny + je];
            for (int c = 0; c < 3; ++c) sum[c] += s * rs[3 * a + c];
        }
        for (int c = 0; c < 3; ++c) {
            const float di = dinv[3 * nd + c];
            if (di != 0.0f) best = fmaxf(best, 1.0001f * sum[c] * di);
        }
    }
    block_max_to(&cst->lam_bits, best);
}

// Free-DOF bits of every fine element (bit 3a + c: corner a, component c).
__global__ void __launch_bounds__(256) k_mg_elem_mask(GridDev g, const float* __restrict__ dinv0,
                                                      unsigned* __restrict__ emask) {
    const long long m = (long long)g.nx * g.ny * g.nz;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < m;
         e += (long long)gridDim.x * blockDim.x) {
        const unsigned q = unsigned(e) / unsigned(g.ny);
        const int je = int(unsigned(e) - q * unsigned(g.ny)), ke = int(q / unsigned(g.nx));
        const int ie = int(q - unsigned(ke) * unsigned(g.nx));
        unsigned mask = 0;
        for (int a = 0; a < 8; ++a) {
            const long long fn = ((long long)(ke + OZ(a)) * g.NI + (ie + OX(a))) * g.NJ + (je + OY(a));
            for (int c = 0; c < 3; ++c)
                if (dinv0[3 * fn + c] != 0.0f) mask |= 1u << (3 * a + c);
        }
        emask[e] = mask;
    }
}

// Level-1 stencil straight from the fine elements: entry (coarse node I,
// neighbour block nb, 3 c + d) of A_1 = P^T A_0 P sums, over the coarse
// elements E holding both I (corner A) and J = I + nb (corner B) and over E's
// fine children e (child position ch), s_e (P_ch^T K0 P_ch)[3A + c][3B + d]
// -- the fixed child blocks M (k_mg_mc) -- when every DOF of e is free; a
// child with fixed DOFs is evaluated with its free mask from K0.  A thread
// owns one (node, neighbour) pair and accumulates its nine entries in FP64
// (coarse elements, then children, in a fixed order), so the element loops
// and the scale and mask loads are shared by nine products.  Pairs are
// numbered neighbour-major (p = nb * nodes + I): the 64 lanes of a wave hold
// consecutive nodes of one neighbour block, so the element loops, the corner
// pair (A, B) and with them every read of M and K0 (staged in LDS) are the
// same for the whole wave -- one broadcast LDS read per value, no divergence
// except where a boundary clips the loops -- and the stores of an entry are
// contiguous.  The masked children (boundary elements only) then take the
// same uniform loops.
__global__ void __launch_bounds__(256) k_mg_level1(GridDev g, MgLevel C, int rx, int ry, int rz,
                                                   const double* __restrict__ s_e, const double* __restrict__ K0,
                                                   const double* __restrict__ Mg, const unsigned* __restrict__ emask) {
    __shared__ double M[8 * 576];
    __shared__ double sK0[576];
    for (int t = threadIdx.x; t < 8 * 576; t += blockDim.x) M[t] = Mg[t];
    for (int t = threadIdx.x; t < 576; t += blockDim.x) sK0[t] = K0[t];
    __syncthreads();
    // 27 x nodes < 2^31 (the plan keeps 243 x nodes of a stencil level below it)
    const unsigned nn = unsigned(C.nnodes), total = 27u * nn;
    for (unsigned pr = blockIdx.x * blockDim.x + threadIdx.x; pr < total; pr += gridDim.x * blockDim.x) {
        const int nb = int(pr / nn);
        const long long In = pr - unsigned(nb) * nn;
        int Ij, Ii, Ik;
        mg_decode(C, In, Ij, Ii, Ik);
        const int Jj = Ij + nb % 3 - 1, Ji = Ii + (nb / 3) % 3 - 1, Jk = Ik + nb / 9 - 1;
        double acc[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) acc[e] = 0.0;
        if (Jj >= 0 && Jj < C.NJ && Ji >= 0 && Ji < C.NI && Jk >= 0 && Jk < C.NK) {
            const int Ej0 = max(max(Ij, Jj) - 1, 0), Ej1 = min(min(Ij, Jj), C.ny - 1);
            const int Ei0 = max(max(Ii, Ji) - 1, 0), Ei1 = min(min(Ii, Ji), C.nx - 1);
            const int Ek0 = max(max(Ik, Jk) - 1, 0), Ek1 = min(min(Ik, Jk), C.nz - 1);
            for (int Ek = Ek0; Ek <= Ek1; ++Ek)
                for (int Ei = Ei0; Ei <= Ei1; ++Ei)
                    for (int Ej = Ej0; Ej <= Ej1; ++Ej) {
                        const int A = corner_of(Ii - Ei, Ij - Ej, Ik - Ek);
                        const int B = corner_of(Ji - Ei, Jj - Ej, Jk - Ek);
                        // the children's scales and masks, loaded together
                        double sv[8];
                        unsigned mk[8];
                        bool partial = false;
#pragma unroll
                        for (int ch = 0; ch < 8; ++ch) {
                            const int cx = ch & 1, cy = (ch >> 1) & 1, cz = ch >> 2;
                            const bool on = cx < rx && cy < ry && cz < rz;
                            const long long fe = on ? ((long long)(Ek * rz + cz) * g.nx + (Ei * rx + cx)) * g.ny +
                                                          (Ej * ry + cy)
                                                    : 0;
                            sv[ch] = on ? s_e[fe] : 0.0;
                            mk[ch] = on ? emask[fe] : 0u;
                            partial = partial || (mk[ch] != 0xFFFFFFu && mk[ch] != 0u);
                        }
                        if (!partial) {
                            // every child fully free (or absent): the fixed blocks
#pragma unroll
                            for (int ch = 0; ch < 8; ++ch) {
                                if (mk[ch] != 0xFFFFFFu) continue;
                                const double* m = M + ch * 576 + (3 * A) * 24 + 3 * B;
#pragma unroll
                                for (int c = 0; c < 3; ++c)
#pragma unroll
                                    for (int d = 0; d < 3; ++d) acc[3 * c + d] += sv[ch] * m[c * 24 + d];
                            }
                        } else {
                            // a child with fixed DOFs (boundary elements): its block from K0
                            // under its free mask, the interpolation weights of its corners
                            // formed once; the children in the same order as above
                            for (int ch = 0; ch < 8; ++ch) {
                                const int cx = ch & 1, cy = (ch >> 1) & 1, cz = ch >> 2;
                                if (cx >= rx || cy >= ry || cz >= rz) continue;
                                const long long fe = ((long long)(Ek * rz + cz) * g.nx + (Ei * rx + cx)) * g.ny +
                                                     (Ej * ry + cy);
                                const double svc = s_e[fe];
                                const unsigned mask = emask[fe];
                                if (mask == 0xFFFFFFu) {
                                    const double* m = M + ch * 576 + (3 * A) * 24 + 3 * B;
#pragma unroll
                                    for (int c = 0; c < 3; ++c)
#pragma unroll
                                        for (int d = 0; d < 3; ++d) acc[3 * c + d] += svc * m[c * 24 + d];
                                } else if (mask) {
                                    // B's corner weights as exponents (w = 2^-k; k >= 16: 0),
                                    // five bits each, so the loops below stay rolled
                                    unsigned long long pbc = 0;
#pragma unroll
                                    for (int bb = 0; bb < 8; ++bb) {
                                        const unsigned k = mg_wexp(cx + OX(bb), OX(B), rx) +
                                                           mg_wexp(cy + OY(bb), OY(B), ry) +
                                                           mg_wexp(cz + OZ(bb), OZ(B), rz);
                                        pbc |= (unsigned long long)min(k, 31u) << (5 * bb);
                                    }
                                    for (int a = 0; a < 8; ++a) {
                                        const double pa = mg_w(cx + OX(a), OX(A), rx) * mg_w(cy + OY(a), OY(A), ry) *
                                                          mg_w(cz + OZ(a), OZ(A), rz);
                                        if (pa == 0.0) continue;
                                        for (int bb = 0; bb < 8; ++bb) {
                                            const unsigned k = unsigned(pbc >> (5 * bb)) & 31u;
                                            if (k >= 16u) continue;
                                            const double w = svc * pa * ldexp(1.0, -int(k));
#pragma unroll
                                            for (int c = 0; c < 3; ++c) {
                                                if (!((mask >> (3 * a + c)) & 1u)) continue;
#pragma unroll
                                                for (int d = 0; d < 3; ++d)
                                                    if ((mask >> (3 * bb + d)) & 1u)
                                                        acc[3 * c + d] += w * sK0[(3 * a + c) * 24 + 3 * bb + d];
                                            }
                                        }
                                    }
                                }
                            }
                        }
                    }
        }
#pragma unroll
        for (int e = 0; e < 9; ++e) C.S[(long long)(nb * 9 + e) * C.nnodes + In] = float(acc[e]);
    }
}

// Stencil of level l+1 from the stencil of level l (A_{l+1} = P^T A_l P; A_l
// is zero at its masked DOFs, so P needs no mask here): 1-32 lanes per stencil
// entry (coarse node I, neighbour block nb, entry q = 3 c + d of the block).
__global__ void __launch_bounds__(256) k_mg_rap(MgLevel F, MgLevel C, int rx, int ry, int rz, int lg) {
    // 2^lg lanes per entry (G = 1..32, so that a launch has about a million
    // threads): lane l takes the fine nodes at offsets o = l, l + G, ... < 27,
    // (o % 3, o / 3 % 3, o / 9) - 1, of I's support (offsets along kept axes
    // and outside the fine grid take nothing) and sums their stencil
    // neighbours inside J's support; the lanes' sums are added by a fixed
    // butterfly (deterministic for a given G).  With one thread walking all 27
    // fine nodes, the small levels' entries were a serial chain of loads.
    const long long total = 243LL * C.nnodes;
    const int G = 1 << lg;
    const int l = threadIdx.x & (G - 1);
    // (the grid stride is a multiple of G: the lanes of a group stay together)
    for (long long t = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> lg; t < total;
         t += ((long long)gridDim.x * blockDim.x) >> lg) {
        const int e = int(unsigned(t) / unsigned(C.nnodes));   // nb * 9 + q
        const int nb = e / 9, q = e % 9;
        const long long In = t - (long long)e * C.nnodes;
        int Ij, Ii, Ik;
        mg_decode(C, In, Ij, Ii, Ik);
        const int Jj = Ij + nb % 3 - 1, Ji = Ii + (nb / 3) % 3 - 1, Jk = Ik + nb / 9 - 1;
        double acc = 0.0;
        if (Jj >= 0 && Jj < C.NJ && Ji >= 0 && Ji < C.NI && Jk >= 0 && Jk < C.NK)
        for (int o = l; o < 27; o += G) {
            const int oj = o % 3 - 1, oi = (o / 3) % 3 - 1, ok = o / 9 - 1;
            const int fj = Ij * ry + oj, fi = Ii * rx + oi, fk = Ik * rz + ok;
            if (!((ry == 2 || oj == 0) && (rx == 2 || oi == 0) && (rz == 2 || ok == 0) && fj >= 0 && fj < F.NJ &&
                  fi >= 0 && fi < F.NI && fk >= 0 && fk < F.NK))
                continue;
            const double wi = (oj ? 0.5 : 1.0) * (oi ? 0.5 : 1.0) * (ok ? 0.5 : 1.0);
            const long long fn = ((long long)fk * F.NI + fi) * F.NJ + fj;
            // stencil neighbours g = f + s of f inside J's support (|g - r J| <= r - 1
            // per axis) and inside the fine grid
            const int sj0 = max(-1, max(Jj * ry - (ry - 1) - fj, -fj)),
                      sj1 = min(1, min(Jj * ry + (ry - 1) - fj, F.NJ - 1 - fj));
            const int si0 = max(-1, max(Ji * rx - (rx - 1) - fi, -fi)),
                      si1 = min(1, min(Ji * rx + (rx - 1) - fi, F.NI - 1 - fi));
            const int sk0 = max(-1, max(Jk * rz - (rz - 1) - fk, -fk)),
                      sk1 = min(1, min(Jk * rz + (rz - 1) - fk, F.NK - 1 - fk));
            for (int sk = sk0; sk <= sk1; ++sk)
                for (int si = si0; si <= si1; ++si)
                    for (int sj = sj0; sj <= sj1; ++sj) {
                        const int sn = ((sk + 1) * 3 + si + 1) * 3 + sj + 1;
                        const int uj = fj + sj - Jj * ry, ui = fi + si - Ji * rx, uk = fk + sk - Jk * rz;
                        const double w = wi * (uj ? 0.5 : 1.0) * (ui ? 0.5 : 1.0) * (uk ? 0.5 : 1.0);
                        acc += w * double(F.S[(long long)(sn * 9 + q) * F.nnodes + fn]);
                    }
        }
        for (int off = G >> 1; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, G);
        if (l == 0) C.S[(long long)e * C.nnodes + In] = float(acc);
    }
}

// The RAP of a large level (k_mg_rap with one lane per entry reads each fine
// stencil word for up to 27 entries at distant times, from every XCD: about 4 GB
// of fetch for the 137 MB level-1 stencil of a 256 x 64 x 64 grid).  Here a
// block computes one entry e (neighbour block and 3 x 3 position) for a chunk of
// MG_RAP_CHUNK consecutive coarse nodes, and the blocks are numbered so that XCD
// x walks a contiguous run of (chunk, entry) pairs, chunk-major: the 243 entries
// of a chunk run back to back on one XCD while the fine rows they read (about
// 2 MB) stay in its L2.  Same sums, in the same order, as k_mg_rap with G = 1.
constexpr int MG_RAP_CHUNK = 256;
__global__ void __launch_bounds__(MG_RAP_CHUNK) k_mg_rap_chunked(MgLevel F, MgLevel C, int rx, int ry, int rz) {
    const int nblk = int(gridDim.x);
    const int L = xcd_local_tile(int(blockIdx.x), nblk);     // (chunk, entry), entry fastest
    const int e = L % 243, chunk = L / 243;
    const long long In = (long long)chunk * MG_RAP_CHUNK + threadIdx.x;
    if (In >= C.nnodes) return;
    const int nb = e / 9, q = e % 9;
    int Ij, Ii, Ik;
    mg_decode(C, In, Ij, Ii, Ik);
    const int Jj = Ij + nb % 3 - 1, Ji = Ii + (nb / 3) % 3 - 1, Jk = Ik + nb / 9 - 1;
    double acc = 0.0;
    if (Jj >= 0 && Jj < C.NJ && Ji >= 0 && Ji < C.NI && Jk >= 0 && Jk < C.NK)
        for (int o = 0; o < 27; ++o) {
            const int oj = o % 3 - 1, oi = (o / 3) % 3 - 1, ok = o / 9 - 1;
            const int fj = Ij * ry + oj, fi = Ii * rx + oi, fk = Ik * rz + ok;
            if (!((ry == 2 || oj == 0) && (rx == 2 || oi == 0) && (rz == 2 || ok == 0) && fj >= 0 && fj < F.NJ &&
                  fi >= 0 && fi < F.NI && fk >= 0 && fk < F.NK))
                continue;
            const double wi = (oj ? 0.5 : 1.0) * (oi ? 0.5 : 1.0) * (ok ? 0.5 : 1.0);
            const long long fn = ((long long)fk * F.NI + fi) * F.NJ + fj;
            const int sj0 = max(-1, max(Jj * ry - (ry - 1) - fj, -fj)),
                      sj1 = min(1, min(Jj * ry + (ry - 1) - fj, F.NJ - 1 - fj));
            const int si0 = max(-1, max(Ji * rx - (rx - 1) - fi, -fi)),
                      si1 = min(1, min(Ji * rx + (rx - 1) - fi, F.NI - 1 - fi));
            const int sk0 = max(-1, max(Jk * rz - (rz - 1) - fk, -fk)),
                      sk1 = min(1, min(Jk * rz + (rz - 1) - fk, F.NK - 1 - fk));
            for (int sk = sk0; sk <= sk1; ++sk)
                for (int si = si0; si <= si1; ++si)
                    for (int sj = sj0; sj <= sj1; ++sj) {
                        const int sn = ((sk + 1) * 3 + si + 1) * 3 + sj + 1;
                        const int uj = fj + sj - Jj * ry, ui = fi + si - Ji * rx, uk = fk + sk - Jk * rz;
                        const double w = wi * (uj ? 0.5 : 1.0) * (ui ? 0.5 : 1.0) * (uk ? 0.5 : 1.0);
                        acc += w * double(F.S[(long long)(sn * 9 + q) * F.nnodes + fn]);
                    }
        }
    C.S[(long long)e * C.nnodes + In] = float(acc);
}

// Inverse diagonal (0 where the row is zero: a masked DOF) and the Gershgorin
// bound of a stencil level.
__global__ void __launch_bounds__(256) k_mg_diag(MgLevel L, MgConst* cst) {
    // a thread per row (n, c), its 81 stencil words loaded at once
    float best = 0.0f;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < L.ndof;
         t += (long long)gridDim.x * blockDim.x) {
        const int c = int(unsigned(t) / unsigned(L.nnodes));
        const long long n = t - (long long)c * L.nnodes;
        float w[81];
#pragma unroll
        for (int k = 0; k < 81; ++k) w[k] = L.S[(long long)((k / 3) * 9 + 3 * c + k % 3) * L.nnodes + n];
        const float dg = w[13 * 3 + c];
        const float di = dg > 0.0f ? 1.0f / dg : 0.0f;
        L.dinv[3 * n + c] = di;
        float sum = 0.0f;
#pragma unroll
        for (int k = 0; k < 81; ++k) sum += fabsf(w[k]);
        if (di != 0.0f) best = fmaxf(best, 1.0001f * sum * di);
    }
    block_max_to(&cst->lam_bits, best);
}

// Chebyshev coefficients of every level from its bound.
__global__ void k_mg_consts(MgConst* cst, int nlev) {
    const int l = threadIdx.x;
    if (l >= nlev) return;
    float lam = __uint_as_float(cst[l].lam_bits);
    if (!(lam > 0.0f)) lam = 1.0f;     // a level with no free DOF: any value
    const float a = lam / MG_CHEB_RATIO;
    const float theta = 0.5f * (lam + a), delta = 0.5f * (lam - a);
    const float sigma = theta / delta;
    // the three-term recurrence of the scaled Chebyshev polynomials (Saad,
    // Iterative Methods for Sparse Linear Systems, Alg. 12.1): rho_0 = 1/sigma,
    // rho_k = 1/(2 sigma - rho_{k-1}); step k: ca = rho_k rho_{k-1}, cb = 2 rho_k / delta
    float* c = &cst[l].c0;
    c[0] = 1.0f / theta;
    float rho = 1.0f / sigma;
    for (int k = 1; k < MG_CHEB_MAXDEG; ++k) {
        const float rn = 1.0f / (2.0f * sigma - rho);
        c[2 * k - 1] = rn * rho;
        c[2 * k] = 2.0f * rn / delta;
        rho = rn;
    }
}

// Dense matrix of the coarsest level (FP64, row-major, n = L.ndof), identity
// on masked rows; A must be zero on entry.
__global__ void __launch_bounds__(256) k_mg_dense(MgLevel L, double* __restrict__ A) {
    // a thread per (row, neighbour block): the 27 blocks of a row are
    // independent, so their loads need not wait on one another
    const long long n = L.ndof;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < 27 * n;
         t += (long long)gridDim.x * blockDim.x) {
        const long long row = t / 27;
        const int nb = int(t - row * 27);
        const long long nd = row / 3;
        const int c = int(row % 3);
        if (L.dinv[row] == 0.0f) {
            if (nb == 13) A[row * n + row] = 1.0;
            continue;
        }
        int j, i, k;
        mg_decode(L, nd, j, i, k);
        const int mj = j + nb % 3 - 1, mi = i + (nb / 3) % 3 - 1, mk = k + nb / 9 - 1;
        if (mj < 0 || mj >= L.NJ || mi < 0 || mi >= L.NI || mk < 0 || mk >= L.NK) continue;
        const long long m = ((long long)mk * L.NI + mi) * L.NJ + mj;
        for (int d = 0; d < 3; ++d)
            if (L.dinv[3 * m + d] != 0.0f)
                A[row * n + 3 * m + d] = double(L.S[(long long)(nb * 9 + 3 * c + d) * L.nnodes + nd]);
    }
}

// In-place Gauss-Jordan inversion of the SPD n x n matrix A (one block, no
// pivoting: the pivots of an SPD matrix are positive), then Ainv = float(A).
// *flag = 1 if a pivot is not above 1e-10 of its row's original diagonal (a
// safety net for an exactly singular coarse operator; the host does not use
// the V-cycle when the supports leave rigid-body motions free).  INLDS: the
// matrix is held in registers (n <= MG_LDS_GJ; dynamic LDS 40 n bytes for the
// row/column buffers and the diagonal), else in global memory (dynamic LDS
// 16 n bytes).
template <bool INLDS, int PER = 1>
__global__ void __launch_bounds__(1024) k_mg_gauss_jordan(double* __restrict__ Ag, int n, float* __restrict__ Ainv,
                                                         int* flag) {
    extern __shared__ double smem_gj[];
    __shared__ int s_bad;
    const int nn = n * n;
    if (threadIdx.x == 0) s_bad = 0;
    if constexpr (INLDS) {
        // n <= MG_LDS_GJ: thread t holds entries t + 1024 e, e < PER, of A in
        // registers (PER = ceil(n^2 / 1024): every slot a thread walks in a step
        // costs its VALU time, used or not); step k passes row k and column k through
        // LDS buffers of alternating parity, so a step needs one barrier (the
        // buffers of step k + 1 are not those any thread may still read in
        // step k).  Same arithmetic, in the same order, as the LDS form below.
        double* buf = smem_gj;               // [2][row k (n), column k (n)]
        double* diag0 = smem_gj + 4 * n;     // the original diagonal
        double a[PER];
        int ri[PER], ci[PER];
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int t = int(threadIdx.x) + 1024 * e;
            ri[e] = t < nn ? t / n : -1;
            ci[e] = t < nn ? t - (t / n) * n : -1;
            a[e] = t < nn ? Ag[t] : 0.0;
            if (t < nn && ri[e] == ci[e]) diag0[ri[e]] = a[e];
        }
        int bad = 0;
        for (int k = 0; k < n; ++k) {
            double* fr = buf + (k & 1) * 2 * n;
            double* fc = fr + n;
#pragma unroll
            for (int e = 0; e < PER; ++e) {
                if (ri[e] == k) fr[ci[e]] = a[e];
                if (ci[e] == k) fc[ri[e]] = a[e];
            }
            __syncthreads();
            double p = fc[k];
            if (!(p > 1e-10 * diag0[k])) {
                bad = 1;
                p = 1.0;
            }
            const double ip = 1.0 / p;
#pragma unroll
            for (int e = 0; e < PER; ++e) {
                const int i = ri[e], j = ci[e];
                if (i < 0) continue;
                if (i == k) a[e] = (j == k) ? ip : a[e] * ip;
                else if (j == k) a[e] = -fc[i] * ip;
                else a[e] -= fc[i] * (fr[j] * ip);
            }
        }
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int t = int(threadIdx.x) + 1024 * e;
            if (t < nn) Ainv[t] = float(a[e]);
        }
        if (threadIdx.x == 0) *flag = bad;
        return;
    } else {
        double* fcol = smem_gj;         // column k before the step
        double* diag0 = smem_gj + n;    // the original diagonal
        double* A = Ag;
        __shared__ double s_piv;
        __syncthreads();
        for (int j = threadIdx.x; j < n; j += blockDim.x) diag0[j] = A[(long long)j * n + j];
        __syncthreads();
        for (int k = 0; k < n; ++k) {
            if (threadIdx.x == 0) {
                double p = A[(long long)k * n + k];
                if (!(p > 1e-10 * diag0[k])) {
                    s_bad = 1;
                    p = 1.0;
                }
                s_piv = 1.0 / p;
            }
            __syncthreads();
            const double ip = s_piv;
            // scale row k (A[k][k] becomes 1 / pivot), keep the old column k
            for (int j = threadIdx.x; j < n; j += blockDim.x) {
                fcol[j] = A[(long long)j * n + k];
                A[(long long)k * n + j] = (j == k) ? ip : A[(long long)k * n + j] * ip;
            }
            __syncthreads();
            for (int t = threadIdx.x; t < nn; t += blockDim.x) {
                const int i = t / n, j = t - i * n;
                if (i == k) continue;
                const double f = fcol[i];
                if (j == k) A[t] = -f * ip;
                else A[t] -= f * A[(long long)k * n + j];
            }
            __syncthreads();
        }
        for (int t = threadIdx.x; t < nn; t += blockDim.x) Ainv[t] = float(A[t]);
        if (threadIdx.x == 0) *flag = s_bad;
    }
}

// Inverse of the SPD n x n matrix A (128 < n <= MG_SWEEP_MAX: a 3 x 3 x 3
// coarsest grid) by the symmetric sweep operator (Goodnight, "A tutorial on
// the SWEEP operator", The American Statistician 33, 1979): sweeping pivot k
// of a symmetric matrix, p = a_kk,
//     a_ij -= a_ik a_kj / p (i, j != k),  a_ik = a_kj = a_.k / p,  a_kk = -1 / p,
// keeps it symmetric, and sweeping every pivot leaves -A^-1.  Only the upper
// triangle is held: thread t owns entries t + 1024 e, e < PER, of the packed
// upper triangle in registers (PER = ceil(n (n+1) / 2048), half the slots of
// the square form, so it fits where the square form spills); step k passes
// row k through an LDS buffer of alternating parity, one barrier per step.
// The pivots are the Schur complements' diagonals, positive for an SPD matrix;
// *flag = 1 if one is not above 1e-10 of its original diagonal.  Dynamic LDS
// 24 n bytes (row buffers of both parities and the diagonal).
constexpr int MG_SWEEP_MAX = 192;
template <int PER>
__global__ void __launch_bounds__(1024) k_mg_sweep_sym(const double* __restrict__ Ag, int n,
                                                       float* __restrict__ Ainv, int* flag) {
    extern __shared__ double smem_sw[];
    double* buf = smem_sw;                   // [2][n]: row k of step k, by parity
    double* diag0 = smem_sw + 2 * n;
    const int nt = n * (n + 1) / 2;
    double a[PER];
    int rc[PER];                             // (row << 16) | column, row <= column; -1 past the triangle
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int t = int(threadIdx.x) + 1024 * e;
        int i = -1, j = -1;
        if (t < nt) {
            // row i holds the n - i entries (i, i .. n-1); rows before it hold
            // i n - i (i - 1) / 2 entries
            i = int((2.0 * n + 1.0 - sqrt((2.0 * n + 1.0) * (2.0 * n + 1.0) - 8.0 * t)) * 0.5);
            while (i > 0 && i * n - i * (i - 1) / 2 > t) --i;
            while ((i + 1) * n - (i + 1) * i / 2 <= t) ++i;
            j = i + (t - (i * n - i * (i - 1) / 2));
        }
        rc[e] = t < nt ? (i << 16) | j : -1;
        a[e] = t < nt ? Ag[i * n + j] : 0.0;
        if (t < nt && i == j) diag0[i] = a[e];
    }
    int bad = 0;
    for (int k = 0; k < n; ++k) {
        double* r = buf + (k & 1) * n;       // row k: r[m] = a_km
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            if (rc[e] < 0) continue;
            const int i = rc[e] >> 16, j = rc[e] & 0xFFFF;
            if (i == k) r[j] = a[e];
            else if (j == k) r[i] = a[e];
        }
        __syncthreads();
        double p = r[k];
        if (!(p > 1e-10 * diag0[k])) {
            bad = 1;
            p = 1.0;
        }
        const double ip = 1.0 / p;
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            if (rc[e] < 0) continue;
            const int i = rc[e] >> 16, j = rc[e] & 0xFFFF;
            if (i == k && j == k) a[e] = -ip;
            else if (i == k || j == k) a[e] *= ip;
            else a[e] -= r[i] * (r[j] * ip);
        }
    }
    // -A^-1 is held: both triangles of Ainv
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        if (rc[e] < 0) continue;
        const int i = rc[e] >> 16, j = rc[e] & 0xFFFF;
        const float v = float(-a[e]);
        Ainv[i * n + j] = v;
        Ainv[j * n + i] = v;
    }
    if (threadIdx.x == 0) *flag = bad;
}

// x = Ainv b on the coarsest level: one wave per row, lanes over the columns
// (Ainv is symmetric, so row r is read as column r), a fixed-order wave sum.
__global__ void __launch_bounds__(256) k_mg_dense_solve(const float* __restrict__ Ainv, int n,
                                                        const float* __restrict__ b, float* __restrict__ x,
                                                        const PcgState* st) {
    if (!mg_live(st)) return;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= n) return;
    float acc = 0.0f;
    for (int j = lane; j < n; j += 64) acc += Ainv[(long long)row * n + j] * b[j];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) x[row] = acc;
}

// Row (n, c) of y = A x on a stencil level.  Kernels over a level's DOFs use
// component-major thread indices t = c * nnodes + n, so the lanes of a wave
// read adjacent stencil words.
__device__ __forceinline__ float mg_row(const MgLevel& L, const float* __restrict__ x, long long n, int c) {
    int j, i, k;
    mg_decode(L, n, j, i, k);
    // Branch-free over all 27 neighbours, so every load of the row is issued at
    // once: the stencil blocks of neighbours outside the grid are exactly zero
    // (the setup writes them so), and their x index is clamped to the node itself.
    const long long sj = L.NJ, si = (long long)L.NI * L.NJ;
    // (one plane of 9 neighbours per pass: its 54 loads in flight together, at
    // a third of the registers of all 27)
    const float* Sc = L.S + (long long)(3 * c) * L.nnodes + n;
    float y = 0.0f;
#pragma unroll 1
    for (int dk = -1; dk <= 1; ++dk) {
        const bool kin = unsigned(k + dk) < unsigned(L.NK);
#pragma unroll
        for (int p = 0; p < 9; ++p) {
            const int dj = p % 3 - 1, di = p / 3 - 1;
            const bool in = kin && unsigned(j + dj) < unsigned(L.NJ) && unsigned(i + di) < unsigned(L.NI);
            const long long m = 3 * (in ? n + dk * si + di * sj + dj : n);
            const float* s = Sc + (long long)(((dk + 1) * 9 + p) * 9) * L.nnodes;
            y += s[0] * x[m] + s[L.nnodes] * x[m + 1] + s[2 * L.nnodes] * x[m + 2];
        }
    }
    return y;
}

// The stencil-level sweeps of the cycle (k_mg_cheb_st, k_mg_resid_st) on a
// small level are bound by dependent memory round trips, not bytes: a row with
// its three neighbour planes summed one after another, then the pointwise
// words, then the stores, behind a load of the solve state, is five round
// trips.  Here four lanes (a quad) share one row: lane dk + 1 sums neighbour
// plane dk (9 blocks, 27 stencil and 27 x words, all in flight together; lane
// 3 repeats plane 0 and drops it), every lane loads the row's pointwise words
// and the solve state at the same time, the quad adds its three plane sums in
// a fixed order (DPP quad swaps: every lane gets the same bits) and lane 3
// stores the result.  The stores are bounds-checked buffer stores whose offset
// is out of range once the solve has stopped, so no load waits for the state.
// Plane dk of row (n, c): its 27 stencil words and 9 x triples, loaded all at
// once (mg_plane_load), then summed (mg_plane_sum).  Branch-free as mg_row:
// blocks of neighbours outside the grid are stored zeros and their x index is
// clamped to the node itself.
struct MgPlane {
    float s[9][3], x[9][3];
};
__device__ __forceinline__ void mg_plane_load(const MgLevel& L, const float* __restrict__ x, long long n, int c,
                                              int dk, MgPlane& P) {
    int j, i, k;
    mg_decode(L, n, j, i, k);
    const long long sj = L.NJ, si = (long long)L.NI * L.NJ;
    const bool kin = unsigned(k + dk) < unsigned(L.NK);
    const float* Sc = L.S + (long long)(3 * c + (dk + 1) * 81) * L.nnodes + n;
#pragma unroll
    for (int p = 0; p < 9; ++p) {
        const int dj = p % 3 - 1, di = p / 3 - 1;
        const bool in = kin && unsigned(j + dj) < unsigned(L.NJ) && unsigned(i + di) < unsigned(L.NI);
        const long long m = 3 * (in ? n + dk * si + di * sj + dj : n);
        const float* sp = Sc + (long long)(p * 9) * L.nnodes;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            P.s[p][d] = sp[d * L.nnodes];
            P.x[p][d] = x[m + d];
        }
    }
}
__device__ __forceinline__ float mg_plane_sum(const MgPlane& P) {
    float y = 0.0f;
#pragma unroll
    for (int p = 0; p < 9; ++p) y += P.s[p][0] * P.x[p][0] + P.s[p][1] * P.x[p][1] + P.s[p][2] * P.x[p][2];
    return y;
}

// Sum of the plane sums of lanes 0..2 of a quad (lane 3 passes 0), the same
// bits in all four lanes.
__device__ __forceinline__ float mg_quad_sum(float v) {
    const float a = v + __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0xB1, 0xF, 0xF,
                                                                         false));   // quad_perm [1,0,3,2]
    return a + __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, a), 0x4E, 0xF, 0xF,
                                                                  false));          // quad_perm [2,3,0,1]
}

// Vector (per-lane) loads of the solve state and the smoother constants.  The
// address is offset by a zero the compiler cannot see through, so the value
// is not taken as wave-uniform: it stays in a vector register, joins the data
// loads' memory counter and is only waited for where it is used (a store's
// offset), instead of forcing a wait for a scalar branch before the data loads
// are issued.
__device__ __forceinline__ int opaque_zero() {
    int z;
    asm volatile("v_mov_b32 %0, 0" : "=v"(z));
    return z;
}
__device__ __forceinline__ bool mg_live_v(const PcgState* st) {
    if (!st) return true;
    return __hip_atomic_load(const_cast<int*>(&st->status) + opaque_zero(), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT) == ST_RUNNING;
}
__device__ __forceinline__ float ld_vf(const float* p) {
    return __hip_atomic_load(const_cast<float*>(p) + opaque_zero(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Chebyshev step on a stencil level: x_out = x_in + d, d = ca d + cb dinv (b - A x_in).
// QUAD (small levels, latency-bound): a quad per row, 4 ndof threads.  Row
// form (large levels, where a wave's loads should cover 64 consecutive
// stencil words): a thread per row, the three planes summed in turn as
// mg_row, with the pointwise words, the constants and the solve state loaded
// before the first plane.  Both grid-stride.
template <bool QUAD>
__global__ void __launch_bounds__(256) k_mg_cheb_st(MgLevel L, const float* __restrict__ xin, float* __restrict__ xout,
                                                    const MgConst* cst, int step, const PcgState* st) {
    const unsigned vbytes = unsigned(L.ndof) * sizeof(float);
    const int lane = QUAD ? int(threadIdx.x & 3) : 3;
    const long long nt = QUAD ? 4 * L.ndof : L.ndof;
    // QUAD: 4 ndof is a multiple of 4 and so is every quad's first thread: a
    // quad is wholly in or out of range, and its DPP sums see only active lanes
    for (long long t4 = (long long)blockIdx.x * blockDim.x + threadIdx.x; t4 < nt;
         t4 += (long long)gridDim.x * blockDim.x) {
        const long long t = QUAD ? t4 >> 2 : t4;
        const int c = int(unsigned(t) / unsigned(L.nnodes));
        const long long n = t - c * L.nnodes, q = 3 * n + c;
        float y;
        if constexpr (QUAD) {
            MgPlane P;
            mg_plane_load(L, xin, n, c, lane == 3 ? 0 : lane - 1, P);
            const float dq = L.d[q], dv = L.dinv[q], bv = L.b[q], xv = xin[q];
            const float ca = ld_vf(&cst->c1a), cb = ld_vf(step ? &cst->c1b : &cst->c0);
            const bool live = mg_live_v(st);
            __builtin_amdgcn_sched_barrier(0);     // every load above is issued before any arithmetic
            y = mg_quad_sum(lane == 3 ? 0.0f : mg_plane_sum(P));
            const float dn = (step ? ca * dq : 0.0f) + cb * dv * (bv - y);
            const int off = (live && lane == 3) ? int(sizeof(float) * q) : OOB;
            Buf<float>::st(dn, buf_rsrc(L.d, vbytes), off);
            Buf<float>::st(xv + dn, buf_rsrc(xout, vbytes), off);
        } else {
            const bool live = mg_live_v(st);
            const float ca = ld_vf(&cst->c1a), cb = ld_vf(step ? &cst->c1b : &cst->c0);
            const float dq = L.d[q], dv = L.dinv[q], bv = L.b[q], xv = xin[q];
            y = mg_row(L, xin, n, c);
            const float dn = (step ? ca * dq : 0.0f) + cb * dv * (bv - y);
            const int off = live ? int(sizeof(float) * q) : OOB;
            Buf<float>::st(dn, buf_rsrc(L.d, vbytes), off);
            Buf<float>::st(xv + dn, buf_rsrc(xout, vbytes), off);
        }
    }
}

// Masked residual r = (dinv != 0) (b - A x) on a stencil level (the input of
// the restriction).  QUAD or row form, as k_mg_cheb_st.
template <bool QUAD>
__global__ void __launch_bounds__(256) k_mg_resid_st(MgLevel L, const float* __restrict__ x, float* __restrict__ r,
                                                     const PcgState* st) {
    const unsigned vbytes = unsigned(L.ndof) * sizeof(float);
    const int lane = QUAD ? int(threadIdx.x & 3) : 3;
    const long long nt = QUAD ? 4 * L.ndof : L.ndof;
    for (long long t4 = (long long)blockIdx.x * blockDim.x + threadIdx.x; t4 < nt;
         t4 += (long long)gridDim.x * blockDim.x) {
        const long long t = QUAD ? t4 >> 2 : t4;
        const int c = int(unsigned(t) / unsigned(L.nnodes));
        const long long n = t - c * L.nnodes, q = 3 * n + c;
        if constexpr (QUAD) {
            MgPlane P;
            mg_plane_load(L, x, n, c, lane == 3 ? 0 : lane - 1, P);
            const float dv = L.dinv[q], bv = L.b[q];
            const bool live = mg_live_v(st);
            __builtin_amdgcn_sched_barrier(0);
            const float y = mg_quad_sum(lane == 3 ? 0.0f : mg_plane_sum(P));
            const int off = (live && lane == 3) ? int(sizeof(float) * q) : OOB;
            Buf<float>::st(dv != 0.0f ? bv - y : 0.0f, buf_rsrc(r, vbytes), off);
        } else {
            const bool live = mg_live_v(st);
            const float dv = L.dinv[q], bv = L.b[q];
            const float y = mg_row(L, x, n, c);
            Buf<float>::st(dv != 0.0f ? bv - y : 0.0f, buf_rsrc(r, vbytes), live ? int(sizeof(float) * q) : OOB);
        }
    }
}

// Levels up to this many DOFs run the stencil sweeps in the quad form.
constexpr long long MG_QUAD_DOF = 32768;

// Coarse node In of the restriction bC = P^T rF (rF the fine level's masked
// residual), all three components: the 27 fine nodes around the node's fine
// position (weights 1, 1/2, 1/4, 1/8 by the number of offset axes; a kept axis
// has only the centre), each a 12-byte read.  mg_restrict_load issues all 27
// reads, mg_restrict_sum adds them.
struct MgRestrict {
    float v[27][3], w[27];
};
// OPAQUE: the clamped index passes through an empty asm statement, so the
// compiler cannot tell that absent neighbours read the centre; otherwise it
// loads the centre once and branches around every other load, which puts a
// memory wait between them.  All 27 reads are then issued at once.
template <bool OPAQUE>
__device__ __forceinline__ void mg_restrict_load(const MgLevel& F, const MgLevel& C, int rx, int ry, int rz,
                                                 const float* __restrict__ rF, long long In, MgRestrict& R) {
    int Ij, Ii, Ik;
    mg_decode(C, In, Ij, Ii, Ik);
    const int fj = Ij * ry, fi = Ii * rx, fk = Ik * rz;
    // all 27 offsets, branch-free: weight 0 (and the index clamped to the
    // centre) for an offset along a kept axis or outside the grid
#pragma unroll
    for (int o = 0; o < 27; ++o) {
        const int oj = o % 3 - 1, oi = (o / 3) % 3 - 1, ok = o / 9 - 1;
        // (bitwise operators: no short-circuit branches between the loads)
        const bool in = ((oj == 0) | ((ry == 2) & (unsigned(fj + oj) < unsigned(F.NJ)))) &
                        ((oi == 0) | ((rx == 2) & (unsigned(fi + oi) < unsigned(F.NI)))) &
                        ((ok == 0) | ((rz == 2) & (unsigned(fk + ok) < unsigned(F.NK))));
        R.w[o] = in ? (oj ? 0.5f : 1.0f) * (oi ? 0.5f : 1.0f) * (ok ? 0.5f : 1.0f) : 0.0f;
        long long f = 3 * (in ? ((long long)(fk + ok) * F.NI + (fi + oi)) * F.NJ + (fj + oj)
                              : ((long long)fk * F.NI + fi) * F.NJ + fj);
        if constexpr (OPAQUE) asm volatile("" : "+v"(f));
#pragma unroll
        for (int c = 0; c < 3; ++c) R.v[o][c] = rF[f + c];
    }
}
__device__ __forceinline__ void mg_restrict_sum(const MgRestrict& R, float (&acc)[3]) {
    acc[0] = acc[1] = acc[2] = 0.0f;
#pragma unroll
    for (int o = 0; o < 27; ++o)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[c] += R.w[o] * R.v[o][c];
}
__device__ __forceinline__ void mg_restrict_node(const MgLevel& F, const MgLevel& C, int rx, int ry, int rz,
                                                 const float* __restrict__ rF, long long In, float (&acc)[3]) {
    MgRestrict R;
    mg_restrict_load<false>(F, C, rx, ry, rz, rF, In, R);
    mg_restrict_sum(R, acc);
}

// Fine node fn of the coarse correction xF += M_F P xC.  BUF: xF is global
// memory, stored through bounds-checked buffer stores (dropped unless live);
// otherwise plain stores (the bottom kernel's LDS views).
template <bool BUF>
__device__ __forceinline__ void mg_prolong_node(const MgLevel& F, const MgLevel& C, int rx, int ry, int rz,
                                                float* __restrict__ xF, const float* __restrict__ xC,
                                                const float* __restrict__ maskF, long long fn,
                                                const PcgState* st = nullptr) {
    int fj, fi, fk;
    mg_decode(F, fn, fj, fi, fk);
    // parents per axis: f = r I (weight 1) or, on a halved axis at an odd f,
    // I = (f -+ 1) / 2 (0.5 each); all 8 candidates read, the absent second
    // parent of an axis with weight 0 at the first one's index
    const int j0 = fj / ry, i0 = fi / rx, k0 = fk / rz;
    const bool tj = ry == 2 && (fj & 1), ti = rx == 2 && (fi & 1), tk = rz == 2 && (fk & 1);
    float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const int e = a & 1, b = (a >> 1) & 1, g = a >> 2;
        const float w = (tj ? 0.5f : (e ? 0.0f : 1.0f)) * (ti ? 0.5f : (b ? 0.0f : 1.0f)) *
                        (tk ? 0.5f : (g ? 0.0f : 1.0f));
        const long long I = 3 * (((long long)(k0 + (tk ? g : 0)) * C.NI + (i0 + (ti ? b : 0))) * C.NJ +
                                 (j0 + (tj ? e : 0)));
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[c] += w * xC[I + c];
    }
    if constexpr (BUF) {
        // (live = false: the solve has stopped; the stores are dropped by the
        // buffer range check, so no load waits for the solve state)
        const unsigned vbytes = unsigned(F.ndof) * sizeof(float);
        float m[3], xv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            m[c] = maskF[3 * fn + c];
            xv[c] = xF[3 * fn + c];
        }
        const bool live = mg_live_v(st);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const long long q = 3 * fn + c;
            const int off = (live && m[c] != 0.0f) ? int(sizeof(float) * q) : OOB;
            Buf<float>::st(xv[c] + acc[c], buf_rsrc(xF, vbytes), off);
        }
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c)
            if (maskF[3 * fn + c] != 0.0f) xF[3 * fn + c] += acc[c];
    }
}

// Restriction bC = P^T rF, one thread per coarse node; with pre0 the coarse
// level's first smoothing step from zero follows at once in the same thread:
// x0 = d = c0 dinv b (cstC = the coarse level's constants).
__global__ void __launch_bounds__(256) k_mg_restrict(MgLevel F, MgLevel C, int rx, int ry, int rz,
                                                     const float* __restrict__ rF, const MgConst* cstC, int pre0,
                                                     const PcgState* st) {
    const unsigned vbytes = unsigned(C.ndof) * sizeof(float);
    for (long long In = (long long)blockIdx.x * blockDim.x + threadIdx.x; In < C.nnodes;
         In += (long long)gridDim.x * blockDim.x) {
        MgRestrict R;
        mg_restrict_load<true>(F, C, rx, ry, rz, rF, In, R);
        float b[3], dv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) dv[c] = C.dinv[3 * In + c];
        // the solve state and c0 are loaded beside the data and only gate or
        // scale the stores
        const float c0 = ld_vf(&cstC->c0);
        const bool live = mg_live_v(st);
        __builtin_amdgcn_sched_barrier(0);
        mg_restrict_sum(R, b);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const long long q = 3 * In + c;
            const int off = live ? int(sizeof(float) * q) : OOB;
            Buf<float>::st(b[c], buf_rsrc(C.b, vbytes), off);
            if (pre0) {
                const float d = c0 * dv[c] * b[c];
                Buf<float>::st(d, buf_rsrc(C.d, vbytes), off);
                Buf<float>::st(d, buf_rsrc(C.x0, vbytes), off);
            }
        }
    }
}

// Coarse correction: xF += M_F P xC.
__global__ void __launch_bounds__(256) k_mg_prolong(MgLevel F, MgLevel C, int rx, int ry, int rz,
                                                    float* __restrict__ xF, const float* __restrict__ xC,
                                                    const float* __restrict__ maskF, const PcgState* st) {
    for (long long fn = (long long)blockIdx.x * blockDim.x + threadIdx.x; fn < F.nnodes;
         fn += (long long)gridDim.x * blockDim.x)
        mg_prolong_node<true>(F, C, rx, ry, rz, xF, xC, maskF, fn, st);
}

// The small levels of a cycle in one workgroup: levels lb .. nl-1 (the last
// the coarsest, inverted densely), given b of level lb; leaves that level's
// result in its x1.  Phases are separated by workgroup barriers.  Staged form
// (lds_words > 0): the levels' stencils, inverse diagonals and vectors and the
// dense inverse are copied into LDS once per launch (all loads of a thread in
// flight together), every phase then runs on LDS, and only level lb's x1 is
// written back.  Unstaged form: the phases work on the levels' global vectors
// (all waves of a workgroup share one CU and its L1, so their global stores
// are visible to each other after the barrier).
constexpr int MG_BOTTOM_LDS_BYTES = 152 * 1024;   // LDS budget of the staged form

struct MgBottom {
    MgLevel L[MG_MAX_LEVELS];
    int lb, nl;
    const MgConst* cst;
    const float* Ainv;
    // staged form: word offset of each level's block in LDS (a level above
    // the coarsest: S, dinv, b, x0, x1, d, y; the coarsest: b, x1, Ainv)
    int lds_off[MG_MAX_LEVELS];
    int lds_words;
};

// LDS words of the staged form of levels lb .. nl-1 (offsets into off); every
// array starts at a multiple of 4 words (16-byte copies).
__host__ __device__ constexpr long long mg_pad4(long long w) { return (w + 3) & ~3LL; }
__host__ __device__ inline long long mg_bottom_layout(const MgLevel* L, int lb, int nl, int* off) {
    long long w = 0;
    for (int l = lb; l < nl; ++l) {
        if (off) off[l] = int(w);
        const long long n = mg_pad4(L[l].ndof);
        w += (l + 1 < nl) ? mg_pad4(243 * L[l].nnodes) + 6 * n : mg_pad4(L[l].ndof * L[l].ndof) + 2 * n;
    }
    return w;
}

template <bool STAGED>
__device__ __forceinline__ MgLevel mg_bottom_view(const MgBottom& B, float* smem, int l) {
    MgLevel L = B.L[l];
    if constexpr (STAGED) {
        float* p = smem + B.lds_off[l];
        const long long n = mg_pad4(L.ndof);
        if (l + 1 < B.nl) {
            L.S = p;
            p += mg_pad4(243 * L.nnodes);
            L.dinv = p;
            L.b = p + n;
            L.x0 = p + 2 * n;
            L.x1 = p + 3 * n;
            L.d = p + 4 * n;
            L.y = p + 5 * n;
        } else {
            // (every pointer of a staged view points into LDS, so the compiler
            // addresses all of them as LDS; the coarsest uses b and x1 only)
            L.S = L.dinv = L.x0 = L.d = L.y = p;
            L.b = p;
            L.x1 = p + n;      // the dense inverse follows at p + 2 n
        }
    }
    return L;
}

// dst[0..cnt) = src[0..cnt) (both 16-byte aligned): 16-byte loads, 8 of them
// in flight per thread, then the tail words
__device__ __forceinline__ void mg_stage(float* dst, const float* __restrict__ src, long long cnt) {
    constexpr int U = 8;
    const long long n4 = cnt >> 2;
    const float4* s4 = reinterpret_cast<const float4*>(src);
    float4* d4 = reinterpret_cast<float4*>(dst);
    for (long long i0 = threadIdx.x; i0 < n4; i0 += (long long)U * blockDim.x) {
        float4 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long i = i0 + (long long)u * blockDim.x;
            v[u] = i < n4 ? s4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long i = i0 + (long long)u * blockDim.x;
            if (i < n4) d4[i] = v[u];
        }
    }
    for (long long i = 4 * n4 + threadIdx.x; i < cnt; i += blockDim.x) dst[i] = src[i];
}

__device__ __forceinline__ void mg_bottom_cheb(const MgLevel& L, const MgConst& k, const float* xin, float* xout,
                                               int step) {
    const float ca = step ? k.c1a : 0.0f, cb = step ? k.c1b : k.c0;
    for (long long t = threadIdx.x; t < L.ndof; t += blockDim.x) {
        const int c = int(unsigned(t) / unsigned(L.nnodes));
        const long long n = t - c * L.nnodes, q = 3 * n + c;
        const float dn = (step ? ca * L.d[q] : 0.0f) + cb * L.dinv[q] * (L.b[q] - mg_row(L, xin, n, c));
        L.d[q] = dn;
        xout[q] = xin[q] + dn;
    }
}

template <bool STAGED>
__global__ void __launch_bounds__(MG_BOTTOM_THREADS) k_mg_bottom(MgBottom B, const PcgState* st) {
    extern __shared__ float mg_smem[];
    float* smem = mg_smem;
    if (!mg_live(st)) return;
    const float* Ainv = B.Ainv;
    if constexpr (STAGED) {
        for (int l = B.lb; l + 1 < B.nl; ++l) {
            const MgLevel V = mg_bottom_view<true>(B, smem, l);
            mg_stage(V.S, B.L[l].S, 243 * B.L[l].nnodes);
            mg_stage(V.dinv, B.L[l].dinv, V.ndof);
        }
        {
            const MgLevel V = mg_bottom_view<true>(B, smem, B.lb);
            mg_stage(V.b, B.L[B.lb].b, V.ndof);
            const MgLevel Vc = mg_bottom_view<true>(B, smem, B.nl - 1);
            float* sA = Vc.x1 + mg_pad4(Vc.ndof);
            mg_stage(sA, B.Ainv, Vc.ndof * Vc.ndof);
            Ainv = sA;
        }
        __syncthreads();
    }
    for (int l = B.lb; l + 1 < B.nl; ++l) {
        const MgLevel F = mg_bottom_view<STAGED>(B, smem, l);
        const MgLevel C = mg_bottom_view<STAGED>(B, smem, l + 1);
        const MgConst k = B.cst[l];
        for (long long q = threadIdx.x; q < F.ndof; q += blockDim.x) {
            const float d = k.c0 * F.dinv[q] * F.b[q];
            F.d[q] = d;
            F.x0[q] = d;
        }
        __syncthreads();
        mg_bottom_cheb(F, k, F.x0, F.x1, 1);
        __syncthreads();
        for (long long t = threadIdx.x; t < F.ndof; t += blockDim.x) {
            const int c = int(unsigned(t) / unsigned(F.nnodes));
            const long long n = t - c * F.nnodes, q = 3 * n + c;
            const float y = mg_row(F, F.x1, n, c);
            F.y[q] = F.dinv[q] != 0.0f ? F.b[q] - y : 0.0f;
        }
        __syncthreads();
        for (long long In = threadIdx.x; In < C.nnodes; In += blockDim.x) {
            float b[3];
            mg_restrict_node(F, C, F.rx, F.ry, F.rz, F.y, In, b);
#pragma unroll
            for (int c = 0; c < 3; ++c) C.b[3 * In + c] = b[c];
        }
        __syncthreads();
    }
    {
        // coarsest: x1 = Ainv b (Ainv symmetric), a fixed-order sum per row
        const MgLevel Lc = mg_bottom_view<STAGED>(B, smem, B.nl - 1);
        const int n = int(Lc.ndof);
        if constexpr (STAGED) {
            // thread per row, column j of Ainv read as row j (adjacent threads, adjacent words)
            for (int row = threadIdx.x; row < n; row += blockDim.x) {
                float acc = 0.0f;
                for (int j = 0; j < n; ++j) acc += Ainv[j * n + row] * Lc.b[j];
                Lc.x1[row] = acc;
            }
        } else {
            // one wave per row, lanes over the columns
            const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
            for (int row = wv; row < n; row += nw) {
                float acc = 0.0f;
                for (int j = lane; j < n; j += 64) acc += Ainv[(long long)row * n + j] * Lc.b[j];
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
                if (lane == 0) Lc.x1[row] = acc;
            }
        }
        __syncthreads();
    }
    for (int l = B.nl - 2; l >= B.lb; --l) {
        const MgLevel F = mg_bottom_view<STAGED>(B, smem, l);
        const MgLevel C = mg_bottom_view<STAGED>(B, smem, l + 1);
        const MgConst k = B.cst[l];
        for (long long fn = threadIdx.x; fn < F.nnodes; fn += blockDim.x)
            mg_prolong_node<false>(F, C, F.rx, F.ry, F.rz, F.x1, C.x1, F.dinv, fn);
        __syncthreads();
        mg_bottom_cheb(F, k, F.x1, F.x0, 0);
        __syncthreads();
        mg_bottom_cheb(F, k, F.x0, F.x1, 1);
        __syncthreads();
    }
    if constexpr (STAGED) {
        const MgLevel V = mg_bottom_view<true>(B, smem, B.lb);
        for (long long q = threadIdx.x; q < V.ndof; q += blockDim.x) B.L[B.lb].x1[q] = V.x1[q];
    }
}

// r.z of the FP32 iterations with the V-cycle's z; the last block sets beta
// (mg_rz_finish).
__global__ void __launch_bounds__(256) k_mg_rz(long long n, const float* __restrict__ r, const float* __restrict__ z,
                                               PcgState* st, double* partials) {
    __shared__ double sRed[16];
    if (st->status != ST_RUNNING) return;
    double v[1] = {0.0};
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        v[0] += double(r[i]) * double(z[i]);
    double tot[1];
    if (reduce_last_block<1>(v, partials, &st->cnt[3], tot, sRed) && threadIdx.x == 0) mg_rz_finish(st, tot[0]);
}

}  // namespace h8
