// h8_recovery.h -- stress / strain recovery on the structured brick grid.
//
// The reference evaluates sigma = D_e B(p) u_e at the 2x2x2 Gauss points and
// moves it to the corners with a lumped projection (fe/recovery.py:234-330):
//     sigma_a = sum_e sum_p N_c(p) sigma_e(p) |J| w / sum_e sum_p N_c(p) |J| w
// over the elements e sharing node a (c = the node's corner in e); the element
// value is the volume average.  On the brick every element has the same
// Jacobian, the Gauss weights are 1 and sum_p N_c(p) = 1, so with D_e = s_e D
//     sigma_a = D (1/n_a) sum_e s_e eps_c(u_e),   eps_c = sum_p N_c(p) B(p) u_e,
// n_a = the number of elements touching node a.  eps_c is the symmetric part of
// the projected displacement gradient G_c[d][i] = sum_b A[c][d][b] u_b[i] with
//     A[c][d][b] = sum_p N_c(p) dN_b/dx_d(p)          (corner projection, c < 8)
//     A[8][d][b] = (1/8) sum_p dN_b/dx_d(p)           (volume average)
// so a node costs 72 multiply-adds per touching element plus one 6x6 product,
// instead of a 6x24 product per Gauss point.  The table holds A (216 values)
// followed by D (36 values; the identity for strain).
#pragma once
#include "h8_device.h"
#include "h8_node.h"

namespace h8 {

constexpr int REC_A = 9 * 3 * 8;           // A[c][d][b]
constexpr int REC_WORDS = REC_A + 36;      // then D (row-major)

// One thread per table entry: A from the brick spacing, then D (or the identity).
__global__ void __launch_bounds__(256) k_recovery_table(double hx, double hy, double hz, const double* D,
                                                        double* tab) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < REC_A) {
        const int c = t / 24, d = (t / 8) % 3, b = t % 8;
        const double g = 0.57735026918962576451;      // 1/sqrt(3)
        const double h = d == 0 ? hx : (d == 1 ? hy : hz);
        const double sb[3] = {OX(b) ? 1.0 : -1.0, OY(b) ? 1.0 : -1.0, OZ(b) ? 1.0 : -1.0};
        double acc = 0.0;
        for (int p = 0; p < 8; ++p) {
            const double xp[3] = {OX(p) ? g : -g, OY(p) ? g : -g, OZ(p) ? g : -g};
            // dN_b/dxi_d = (1/8) s_d prod_{e != d} (1 + s_e x_e); dN/dx_d = (2/h_d) dN/dxi_d
            double dn = 0.125 * sb[d];
            for (int e = 0; e < 3; ++e)
                if (e != d) dn *= 1.0 + sb[e] * xp[e];
            dn *= 2.0 / h;
            double w;
            if (c < 8) {
                const double sc[3] = {OX(c) ? 1.0 : -1.0, OY(c) ? 1.0 : -1.0, OZ(c) ? 1.0 : -1.0};
                w = 0.125 * (1.0 + sc[0] * xp[0]) * (1.0 + sc[1] * xp[1]) * (1.0 + sc[2] * xp[2]);
            } else {
                w = 0.125;
            }
            acc += w * dn;
        }
        tab[t] = acc;
    } else if (t < REC_WORDS) {
        const int i = t - REC_A;
        tab[t] = D ? D[i] : ((i / 6 == i % 6) ? 1.0 : 0.0);
    }
}

// Voigt strain [xx, yy, zz, xy, yz, xz] (engineering shear) of a gradient
// G[d][i] = du_i/dx_d, then D eps, then the von Mises value of the result
// (fe/recovery.py:364-375).
__device__ __forceinline__ void rec_finish(const __attribute__((address_space(4))) double* D, const double (&G)[3][3],
                                           double scale, double (&sig)[6], double& vm) {
    const double e[6] = {G[0][0], G[1][1], G[2][2], G[1][0] + G[0][1], G[2][1] + G[1][2], G[2][0] + G[0][2]};
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        double a = 0.0;
#pragma unroll
        for (int t = 0; t < 6; ++t) a += D[6 * s + t] * e[t];
        sig[s] = scale * a;
    }
    const double dxy = sig[0] - sig[1], dyz = sig[1] - sig[2], dzx = sig[2] - sig[0];
    vm = sqrt(0.5 * (dxy * dxy + dyz * dyz + dzx * dzx) + 3.0 * (sig[3] * sig[3] + sig[4] * sig[4] + sig[5] * sig[5]));
}

// Nodal recovery: one block per TJ x TI x TK node brick (one thread per node);
// the brick's displacements plus a one-node halo and the scales of the
// (TJ+1)(TI+1)(TK+1) elements touching it are staged in LDS.  Every node sums
// its elements in a fixed order (deterministic).  s_e NULL: every scale is 1.
template <int TJ, int TI, int TK>
__global__ void __launch_bounds__(TJ* TI* TK) k_recover(GridDev g, const double* tab, const double* s_e,
                                                       const double* u, double* nodal, double* vm_nodal) {
    constexpr int NT = TJ * TI * TK;
    constexpr int HJ = TJ + 2, HI = TI + 2, HK = TK + 2, HN = HJ * HI * HK;
    constexpr int EJ = TJ + 1, EI = TI + 1, EK = TK + 1, EN = EJ * EI * EK;
    __shared__ double sU[3 * HN];
    __shared__ double sS[EN];
    int j0, i0, k0;
    tile_origin(g, xcd_local_tile(blockIdx.x, gridDim.x), TJ, TI, TK, j0, i0, k0);
    for (int t = threadIdx.x; t < 3 * HN; t += NT) {
        const int hn = t / 3, c = t - 3 * hn;
        const int hj = hn % HJ, hi = (hn / HJ) % HI, hk = hn / (HJ * HI);
        const int j = j0 - 1 + hj, i = i0 - 1 + hi, k = k0 - 1 + hk;
        const bool in = j >= 0 && j < g.NJ && i >= 0 && i < g.NI && k >= 0 && k < g.NK;
        sU[t] = in ? u[3LL * (((long long)k * g.NI + i) * g.NJ + j) + c] : 0.0;
    }
    for (int t = threadIdx.x; t < EN; t += NT) {
        const int ej = t % EJ, ei = (t / EJ) % EI, ek = t / (EJ * EI);
        const int je = j0 - 1 + ej, ie = i0 - 1 + ei, ke = k0 - 1 + ek;
        const bool in = je >= 0 && je < g.ny && ie >= 0 && ie < g.nx && ke >= 0 && ke < g.nz;
        sS[t] = in ? (s_e ? s_e[((long long)ke * g.nx + ie) * g.ny + je] : 1.0) : 0.0;
    }
    __syncthreads();
    const int lj = threadIdx.x % TJ, li = (threadIdx.x / TJ) % TI, lk = threadIdx.x / (TJ * TI);
    const int j = j0 + lj, i = i0 + li, k = k0 + lk;
    if (j >= g.NJ || i >= g.NI || k >= g.NK) return;
    const auto A = as_const(tab);
    double H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    int cnt = 0;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int ex = OX(e), ey = OY(e), ez = OZ(e);     // element (i-1+ex, j-1+ey, k-1+ez)
        const int ie = i - 1 + ex, je = j - 1 + ey, ke = k - 1 + ez;
        if (ie < 0 || ie >= g.nx || je < 0 || je >= g.ny || ke < 0 || ke >= g.nz) continue;
        const int c = corner_of(1 - ex, 1 - ey, 1 - ez);  // this node's corner in the element
        double G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const int h = 3 * (((lk + ez + OZ(b)) * HI + (li + ex + OX(b))) * HJ + (lj + ey + OY(b)));
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double a = A[(c * 3 + d) * 8 + b];
#pragma unroll
                for (int q = 0; q < 3; ++q) G[d][q] += a * sU[h + q];
            }
        }
        const double s = sS[((lk + ez) * EI + (li + ex)) * EJ + (lj + ey)];
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int q = 0; q < 3; ++q) H[d][q] += s * G[d][q];
        ++cnt;
    }
    double sig[6], vm;
    rec_finish(A + REC_A, H, 1.0 / double(cnt), sig, vm);
    const long long node = ((long long)k * g.NI + i) * g.NJ + j;
    if (nodal) {
#pragma unroll
        for (int s = 0; s < 6; ++s) nodal[6 * node + s] = sig[s];
    }
    if (vm_nodal) vm_nodal[node] = vm;
}

// Element averages: one thread per element, sigma_e = s_e D Bbar u_e.
__global__ void __launch_bounds__(256) k_recover_elem(GridDev g, const double* tab, const double* s_e,
                                                      const double* u, double* elem, double* vm_elem) {
    const long long m = (long long)g.nx * g.ny * g.nz;
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    const int je = int(e % g.ny), ie = int((e / g.ny) % g.nx), ke = int(e / ((long long)g.ny * g.nx));
    const auto A = as_const(tab) + 8 * 24;                 // the volume-average rows
    double G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const long long nb = ((long long)(ke + OZ(b)) * g.NI + (ie + OX(b))) * g.NJ + (je + OY(b));
        const double ub[3] = {u[3 * nb], u[3 * nb + 1], u[3 * nb + 2]};
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int q = 0; q < 3; ++q) G[d][q] += A[d * 8 + b] * ub[q];
    }
    double sig[6], vm;
    rec_finish(as_const(tab) + REC_A, G, s_e ? s_e[e] : 1.0, sig, vm);
    if (elem) {
#pragma unroll
        for (int s = 0; s < 6; ++s) elem[6 * e + s] = sig[s];
    }
    if (vm_elem) vm_elem[e] = vm;
}

}  // namespace h8
