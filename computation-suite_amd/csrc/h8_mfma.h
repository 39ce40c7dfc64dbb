// h8_mfma.h -- element-batched matrix-free K*u on MFMA for the structured grid.
//
// A block owns a TJ x TI x TK brick of nodes and evaluates every element that
// touches it: Y = K0 * P_batch on the matrix cores, where the 24 x 16 operand
// P_batch holds the gathered 24-vectors of 16 elements (one column each) and
// K0 (24 rows padded to 32, i.e. two 16-row blocks) stays in registers as the
// A operand for the whole block (12 fragments).  One 16x16x4 MFMA per
// (row block, k step): 12 per batch of 16 elements.  The 24 outputs of each
// element, scaled by s_e, are added into an LDS accumulator over the node
// halo.  Elements are processed in 8 parity colours with a barrier between
// colours: two elements of one colour never share a node, so the LDS adds
// are race-free without atomics and the summation order is fixed (the result
// is deterministic).  Owned nodes are then written out.  The same fused modes
// as the node-centric kernel (h8_node.h) are provided.
//
// MFMA operand layout used (16x16x4, wave64): lane l supplies A[l%16][l/16]
// and B[l/16][l%16]; it receives D[4*(l/16)+r][l%16] (f32) or
// D[4*r+l/16][l%16] (f64), r = 0..3 (checked with a lane-layout probe).  Every LDS and
// global index below is formed from in-range loop counters, so any layout
// mismatch would show up as wrong values in the parity tests, never as an
// out-of-bounds access.
#pragma once
#include "h8_device.h"
#include "h8_node.h"

namespace h8 {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
// Output row held by (lane, r) of a 16x16x4 MFMA (column = lane % 16).
template <> struct Mfma<float> {
    typedef f32x4 acc_t;
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
    static __device__ __forceinline__ acc_t run(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma<double> {
    typedef f64x4 acc_t;
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * r + (lane >> 4); }
    static __device__ __forceinline__ acc_t run(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

template <typename T, int TJ, int TI, int TK, int MODE>
__global__ void __launch_bounds__(256)
k_mfma(NodeArgs<T> A) {
    constexpr int NT = 256;
    constexpr int HJ = TJ + 2, HI = TI + 2, HK = TK + 2;
    constexpr int HPL = HJ * HI * HK;
    constexpr int EJ = TJ + 1, EI = TI + 1, EK = TK + 1;
    constexpr int NE = EJ * EI * EK;
    constexpr int NOWN = TJ * TI * TK;
    typedef typename Mfma<T>::acc_t acc_t;

    __shared__ T sP[3 * HPL];
    __shared__ T sA[3 * HPL];
    __shared__ T sS[NE];
    __shared__ double sRed[16 * 4];

    const GridDev& g = A.g;
    if constexpr (is_dir(MODE)) {
        if (A.st->status != ST_RUNNING) return;
    }
    if constexpr (MODE == MODE_RESID) {
        if (A.only_on_replace && A.st->status != ST_REPLACE) return;
    }
    int j0, i0, k0;
    tile_origin(g, xcd_local_tile(blockIdx.x, gridDim.x), TJ, TI, TK, j0, i0, k0);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;

    T beta = T(0), alpha_prev = T(0);
    int first = 1, pending = 0;
    T* p_new = nullptr;
    const T* p_old = nullptr;
    if constexpr (is_dir(MODE)) {
        beta = T(A.st->beta);
        first = A.st->first;
        alpha_prev = T(A.st->alpha);
        pending = A.st->pending;
        const bool odd = (A.st->k & 1) != 0;
        p_new = odd ? A.p1 : A.p0;
        p_old = odd ? A.p0 : A.p1;
    }

    // ---- K0 as the A operand: rows 16*rb + l%16 (< 24), columns 4*s + l/16
    T a_frag[2][6];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            const int row = 16 * rb + (lane & 15);
            const int col = 4 * s + (lane >> 4);
            a_frag[rb][s] = row < 24 ? A.K0[row * 24 + col] : T(0);
        }

    // ---- stage the halo and element scales (same layout as the node kernel), zero accumulators
    stage_halo<T, TJ, TI, TK, MODE, NT>(A, j0, i0, k0, beta, first, p_old, p_new, alpha_prev, pending, sP, sS, sA);
    __syncthreads();

    // ---- 8 parity colours; batches of 16 same-colour elements per MFMA pass
#pragma unroll 1
    for (int colour = 0; colour < 8; ++colour) {
        const int pj = colour & 1, pi = (colour >> 1) & 1, pk = colour >> 2;
        const int cj = (EJ - pj + 1) / 2, ci = (EI - pi + 1) / 2, ck = (EK - pk + 1) / 2;
        const int count = cj * ci * ck;
        const int nbatch = (count + 15) / 16;
#pragma unroll 1
        for (int b = wave; b < nbatch; b += NT / 64) {
            const int t = b * 16 + (lane & 15);
            const bool valid = t < count;
            const int tt = valid ? t : 0;
            const int ej = pj + 2 * (tt % cj);
            const int ei = pi + 2 * ((tt / cj) % ci);
            const int ek = pk + 2 * (tt / (cj * ci));
            const T se = valid ? sS[(ek * EI + ei) * EJ + ej] : T(0);
            acc_t acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < 6; ++s) {
                const int kdof = 4 * s + (lane >> 4);   // element DOF 0..23
                const int a = kdof / 3, d = kdof - 3 * (kdof / 3);
                // offset by the element's first corner (K0 annihilates translations)
                const T bv = sP[3 * ((ek + OZ(a)) * HI * HJ + (ei + OX(a)) * HJ + ej + OY(a)) + d] -
                             sP[3 * (ek * HI * HJ + ei * HJ + ej) + d];
                acc0 = Mfma<T>::run(a_frag[0][s], bv, acc0);
                acc1 = Mfma<T>::run(a_frag[1][s], bv, acc1);
            }
            if (valid) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row0 = Mfma<T>::row(lane, r);        // rows 0..15
                    {
                        const int a = row0 / 3, c = row0 - 3 * (row0 / 3);
                        const int li = 3 * ((ek + OZ(a)) * HI * HJ + (ei + OX(a)) * HJ + ej + OY(a)) + c;
                        sA[li] += se * acc0[r];
                    }
                    const int row1 = 16 + row0;                    // rows 16..31 (valid < 24)
                    if (row1 < 24) {
                        const int a = row1 / 3, c = row1 - 3 * (row1 / 3);
                        const int li = 3 * ((ek + OZ(a)) * HI * HJ + (ei + OX(a)) * HJ + ej + OY(a)) + c;
                        sA[li] += se * acc1[r];
                    }
                }
            }
        }
        __syncthreads();
    }

    // ---- owned nodes: write out (+ fused reductions)
    double part[2] = {0.0, 0.0};
#pragma unroll 1
    for (int o = threadIdx.x; o < NOWN; o += NT) {
        const int tj = o % TJ, ti = (o / TJ) % TI, tk = o / (TJ * TI);
        const int j = j0 + tj, i = i0 + ti, k = k0 + tk;
        if (!(j < g.NJ && i < g.NI && k < g.NK)) continue;
        const long long gn = 3LL * ((long long)k * g.NI * g.NJ + (long long)i * g.NJ + j);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int li = 3 * ((tk + 1) * HI * HJ + (ti + 1) * HJ + tj + 1) + c;
            const T qv = sA[li];
            if constexpr (MODE == MODE_PLAIN) {
                A.y[gn + c] = qv;
            } else if constexpr (is_dir(MODE)) {
                A.q[gn + c] = qv;
                part[0] += double(sP[li]) * double(qv);
            } else {
                const double msk = A.dinv[gn + c];
                const double r64 = (msk != 0.0) ? (A.f[gn + c] - double(qv)) : 0.0;
                const float r32 = float(r64);
                A.r32[gn + c] = r32;
                part[0] += r64 * r64;
                part[1] += double(r32) * double(A.dinv32[gn + c] * r32);
            }
        }
    }

    if constexpr (is_dir(MODE)) {
        double v[1] = {part[0]}, tot[1];
        if (reduce_last_block<1>(v, A.partials, &A.st->cnt[0], tot, sRed)) {
            if (threadIdx.x == 0) dir_finish(A.st, tot[0]);
        }
    } else if constexpr (MODE == MODE_RESID) {
        double v[2] = {part[0], part[1]}, tot[2];
        if (reduce_last_block<2>(v, A.partials, &A.st->cnt[2], tot, sRed)) {
            if (threadIdx.x == 0) resid_finish(A.st, tot[0], tot[1]);
        }
    }
}

}  // namespace h8
