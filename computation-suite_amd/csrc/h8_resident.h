// h8_resident.h -- the whole Jacobi-PCG loop in one cooperative launch, with
// every block's share of the vectors resident in LDS.
//
// On grids up to about a million DOFs an iteration of the streaming kernels is
// bound by the fixed cost of its two launches (dispatch, ramp-up, drain and
// the last-block reduction), not by the bytes it moves.  Here one block per CU
// owns a brick of nodes and keeps, for the whole solve, r, x, dinv and q of
// its nodes, p of its nodes plus a one-node halo shell, and the scales of the
// elements touching its nodes in LDS.  An iteration exchanges only the
// preconditioned residual z = dinv r of brick-surface nodes (device-scope
// stores and loads, coherent across the XCDs' L2 caches) and takes two
// grid-wide sums.
//
// Iteration (the Hestenes-Stiefel recurrence of the streaming kernels,
// fe/iterative_solvers.py:438-469):
//   p = z + beta p_old on the halo brick; z of shell nodes comes from the
//       exchange, and every block forms p at a shared node with the same
//       operation (an explicit fma), so all copies of p agree bitwise
//   q = K p for the owned nodes: every element touching an owned node
//       (elements on a brick boundary are computed by both bricks), element
//       products through the reflection sectors, outputs added to q corner
//       by corner (8 rounds in a fixed order: deterministic)
//   p.q -> grid sum -> alpha, or stop (not SPD)
//   x += alpha p, r = mask(r - alpha q), z = dinv r, publish surface z;
//   r.z, r.r -> grid sum -> beta, or stop (update_finish)
//
// Grid sums: every block publishes its partial in its slot of a ring entry and
// polls all blocks' slots of that entry, then sums them in block order, so
// every block holds the same total.  A slot holds RING_EMPTY (a signalling
// NaN, which arithmetic never produces) until written.  Entries rotate over 3
// iterations; after the sum of iteration j a block empties its own slot of the
// entry used in iteration j - 1, which every block has read by then, and that
// store completes before the block's next publication, so no block sees a
// stale value when the entry comes round again.  The ring is filled with
// RING_EMPTY before each launch.  Polls are bounded: a block that waits more
// than SPIN_LIMIT polls raises the error word and leaves; the other blocks
// then leave within the same bound, so a launch always ends.  The launch is
// cooperative (all blocks resident at once), so in a correct run no poll
// waits for a block that has not started.
#pragma once
#include "h8_device.h"
#include "h8_node.h"
#include "h8_sector.h"
#include "h8_stream.h"

namespace h8 {

constexpr unsigned long long RING_EMPTY = 0x7FF4DEADBEEF0001ull;
constexpr long long SPIN_LIMIT = 1LL << 22;

// Brick split of the node grid: PJ x PI x PK parts, part a of P over N nodes
// spans [a N / P, (a + 1) N / P).  LDS layouts use the largest extents:
//   p     [BK + 2][HI][HJ][3]   halo brick (HI = BI + 2, HJ = BJ + 2)
//   r, x, dinv, q [BK][BI][BJ][3]  owned nodes
//   s     element scales of the block's (clipped) element range, j fastest
struct ResGeom {
    int PI, PJ, PK;
    int BI, BJ, BK;
    int HI, HJ;
    int nown, nhalo, nelem;
    int o_p, o_r, o_x, o_d, o_q, o_s;
    int lds_bytes;
};

template <typename T>
struct ResArgs {
    GridDev g;
    ResGeom G;
    const T* s_e;
    const T* kh;       // 72-entry sector table
    const T* dinv;
    T* x;
    T* r;
    T* p0;             // direction buffers: iteration k's p is written to p[k & 1]
    T* p1;
    T* zx;             // z exchange (one vector)
    double* ring;      // [3][nb] (p.q), then [3][nb][2] (r.z, r.r)
    unsigned* err;     // set to 1 when a poll exceeded SPIN_LIMIT
    PcgState* st;
};

template <typename T>
__device__ __forceinline__ T ld_dev(const T* p) {
    return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_dev(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

// DPP lane move of a double (two 32-bit halves); lanes of rows outside
// ROW_MASK receive 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, int(unsigned(u)), CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, int(unsigned(u >> 32)), CTRL, ROW_MASK, 0xF, false);
    return __builtin_bit_cast(double, (unsigned long long)(unsigned)lo | ((unsigned long long)(unsigned)hi << 32));
}

// Sum over the 64 lanes of a wave in a fixed order, without LDS: quad swaps,
// rotations within rows of 16 (row_ror: lane l takes lane l - k), then the
// row broadcasts, which leave the total in lane 63; every lane returns it.
// All 64 lanes must be active.
__device__ __forceinline__ double wave_total(double v) {
    v += dpp_f64<0xB1, 0xF>(v);    // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E, 0xF>(v);    // quad_perm [2,3,0,1]
    v += dpp_f64<0x124, 0xF>(v);   // row_ror:4
    v += dpp_f64<0x128, 0xF>(v);   // row_ror:8   (row sums in every lane)
    v += dpp_f64<0x142, 0xA>(v);   // row_bcast:15 into rows 1 and 3
    v += dpp_f64<0x143, 0xC>(v);   // row_bcast:31 into rows 2 and 3
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = __builtin_amdgcn_readlane(unsigned(u), 63);
    const unsigned hi = __builtin_amdgcn_readlane(unsigned(u >> 32), 63);
    return __builtin_bit_cast(double, (unsigned long long)lo | ((unsigned long long)hi << 32));
}

// Grid-wide sum of NV doubles through one ring entry ([nb][NV]); `done`: the
// entry used one iteration earlier, whose slot of this block is emptied.
// Every wave reduces its lanes, wave 0 adds the waves, publishes the block's
// partial, polls the nb slots (lane l: slots l, l + 64, ... in order) and
// reduces them, so `tot` holds the same total in wave 0 of every block.
// Returns false in wave 0 if a poll timed out (true in the other waves, which
// return after the first barrier).  The caller reads `tot` in thread 0 and
// then synchronises the block.
template <int NV>
__device__ bool grid_sum(const double (&v)[NV], double* entry, double* done, double* sh, double (&tot)[NV]) {
    // the block's earlier device-scope stores (z exchange, an emptied slot)
    // complete before its partial is published
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const double w = wave_total(v[i]);
        if (lane == 0) sh[wv * NV + i] = w;
    }
    __syncthreads();
    if (wv != 0) return true;
    const int nb = gridDim.x;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const double bsum = wave_total(lane < nw ? sh[lane * NV + i] : 0.0);
        if (lane == 0) st_dev(&entry[blockIdx.x * NV + i], bsum);
    }
    double acc[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = 0.0;
    int good = 1;
    for (int b = lane; b < nb; b += 64) {
        double val[NV];
        for (long long spin = 0;; ++spin) {
            bool all = true;
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                val[i] = ld_dev(&entry[b * NV + i]);
                all = all && __builtin_bit_cast(unsigned long long, val[i]) != RING_EMPTY;
            }
            if (all) break;
            if (spin >= SPIN_LIMIT) {
                good = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
#pragma unroll
        for (int i = 0; i < NV; ++i) acc[i] += val[i];
    }
#pragma unroll
    for (int i = 0; i < NV; ++i) tot[i] = wave_total(acc[i]);
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i)
            st_dev(reinterpret_cast<unsigned long long*>(&done[blockIdx.x * NV + i]), RING_EMPTY);
    }
    return __all(good) != 0;
}

template <typename T, int NT, bool REFINED>
__global__ void __launch_bounds__(NT) k_resident(ResArgs<T> A) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ PcgState S;
    __shared__ double sRed[16 * 2];
    __shared__ int sFail;
    constexpr bool OFFS = sizeof(T) == 4;   // subtract the corner-0 value first (FP32 rounding)
    const GridDev& g = A.g;
    const ResGeom& G = A.G;
    T* sP = reinterpret_cast<T*>(smem + G.o_p);
    T* sR = reinterpret_cast<T*>(smem + G.o_r);
    T* sX = reinterpret_cast<T*>(smem + G.o_x);
    T* sD = reinterpret_cast<T*>(smem + G.o_d);
    T* sQ = reinterpret_cast<T*>(smem + G.o_q);
    T* sS = reinterpret_cast<T*>(smem + G.o_s);
    if (threadIdx.x == 0) {
        S = *A.st;
        sFail = 0;
    }
    __syncthreads();
    if (S.status != ST_RUNNING) return;

    // the block's node brick, halo brick and element range
    const int b = blockIdx.x;
    const int pj = b % G.PJ, pi = (b / G.PJ) % G.PI, pk = b / (G.PJ * G.PI);
    const int j0 = pj * g.NJ / G.PJ, j1 = (pj + 1) * g.NJ / G.PJ;
    const int i0 = pi * g.NI / G.PI, i1 = (pi + 1) * g.NI / G.PI;
    const int k0 = pk * g.NK / G.PK, k1 = (pk + 1) * g.NK / G.PK;
    const int bj = j1 - j0, bi = i1 - i0, bk = k1 - k0;
    const int hj = bj + 2, hi = bi + 2;
    const int nh = hj * hi * (bk + 2), no = bj * bi * bk;
    const int ej0 = max(j0 - 1, 0), ei0 = max(i0 - 1, 0), ek0 = max(k0 - 1, 0);
    const int cj = min(j1, g.ny) - ej0, ci = min(i1, g.nx) - ei0, ck = min(k1, g.nz) - ek0;
    const int ne = cj * ci * ck;
    auto gdof = [&](int i, int j, int k) { return 3LL * (((long long)k * g.NI + i) * g.NJ + j); };

    // ---- load: owned r, x, dinv; p_old on the halo brick; element scales
    for (int t = threadIdx.x; t < no; t += NT) {
        const int lj = t % bj, li = (t / bj) % bi, lk = t / (bj * bi);
        const int o = 3 * ((lk * G.BI + li) * G.BJ + lj);
        const long long gd = gdof(i0 + li, j0 + lj, k0 + lk);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            sR[o + c] = A.r[gd + c];
            sX[o + c] = A.x[gd + c];
            sD[o + c] = A.dinv[gd + c];
        }
    }
    {
        const T* pg = ((S.k - 1) & 1) ? A.p1 : A.p0;
        const bool first = S.first != 0;
        for (int t = threadIdx.x; t < nh; t += NT) {
            const int hj_ = t % hj, hi_ = (t / hj) % hi, hk_ = t / (hj * hi);
            const int j = j0 - 1 + hj_, i = i0 - 1 + hi_, k = k0 - 1 + hk_;
            const bool in = unsigned(j) < unsigned(g.NJ) && unsigned(i) < unsigned(g.NI) && unsigned(k) < unsigned(g.NK);
            const int o = 3 * ((hk_ * G.HI + hi_) * G.HJ + hj_);
            const long long gd = in ? gdof(i, j, k) : 0;
#pragma unroll
            for (int c = 0; c < 3; ++c) sP[o + c] = (in && !first) ? pg[gd + c] : T(0);
        }
    }
    for (int t = threadIdx.x; t < ne; t += NT) {
        const int lj = t % cj, li = (t / cj) % ci, lk = t / (cj * ci);
        sS[t] = A.s_e[((long long)(ek0 + lk) * g.nx + (ei0 + li)) * g.ny + (ej0 + lj)];
    }
    __syncthreads();

    const int nb = gridDim.x;
    double* ring1 = A.ring;
    double* ring2 = A.ring + 3 * nb;
    long long kexec = S.k;
    bool ok = true;
    for (int it = 0;; ++it) {
        const T beta = T(S.beta);
        const bool first = S.first != 0;
        kexec = S.k;
        // ---- p on the halo brick, q = 0 (the exchange loads of up to PF nodes
        // per thread are issued before any is used)
        constexpr int PF = 4;
        for (int base = 0; base < nh; base += PF * NT) {
            T zv[PF][3];
            int ho[PF];
#pragma unroll
            for (int u = 0; u < PF; ++u) {
                const int t = base + u * NT + int(threadIdx.x);
                ho[u] = -1;
#pragma unroll
                for (int c = 0; c < 3; ++c) zv[u][c] = T(0);
                const int hj_ = t % hj, hi_ = (t / hj) % hi, hk_ = t / (hj * hi);
                const int j = j0 - 1 + hj_, i = i0 - 1 + hi_, k = k0 - 1 + hk_;
                if (t < nh && unsigned(j) < unsigned(g.NJ) && unsigned(i) < unsigned(g.NI) &&
                    unsigned(k) < unsigned(g.NK)) {
                    ho[u] = 3 * ((hk_ * G.HI + hi_) * G.HJ + hj_);
                    const bool own = hj_ >= 1 && hj_ <= bj && hi_ >= 1 && hi_ <= bi && hk_ >= 1 && hk_ <= bk;
                    const long long gd = gdof(i, j, k);
                    if (own) {
                        const int oo = 3 * (((hk_ - 1) * G.BI + (hi_ - 1)) * G.BJ + (hj_ - 1));
#pragma unroll
                        for (int c = 0; c < 3; ++c) zv[u][c] = sD[oo + c] * sR[oo + c];
                    } else if (it == 0) {   // before the first exchange
#pragma unroll
                        for (int c = 0; c < 3; ++c) zv[u][c] = A.dinv[gd + c] * A.r[gd + c];
                    } else {
#pragma unroll
                        for (int c = 0; c < 3; ++c) zv[u][c] = ld_dev(&A.zx[gd + c]);
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < PF; ++u) {
                if (ho[u] < 0) continue;
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    sP[ho[u] + c] = first ? zv[u][c] : fma_t(beta, sP[ho[u] + c], zv[u][c]);
            }
        }
        for (int t = threadIdx.x; t < 3 * G.nown; t += NT) sQ[t] = T(0);
        __syncthreads();

        // ---- q = K p: element products, outputs added corner by corner
        for (int e0 = 0; e0 < ne; e0 += NT) {
            const int t = e0 + int(threadIdx.x);
            const bool ev = t < ne;
            const int tt = ev ? t : 0;
            const int lj = tt % cj, li = (tt / cj) % ci, lk = tt / (cj * ci);
            const int ie = ei0 + li, je = ej0 + lj, ke = ek0 + lk;
            const int hb = ((ke - k0 + 1) * G.HI + (ie - i0 + 1)) * G.HJ + (je - j0 + 1);
            T U[8][3];
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                const int h = 3 * (hb + (OZ(a) * G.HI + OX(a)) * G.HJ + OY(a));
#pragma unroll
                for (int c = 0; c < 3; ++c) U[a][c] = sP[h + c];
            }
            if constexpr (OFFS) {
#pragma unroll
                for (int a = 1; a < 8; ++a)
#pragma unroll
                    for (int c = 0; c < 3; ++c) U[a][c] -= U[0][c];
#pragma unroll
                for (int c = 0; c < 3; ++c) U[0][c] = T(0);
            }
            sector_product<T>(A.kh, U, ev ? sS[tt] : T(0));
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                const int ni = ie + OX(a) - i0, nj = je + OY(a) - j0, nk = ke + OZ(a) - k0;
                if (ev && unsigned(ni) < unsigned(bi) && unsigned(nj) < unsigned(bj) && unsigned(nk) < unsigned(bk)) {
                    const int o = 3 * ((nk * G.BI + ni) * G.BJ + nj);
#pragma unroll
                    for (int c = 0; c < 3; ++c) sQ[o + c] += U[a][c];
                }
                __syncthreads();
            }
        }

        // ---- p.q -> alpha
        const int e = it % 3, ed = (it + 2) % 3;
        double pq[1] = {0.0};
        for (int t = threadIdx.x; t < no; t += NT) {
            const int lj = t % bj, li = (t / bj) % bi, lk = t / (bj * bi);
            const int o = 3 * ((lk * G.BI + li) * G.BJ + lj);
            const int h = 3 * (((lk + 1) * G.HI + li + 1) * G.HJ + lj + 1);
#pragma unroll
            for (int c = 0; c < 3; ++c) pq[0] += double(sP[h + c]) * double(sQ[o + c]);
        }
        {
            double tot[1];
            const bool good = grid_sum<1>(pq, ring1 + e * nb, ring1 + ed * nb, sRed, tot);
            if (threadIdx.x == 0) {
                if (good) dir_finish(&S, tot[0]);
                else sFail = 1;
            }
        }
        __syncthreads();
        if (sFail) {
            ok = false;
            break;
        }
        if (S.status != ST_RUNNING) break;
        const T alpha = T(S.alpha);

        // ---- x += alpha p, r = mask(r - alpha q), z = dinv r (surface z published)
        double rz[2] = {0.0, 0.0};
        for (int t = threadIdx.x; t < no; t += NT) {
            const int lj = t % bj, li = (t / bj) % bi, lk = t / (bj * bi);
            const int o = 3 * ((lk * G.BI + li) * G.BJ + lj);
            const int h = 3 * (((lk + 1) * G.HI + li + 1) * G.HJ + lj + 1);
            const bool surf = (lj == 0 && j0 > 0) || (lj == bj - 1 && j1 < g.NJ) || (li == 0 && i0 > 0) ||
                              (li == bi - 1 && i1 < g.NI) || (lk == 0 && k0 > 0) || (lk == bk - 1 && k1 < g.NK);
            const long long gd = gdof(i0 + li, j0 + lj, k0 + lk);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T d = sD[o + c];
                sX[o + c] += alpha * sP[h + c];
                const T ri = (d != T(0)) ? T(sR[o + c] - alpha * sQ[o + c]) : T(0);
                sR[o + c] = ri;
                const T z = d * ri;
                rz[0] += double(ri) * double(z);
                rz[1] += double(ri) * double(ri);
                if (surf) st_dev(&A.zx[gd + c], z);
            }
        }
        {
            double tot[2];
            const bool good = grid_sum<2>(rz, ring2 + 2 * e * nb, ring2 + 2 * ed * nb, sRed, tot);
            if (threadIdx.x == 0) {
                if (good) update_finish<REFINED>(&S, tot[0], tot[1]);
                else sFail = 1;
            }
        }
        __syncthreads();
        if (sFail) {
            ok = false;
            break;
        }
        if (S.status != ST_RUNNING) break;
    }
    if (!ok) {
        if (threadIdx.x == 0) st_dev(A.err, 1u);
        return;
    }
    // ---- write back the owned x, r and the direction of the last iteration
    T* pw = (kexec & 1) ? A.p1 : A.p0;
    for (int t = threadIdx.x; t < no; t += NT) {
        const int lj = t % bj, li = (t / bj) % bi, lk = t / (bj * bi);
        const int o = 3 * ((lk * G.BI + li) * G.BJ + lj);
        const int h = 3 * (((lk + 1) * G.HI + li + 1) * G.HJ + lj + 1);
        const long long gd = gdof(i0 + li, j0 + lj, k0 + lk);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            A.x[gd + c] = sX[o + c];
            A.r[gd + c] = sR[o + c];
            pw[gd + c] = sP[h + c];
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S.pending = 0;          // x is updated in the loop
        *A.st = S;
    }
}

// Ring entries -> RING_EMPTY, error word -> 0 (before each launch).
__global__ void __launch_bounds__(256) k_ring_reset(unsigned long long* ring, int n, unsigned* err) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) ring[i] = RING_EMPTY;
    if (blockIdx.x == 0 && threadIdx.x == 0) *err = 0u;
}

}  // namespace h8
