// h8_resident.h -- the whole Jacobi-PCG loop in one cooperative launch, with
// every block's share of the vectors resident in LDS.
//
// On grids up to about a million DOFs an iteration of the streaming kernels is
// bound by the fixed cost of its two launches (dispatch, ramp-up, drain and
// the last-block reduction), not by the bytes it moves.  Here one block per CU
// owns a brick of nodes and keeps, for the whole solve, r, x, dinv and q of
// its nodes, p of its nodes plus a one-node halo shell, and the scales of the
// elements touching its nodes in LDS.  An iteration exchanges only the
// preconditioned residual z = dinv r of brick-surface nodes and takes two
// grid-wide sums, all through device-scope stores and loads (coherent across
// the XCDs' L2 caches).
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
//   r.z, r.r -> grid sum -> beta, or stop (update_finish); meanwhile every
//       block collects the z of its shell nodes
//
// Slots and sentinels.  Every exchanged word -- a block's partial sum, a
// surface node's z -- lives in a ring of 3 entries (iteration j uses entry
// j % 3) and holds EMPTY (a signalling-NaN pattern, which arithmetic never
// produces) until its owner writes it, so a reader polls the word itself: a
// value that is not EMPTY is the current one.  An owner empties its words of
// the entry used one iteration earlier once every block is known to have read
// them (after the next grid sum), and those stores complete before the
// owner's next p.q publication, which every reader observes before it polls
// that entry again.  The rings are filled with EMPTY before each launch.
// Polls are bounded: a block that waits more than SPIN_LIMIT polls raises the
// error word and leaves; the other blocks then leave within the same bound, so
// a launch always ends.  The launch is cooperative (all blocks resident at
// once), so in a correct run no poll waits for a block that has not started.
#pragma once
#include "h8_device.h"
#include "h8_node.h"
#include "h8_sector.h"
#include "h8_stream.h"

namespace h8 {

constexpr unsigned long long RING_EMPTY = 0x7FF4DEADBEEF0001ull;
constexpr long long SPIN_LIMIT = 1LL << 22;
constexpr int RES_MAX_BLOCKS = 256;   // grid sums poll up to 4 slots per lane of one wave
// Every partial is published in RES_REPL copies, RES_RSTRIDE doubles apart (4 KiB:
// different memory channels), and block b polls copy b % RES_REPL: the polls of
// all blocks spread over the copies instead of all reading the same words.
constexpr int RES_REPL = 8;
constexpr int RES_RSTRIDE = 512;
constexpr int RES_ENTRY = RES_REPL * RES_RSTRIDE;             // doubles per ring entry
constexpr int RES_RING_DOUBLES = 2 * 3 * RES_ENTRY;           // p.q ring, then (r.z, r.r) ring

// EMPTY pattern of an exchanged vector word (a signalling NaN of T).
template <typename T> struct Empty;
template <> struct Empty<float> {
    typedef unsigned U;
    static constexpr U bits = 0x7FA00001u;
};
template <> struct Empty<double> {
    typedef unsigned long long U;
    static constexpr U bits = RING_EMPTY;
};

// Brick split of the node grid: PJ x PI x PK parts, part a of P over N nodes
// spans [a N / P, (a + 1) N / P).  LDS layouts use the largest extents:
//   p, z  [BK + 2][HI][HJ][3]   halo brick (HI = BI + 2, HJ = BJ + 2); z: shell nodes
//   r, x, dinv, q [BK][BI][BJ][3]  owned nodes
//   s     element scales of the block's (clipped) element range, j fastest
struct ResGeom {
    int PI, PJ, PK;
    int BI, BJ, BK;
    int HI, HJ;
    int nown, nhalo, nelem;
    int o_p, o_z, o_r, o_x, o_d, o_q, o_s;
    int lds_bytes;
};

template <typename T>
struct ResArgs {
    GridDev g;
    ResGeom G;
    const T* s_e;
    const T* kh;       // 72-entry sector table
    const double* kh64; // the same table in FP64 (FP64 element products of FP32 data)
    const T* dinv;
    T* x;
    T* r;
    T* p0;             // direction buffers: iteration k's p is written to p[k & 1]
    T* p1;
    T* zr[3];          // z exchange ring (three vectors)
    double* ring;      // RES_RING_DOUBLES: p.q ring, then (r.z, r.r) ring; entry [RES_REPL][RES_RSTRIDE]
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

// Exact t / d for t, d < 2^16 through a 64-bit multiply by m = ceil(2^32 / d):
// t m / 2^32 = t / d + t e / (d 2^32) with 0 <= e < d, and t e / 2^32 < 1 / d.
struct FastDiv {
    unsigned long long m;
    __device__ void init(unsigned d) { m = ((1ull << 32) + d - 1) / d; }
    __device__ unsigned operator()(unsigned t) const { return unsigned((t * m) >> 32); }
};
// (a, b, c) with t = a + n1 (b + n2 c), 0 <= a < n1, 0 <= b < n2.
struct Decode3 {
    int n1, n2;
    FastDiv f1, f12;
    __device__ void init(int a, int b) {
        n1 = a;
        n2 = b;
        f1.init(unsigned(a));
        f12.init(unsigned(a * b));
    }
    __device__ void operator()(int t, int& a, int& b, int& c) const {
        const int q = int(f1(unsigned(t)));
        c = int(f12(unsigned(t)));
        a = t - q * n1;
        b = q - c * n2;
    }
};

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

// Grid-wide sum of NV doubles through one ring entry ([RES_REPL][nb][NV] with
// copies RES_RSTRIDE apart); `done`: the entry used one iteration earlier,
// whose slots of this block are emptied.  Every wave reduces its lanes, wave 0
// adds the waves, publishes the block's partial (lanes 0..RES_REPL-1: one copy
// each), polls the nb <= 256 slots of its copy (lane l: slots l + 64 u, u < 4,
// all loads of a round in flight together) and reduces them in that fixed
// order, so `tot` holds the same total in wave 0 of every block.  RELEASE: the
// block's earlier device-scope stores complete before the publication.
// Returns false in wave 0 if a poll timed out (true in the other waves, which
// return after the first barrier).  The caller reads `tot` in thread 0 and
// then synchronises the block.
template <int NV, bool RELEASE>
__device__ bool grid_sum(const double (&v)[NV], double* entry, double* done, double* sh, double (&tot)[NV]) {
    if constexpr (RELEASE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0);
    }
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
        if (lane < RES_REPL) st_dev(&entry[lane * RES_RSTRIDE + blockIdx.x * NV + i], bsum);
    }
    const double* copy = entry + (blockIdx.x % RES_REPL) * RES_RSTRIDE;
    constexpr int U = RES_MAX_BLOCKS / 64;
    double val[U][NV];
    bool pend[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        pend[u] = lane + 64 * u < nb;
#pragma unroll
        for (int i = 0; i < NV; ++i) val[u][i] = 0.0;
    }
    int good = 1;
    for (long long spin = 0;; ++spin) {
        bool any = false;
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (pend[u]) {
#pragma unroll
                for (int i = 0; i < NV; ++i) val[u][i] = ld_dev(&copy[(lane + 64 * u) * NV + i]);
            }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (pend[u]) {
                bool all = true;
#pragma unroll
                for (int i = 0; i < NV; ++i) all = all && __builtin_bit_cast(unsigned long long, val[u][i]) != RING_EMPTY;
                pend[u] = !all;
                any = any || !all;
            }
        if (!any) break;
        if (spin >= SPIN_LIMIT) {
            good = 0;
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int u = 0; u < U; ++u) acc += val[u][i];
        tot[i] = wave_total(acc);
    }
    if (lane < RES_REPL) {
#pragma unroll
        for (int i = 0; i < NV; ++i)
            st_dev(reinterpret_cast<unsigned long long*>(&done[lane * RES_RSTRIDE + blockIdx.x * NV + i]), RING_EMPTY);
    }
    return __all(good) != 0;
}

// TA: arithmetic of the element products (T, or double for FP32 data; the
// outputs are rounded to T before the per-node sums).
template <typename T, int NT, bool REFINED, typename TA = T>
__global__ void __launch_bounds__(NT) k_resident(ResArgs<T> A) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ PcgState S;
    __shared__ double sRed[16 * 2];
    __shared__ int sFail;
    constexpr bool OFFS = sizeof(TA) == 4;  // subtract the corner-0 value first (FP32 rounding)
    const GridDev& g = A.g;
    const ResGeom& G = A.G;
    const TA* kh;
    if constexpr (sizeof(TA) == sizeof(T)) kh = A.kh;
    else kh = A.kh64;
    typedef typename Empty<T>::U EU;
    T* sP = reinterpret_cast<T*>(smem + G.o_p);
    T* sZ = reinterpret_cast<T*>(smem + G.o_z);
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
    Decode3 dO, dH, dE;                 // owned nodes, halo nodes, elements (j fastest)
    dO.init(bj, bi);
    dH.init(hj, hi);
    dE.init(cj, ci);
    auto gdof = [&](int i, int j, int k) { return 3LL * (((long long)k * g.NI + i) * g.NJ + j); };

    // ---- load: owned r, x, dinv; p_old on the halo brick; element scales
    for (int t = threadIdx.x; t < no; t += NT) {
        int lj, li, lk;
        dO(t, lj, li, lk);
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
            int hj_, hi_, hk_;
            dH(t, hj_, hi_, hk_);
            const int j = j0 - 1 + hj_, i = i0 - 1 + hi_, k = k0 - 1 + hk_;
            const bool in = unsigned(j) < unsigned(g.NJ) && unsigned(i) < unsigned(g.NI) && unsigned(k) < unsigned(g.NK);
            const int o = 3 * ((hk_ * G.HI + hi_) * G.HJ + hj_);
            const long long gd = in ? gdof(i, j, k) : 0;
#pragma unroll
            for (int c = 0; c < 3; ++c) sP[o + c] = (in && !first) ? pg[gd + c] : T(0);
        }
    }
    for (int t = threadIdx.x; t < ne; t += NT) {
        int lj, li, lk;
        dE(t, lj, li, lk);
        sS[t] = A.s_e[((long long)(ek0 + lk) * g.nx + (ei0 + li)) * g.ny + (ej0 + lj)];
    }
    __syncthreads();

    const int nb = gridDim.x;
    double* ring1 = A.ring;
    double* ring2 = A.ring + 3 * RES_ENTRY;
    long long kexec = S.k;
    bool ok = true;
    for (int it = 0;; ++it) {
        const T beta = T(S.beta);
        const bool first = S.first != 0;
        kexec = S.k;
        // ---- p on the halo brick, q = 0
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
                int hj_, hi_, hk_;
            dH(t, hj_, hi_, hk_);
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
                    } else {                // collected during the previous iteration's grid sum
#pragma unroll
                        for (int c = 0; c < 3; ++c) zv[u][c] = sZ[ho[u] + c];
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
            int lj, li, lk;
            dE(tt, lj, li, lk);
            const int ie = ei0 + li, je = ej0 + lj, ke = ek0 + lk;
            const int hb = ((ke - k0 + 1) * G.HI + (ie - i0 + 1)) * G.HJ + (je - j0 + 1);
            TA U[8][3];
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                const int h = 3 * (hb + (OZ(a) * G.HI + OX(a)) * G.HJ + OY(a));
#pragma unroll
                for (int c = 0; c < 3; ++c) U[a][c] = TA(sP[h + c]);
            }
            if constexpr (OFFS) {
#pragma unroll
                for (int a = 1; a < 8; ++a)
#pragma unroll
                    for (int c = 0; c < 3; ++c) U[a][c] -= U[0][c];
#pragma unroll
                for (int c = 0; c < 3; ++c) U[0][c] = TA(0);
            }
            sector_product<TA>(kh, U, ev ? TA(sS[tt]) : TA(0));
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                const int ni = ie + OX(a) - i0, nj = je + OY(a) - j0, nk = ke + OZ(a) - k0;
                if (ev && unsigned(ni) < unsigned(bi) && unsigned(nj) < unsigned(bj) && unsigned(nk) < unsigned(bk)) {
                    const int o = 3 * ((nk * G.BI + ni) * G.BJ + nj);
#pragma unroll
                    for (int c = 0; c < 3; ++c) sQ[o + c] += T(U[a][c]);
                }
                __syncthreads();
            }
        }

        // ---- p.q -> alpha
        const int e = it % 3, ed = (it + 2) % 3;
        double pq[1] = {0.0};
        for (int t = threadIdx.x; t < no; t += NT) {
            int lj, li, lk;
        dO(t, lj, li, lk);
            const int o = 3 * ((lk * G.BI + li) * G.BJ + lj);
            const int h = 3 * (((lk + 1) * G.HI + li + 1) * G.HJ + lj + 1);
#pragma unroll
            for (int c = 0; c < 3; ++c) pq[0] += double(sP[h + c]) * double(sQ[o + c]);
        }
        {
            double tot[1];
            const bool good = grid_sum<1, true>(pq, ring1 + e * RES_ENTRY, ring1 + ed * RES_ENTRY, sRed, tot);
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

        // ---- x += alpha p, r = mask(r - alpha q), z = dinv r; surface z published
        // in entry it % 3, this block's words of entry (it - 1) % 3 emptied (every
        // block collected them before its p.q publication of this iteration)
        T* zw = A.zr[e];
        EU* zc = reinterpret_cast<EU*>(A.zr[ed]);
        double rz[2] = {0.0, 0.0};
        for (int t = threadIdx.x; t < no; t += NT) {
            int lj, li, lk;
        dO(t, lj, li, lk);
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
                if (surf) {
                    st_dev(&zw[gd + c], z);
                    if (it > 0) st_dev(&zc[gd + c], Empty<T>::bits);
                }
            }
        }
        {
            double tot[2];
            // no release: the z words carry their own EMPTY test
            const bool good = grid_sum<2, false>(rz, ring2 + e * RES_ENTRY, ring2 + ed * RES_ENTRY, sRed, tot);
            // waves 1.. collect the z of the shell nodes into the z layout while
            // wave 0 runs the grid sum (a collection share of wave 0 would start
            // only after its polls and lengthen the phase)
            static_assert(NT >= 128, "the z collection needs waves besides wave 0");
            constexpr int NC = NT - 64;
            const int ct = int(threadIdx.x) - 64;
            const EU* zr = reinterpret_cast<const EU*>(A.zr[e]);
            int zgood = 1;
            for (int base = 0; ct >= 0 && base < nh; base += PF * NC) {
                EU zb[PF][3];
                int ho[PF];
                long long gd[PF];
                bool pend[PF];
#pragma unroll
                for (int u = 0; u < PF; ++u) {
                    const int t = base + u * NC + ct;
                    int hj_, hi_, hk_;
                    dH(t, hj_, hi_, hk_);
                    const int j = j0 - 1 + hj_, i = i0 - 1 + hi_, k = k0 - 1 + hk_;
                    const bool own = hj_ >= 1 && hj_ <= bj && hi_ >= 1 && hi_ <= bi && hk_ >= 1 && hk_ <= bk;
                    pend[u] = t < nh && !own && unsigned(j) < unsigned(g.NJ) && unsigned(i) < unsigned(g.NI) &&
                              unsigned(k) < unsigned(g.NK);
                    ho[u] = 3 * ((hk_ * G.HI + hi_) * G.HJ + hj_);
                    gd[u] = pend[u] ? gdof(i, j, k) : 0;
                }
                for (long long spin = 0;; ++spin) {
#pragma unroll
                    for (int u = 0; u < PF; ++u)
                        if (pend[u]) {
#pragma unroll
                            for (int c = 0; c < 3; ++c) zb[u][c] = ld_dev(&zr[gd[u] + c]);
                        }
                    bool any = false;
#pragma unroll
                    for (int u = 0; u < PF; ++u)
                        if (pend[u]) {
                            if (zb[u][0] != Empty<T>::bits && zb[u][1] != Empty<T>::bits &&
                                zb[u][2] != Empty<T>::bits) {
                                pend[u] = false;
#pragma unroll
                                for (int c = 0; c < 3; ++c) sZ[ho[u] + c] = __builtin_bit_cast(T, zb[u][c]);
                            } else {
                                any = true;
                            }
                        }
                    if (!any) break;
                    if (spin >= SPIN_LIMIT) {
                        zgood = 0;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            if (!zgood) sFail = 1;
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
        int lj, li, lk;
        dO(t, lj, li, lk);
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

// Sum rings -> RING_EMPTY, z rings -> Empty<T>, error word -> 0 (before each launch).
template <typename T>
__global__ void __launch_bounds__(256) k_ring_reset(unsigned long long* ring, int n, unsigned* err, ResArgs<T> A) {
    typedef typename Empty<T>::U EU;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) ring[i] = RING_EMPTY;
    for (int e = 0; e < 3; ++e) {
        EU* z = reinterpret_cast<EU*>(A.zr[e]);
        for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < A.g.ndof; i += stride)
            z[i] = Empty<T>::bits;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *err = 0u;
}

}  // namespace h8
