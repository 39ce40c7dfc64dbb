// h8_sector.h -- matrix-free K*u through the reflection sectors of the brick
// element, marching through the grid in z.
//
// Sector form of the element product.  Every brick element is symmetric under
// the reflections x -> -x, y -> -y, z -> -z about its centre; a reflection g
// maps corner a to a^g and flips the sign of displacement component c when g
// reflects axis c.  K0 commutes with these maps, so the Walsh-Hadamard
// transform over the 8 corners,
//     U^[chi][c] = sum_a (-1)^(chi . a) U[a][c]      (a, chi as 3-bit masks),
// block-diagonalises it: output (chi, c) couples only to inputs (chi', c')
// with chi ^ e_c == chi' ^ e_c' =: psi (e_c = the bit of axis c), i.e. into 8
// sectors psi of 3 components each:
//     Y^[psi ^ e_c][c] = sum_c' Kh[psi][c][c'] U^[psi ^ e_c'][c'],
//     Kh[psi][c][c'] = (1/8) sum_x (-1)^((psi ^ e_c') . x) K0[c][3x + c']
// (the 1/8 of the inverse transform folded in; only K0's first three rows
// enter).  The three translation sectors psi = e_c have a zero row and column
// c (K0 annihilates rigid translations).  One element product is then a
// forward transform (72 add/sub), 57 multiply-adds and an inverse transform
// (72 add/sub) instead of 576 multiply-adds.
//
// Marching.  A block owns a column of TJ x TI = (EJ-1) x (EI-1) nodes over the
// node layers [kb, ke) and walks the element slabs s = kb-1 .. ke-1 (slab s
// lies between node layers s and s+1) upwards with one thread per element of
// the EJ x EI slab (the slab's elements cover the owned nodes plus a one-node
// halo ring).  Node layers are staged into a 3-slot LDS ring (the direction
// p = dinv*r + beta*p_old is formed here in the direction modes, and written out for the
// owned nodes); each element writes the outputs of its four bottom corners
// (for node layer s) and its four top corners (for layer s+1) into per-node
// slots, and every owned node sums its slots in a fixed order (deterministic)
// - the top-corner sum is carried to the next step in a register.  The next
// layer's global loads are issued before the element work of a step and
// stored into the ring after it.  Two barriers per step.  Corner outputs are
// summed in pairs along j across lanes (DPP) before they reach LDS, which
// halves the slot LDS: six FP32 blocks or three FP64 blocks fit per CU
// instead of four and two.
#pragma once
#include "h8_device.h"
#include "h8_node.h"

namespace h8 {

// Corner index of the 3-bit mask (bit 0 = x offset, bit 1 = y, bit 2 = z).
__host__ __device__ constexpr int corner_of_mask(int m) { return corner_of(m & 1, (m >> 1) & 1, (m >> 2) & 1); }

// In-place unnormalised Walsh-Hadamard transform over the corner masks.
template <typename T>
__device__ __forceinline__ void wht8(T (&v)[8]) {
#pragma unroll
    for (int bit = 1; bit < 8; bit <<= 1)
#pragma unroll
        for (int m = 0; m < 8; ++m)
            if (!(m & bit)) {
                const T u = v[m], w = v[m | bit];
                v[m] = u + w;
                v[m | bit] = u - w;
            }
}

// Element product of one element: Y[a][c] (corner a in the reference order)
// from U[a][c] through the sectors.  kh: 72 values Kh[psi][c][c'] (uniform).
template <typename T>
__device__ __forceinline__ void sector_product(const T* kh, T (&U)[8][3], T scale) {
    T V[3][8];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int m = 0; m < 8; ++m) V[c][m] = U[corner_of_mask(m)][c];
        wht8(V[c]);
    }
    // scale the transformed input once (the product is linear)
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int m = 1; m < 8; ++m) V[c][m] *= scale;
    T W[3][8];
#pragma unroll
    for (int c = 0; c < 3; ++c) W[c][0] = T(0);    // the translation outputs of sectors e_c
    T dep = V[0][1];
#pragma unroll
    for (int psi = 0; psi < 8; ++psi) {
        // this sector's 9 table values are loaded after the previous sector's
        // products (data dependence through `dep`), so at most a few sectors'
        // values occupy scalar registers at a time
        const T* kp = kh;
        asm volatile("" : "+s"(kp) : "v"(dep));
        const auto K = as_const(kp);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (psi == (1 << c)) continue;          // zero row of a translation sector
            T acc = T(0);
            bool firstterm = true;
#pragma unroll
            for (int c2 = 0; c2 < 3; ++c2) {
                if (psi == (1 << c2)) continue;     // zero column of a translation sector
                const T kv = K[psi * 9 + c * 3 + c2];
                const T x = V[c2][psi ^ (1 << c2)];
                if (firstterm) { acc = kv * x; firstterm = false; }
                else acc += kv * x;
            }
            W[c][psi ^ (1 << c)] = acc;
            dep = acc;
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        wht8(W[c]);
#pragma unroll
        for (int m = 0; m < 8; ++m) U[corner_of_mask(m)][c] = W[c][m];
    }
}

// ---- element bilinear forms out_e = w_e u_e^T K0 v_e -------------------------
// Through the sectors: with U^ = WHT(u_e) and V^ = WHT(v_e) per component,
//     u_e^T K0 v_e = sum_psi sum_{c,c'} U^[c][psi ^ e_c] Kh[psi][c][c'] V^[c'][psi ^ e_c']
// (Kh carries the 1/8 of the inverse transform): two forward transforms
// (144 add/sub), 57 + 21 multiply-adds, instead of 576.
// A block of BJ x BI threads owns a column of BJ x BI elements and marches
// up through `kd` element slabs; the (BJ+1) x (BI+1) node layers of u and v
// are staged once each into a two-slot LDS ring (component planes), the next
// layer's loads in flight while the current slab is computed.  SAME: u == v
// (the compliance form), one vector staged.
constexpr int BIL_BJ = 16, BIL_BI = 16;

template <bool SAME>
__global__ void __launch_bounds__(BIL_BJ * BIL_BI)
k_elem_bilinear(GridDev g, const double* __restrict__ K0, const double* __restrict__ u,
                const double* __restrict__ v, const double* __restrict__ w, double* __restrict__ out, int kd) {
    constexpr int NT = BIL_BJ * BIL_BI, LJ = BIL_BJ + 1, LI = BIL_BI + 1, LN = LJ * LI, LW = 3 * LN;
    constexpr int PER = (LW + NT - 1) / NT;     // staged words per thread and vector
    constexpr int NV = SAME ? 1 : 2;
    __shared__ double ring[2][NV][LW];          // [slot][vector][c * LN + node]
    const int tj = (g.ny + BIL_BJ - 1) / BIL_BJ, ti = (g.nx + BIL_BI - 1) / BIL_BI;
    const int bj = blockIdx.x % tj, bi = (blockIdx.x / tj) % ti, bk = blockIdx.x / (tj * ti);
    const int j0 = bj * BIL_BJ, i0 = bi * BIL_BI, k0 = bk * kd;
    if (k0 >= g.nz) return;
    const int k1 = min(k0 + kd, g.nz);          // element slabs [k0, k1), node layers k0 .. k1
    const int t = threadIdx.x;
    // the words this thread stages: global order [hi][hj][c] of the layer's rows
    int goff[PER], loff[PER];
    bool ok[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int wd = t + q * NT;
        const int c = wd % 3, nd = wd / 3, hj = nd % LJ, hi = nd / LJ;
        ok[q] = wd < LW && j0 + hj <= g.ny && i0 + hi <= g.nx;
        goff[q] = ok[q] ? 3 * ((i0 + hi) * g.NJ + j0 + hj) + c : 0;
        loff[q] = c * LN + nd;
    }
    const long long layer = 3LL * g.NI * g.NJ;
    double reg[NV][PER];
    auto load = [&](int kl) {
        const double* __restrict__ ul = u + kl * layer;
        const double* __restrict__ vl = v + kl * layer;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            reg[0][q] = ok[q] ? ul[goff[q]] : 0.0;
            if constexpr (!SAME) reg[NV - 1][q] = ok[q] ? vl[goff[q]] : 0.0;
        }
    };
    auto store = [&](int slot) {
#pragma unroll
        for (int q = 0; q < PER; ++q)
            if (t + q * NT < LW) {
#pragma unroll
                for (int vv = 0; vv < NV; ++vv) ring[slot][vv][loff[q]] = reg[vv][q];
            }
    };
    const int ej = t % BIL_BJ, ei = t / BIL_BJ;
    const bool mine = j0 + ej < g.ny && i0 + ei < g.nx;
    const auto kh = as_const(K0 + 576);
    load(k0);
    store(k0 & 1);
    load(k0 + 1);
    for (int k = k0; k < k1; ++k) {
        __syncthreads();                        // the slot of layer k - 1 is free
        store((k + 1) & 1);
        __syncthreads();
        if (k + 2 <= k1) load(k + 2);           // in flight during this slab
        if (mine) {
            // v's transform and the sector products first (W[c][psi ^ e_c]), then
            // u's transform and the dot: at most 45 values live at a time
            double W[3][8];
            {
                double V[3][8];
#pragma unroll
                for (int m = 0; m < 8; ++m) {
                    const int slot = (k + (m >> 2)) & 1;
                    const int nd = (ei + (m & 1)) * LJ + ej + ((m >> 1) & 1);
#pragma unroll
                    for (int c = 0; c < 3; ++c) V[c][m] = ring[slot][NV - 1][c * LN + nd];
                }
#pragma unroll
                for (int c = 0; c < 3; ++c) wht8(V[c]);
#pragma unroll
                for (int c = 0; c < 3; ++c) W[c][0] = 0.0;
#pragma unroll
                for (int psi = 0; psi < 8; ++psi)
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        if (psi == (1 << c)) continue;      // zero row of a translation sector
                        double y = 0.0;
#pragma unroll
                        for (int c2 = 0; c2 < 3; ++c2) {
                            if (psi == (1 << c2)) continue; // zero column
                            y += kh[psi * 9 + c * 3 + c2] * V[c2][psi ^ (1 << c2)];
                        }
                        W[c][psi ^ (1 << c)] = y;
                    }
                if constexpr (SAME) {
                    double acc = 0.0;
#pragma unroll
                    for (int c = 0; c < 3; ++c)
#pragma unroll
                        for (int m = 1; m < 8; ++m) acc += V[c][m] * W[c][m];
                    W[0][0] = acc;
                }
            }
            double acc;
            if constexpr (SAME) {
                acc = W[0][0];
            } else {
                double U[3][8];
#pragma unroll
                for (int m = 0; m < 8; ++m) {
                    const int slot = (k + (m >> 2)) & 1;
                    const int nd = (ei + (m & 1)) * LJ + ej + ((m >> 1) & 1);
#pragma unroll
                    for (int c = 0; c < 3; ++c) U[c][m] = ring[slot][0][c * LN + nd];
                }
                acc = 0.0;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    wht8(U[c]);
#pragma unroll
                    for (int m = 1; m < 8; ++m) acc += U[c][m] * W[c][m];
                }
            }
            const long long e = ((long long)k * g.nx + i0 + ei) * g.ny + j0 + ej;
            out[e] = w[e] * acc;
        }
    }
}

// Layer staging of the marching kernel: a node layer of the column
// ((EJ+1) x (EI+1) nodes, 3 words each, node-interleaved [hi][hj][c] = the
// global order of those words, row by row) is LW words; thread t owns words
// t + NT*u.  The word's row / in-row offset and its range tests do not change
// between layers, so they are computed once per thread (SLayer::init); per
// layer only the k term is added.
template <typename T, int EJ, int EI>
struct SLayer {
    static constexpr int NT = EJ * EI;
    static constexpr int RW = 3 * (EJ + 1);            // words per row
    static constexpr int LW = RW * (EI + 1);           // words per layer
    static constexpr int PASS = (LW + NT - 1) / NT;
    // byte offsets without the k term: ofs = OOB outside the (i, j) domain,
    // sofs = OOB unless the word belongs to an owned node.  Adding a k term
    // (< 2^31 bytes) to OOB = 2^31 keeps it out of range, so no per-word flags
    // are needed (they would occupy scalar registers as lane masks).
    int ofs[PASS], sofs[PASS];
    T v0[PASS], v1[PASS], v2[PASS], xv[PASS];

    __device__ __forceinline__ void init(const GridDev& g, int j0, int i0) {
#pragma unroll
        for (int u = 0; u < PASS; ++u) {
            const int w = threadIdx.x + NT * u;
            const int row = w / RW, r = w - RW * (w / RW);
            const int hj = r / 3;
            const int i = i0 - 1 + row, j = j0 - 1 + hj;
            const bool in = w < LW && i >= 0 && i < g.NI && j >= 0 && j < g.NJ;
            const int o = int(sizeof(T)) * (3 * (i * g.NJ + j0 - 1) + r);
            ofs[u] = in ? o : OOB;
            sofs[u] = (in && row >= 1 && row <= EI - 1 && hj >= 1 && hj <= EJ - 1) ? o : OOB;
        }
    }
    // global loads of layer k.  Branch-free: a layer outside the grid, and the
    // previous direction in the first iteration, get the out-of-range bit ORed
    // into their offsets and load zeros.
    // x_own (MODE_DIRX): the layer is owned and the previous x update is
    // pending (uniform): the owned words of x are loaded for the update
    // applied in store().
    template <int MODE>
    __device__ __forceinline__ void load(const NodeArgs<T>& A, const T* p_old, int first, int k, bool x_own) {
        const GridDev& g = A.g;
        const unsigned vbytes = unsigned(g.ndof) * sizeof(T);
        const bool kin = k >= 0 && k < g.NK;
        const unsigned kbad = kin ? 0u : unsigned(OOB);
        const unsigned kofs = kin ? unsigned(sizeof(T)) * 3u * unsigned(k) * unsigned(g.NI) * unsigned(g.NJ) : 0u;
        const unsigned pbad = first ? unsigned(OOB) : 0u;
        const unsigned xbad = x_own ? 0u : unsigned(OOB);
#pragma unroll
        for (int u = 0; u < PASS; ++u) {
            const unsigned off = (unsigned(ofs[u]) + kofs) | kbad;
            if constexpr (MODE == MODE_CHEB0) {
                v0[u] = Buf<T>::ld(buf_rsrc(A.cb_b, vbytes), int(off));   // b
                v1[u] = Buf<T>::ld(buf_rsrc(A.dinv, vbytes), int(off));
            } else if constexpr (MODE == MODE_DIRZ) {
                v0[u] = Buf<T>::ld(buf_rsrc(A.r, vbytes), int(off));      // z
                v1[u] = T(1);
                v2[u] = Buf<T>::ld(buf_rsrc(p_old, vbytes), int(off | pbad));
            } else if constexpr (is_dir(MODE)) {
                v0[u] = Buf<T>::ld(buf_rsrc(A.r, vbytes), int(off));
                v1[u] = Buf<T>::ld(buf_rsrc(A.dinv, vbytes), int(off));
                v2[u] = Buf<T>::ld(buf_rsrc(p_old, vbytes), int(off | pbad));
                if constexpr (MODE == MODE_DIRX)
                    xv[u] = Buf<T>::ld(buf_rsrc(A.x, vbytes), int((unsigned(sofs[u]) + kofs) | kbad | xbad));
            } else {
                v0[u] = Buf<T>::ld(buf_rsrc(A.u, vbytes), int(off));
                if constexpr (MODE == MODE_RESID) {
                    // + the unfolded FP32 iterate (FP32 word = half the FP64 byte offset;
                    // an empty buffer when there is none, so the load returns 0)
                    const unsigned xbytes = A.x32 ? unsigned(g.ndof) * sizeof(float) : 0u;
                    const int o32 = (off & 0x80000000u) ? OOB : int(off >> 1);
                    v0[u] += T(Buf<float>::ld(buf_rsrc(A.x32, xbytes), o32));
                }
            }
        }
    }
    // form the layer's values, write the owned direction entries (k_owned is
    // wave-uniform), store to LDS
    template <int MODE>
    __device__ __forceinline__ void store(const NodeArgs<T>& A, T beta, T* p_new, T alpha_prev, bool x_own, int k,
                                          bool k_owned, T* slot) {
        const unsigned vbytes = unsigned(A.g.ndof) * sizeof(T);
        const unsigned kbad = k_owned ? 0u : unsigned(OOB);
        const unsigned kofs = k_owned ? unsigned(sizeof(T)) * 3u * unsigned(k) * unsigned(A.g.NI) * unsigned(A.g.NJ) : 0u;
        const unsigned xbad = x_own ? 0u : unsigned(OOB);
#pragma unroll
        for (int u = 0; u < PASS; ++u) {
            const int w = threadIdx.x + NT * u;
            T val;
            if constexpr (is_dir(MODE)) {
                val = v1[u] * v0[u] + beta * v2[u];
                const unsigned so = (unsigned(sofs[u]) + kofs) | kbad;
                Buf<T>::st(val, buf_rsrc(p_new, vbytes), int(so));
                // the previous iteration's x += alpha p_old, for the owned words
                if constexpr (MODE == MODE_DIRX)
                    Buf<T>::st(xv[u] + alpha_prev * v2[u], buf_rsrc(A.x, vbytes), int(so | xbad));
            } else if constexpr (MODE == MODE_CHEB0) {
                val = beta * v1[u] * v0[u];        // x0 = c0 dinv b (beta carries c0)
            } else {
                val = v0[u];
            }
            if (u + 1 < PASS || w < LW) slot[w] = val;
        }
    }
};

// Waves per SIMD the register allocator aims for.  FP64: the 16 x 16 column's
// LDS ring (45.9 KB per 256-thread block) fits three blocks in a CU's 160 KB,
// so three waves is the most the sector kernel can have whatever its
// registers (118 VGPRs would allow four); asking for four changed nothing in
// the sector kernel's time and pushed the persistent batched kernel, which
// shares this target, to 201 VGPRs and two waves.
template <typename T> struct SectorWaves;
template <> struct SectorWaves<double> { static constexpr int value = 3; };
template <> struct SectorWaves<float> { static constexpr int value = 6; };

// LDS of one march: the 3-slot layer ring and the per-node corner slots.
template <typename T, int EJ, int EI>
struct SectorSmem {
    // pair slots of the bottom (lo) and top (hi) node layer, component-major: value
    // k = 3 x + c (x = the corner column OX, c the component) of thread t at
    // k * EJ * EI + t, so lane-adjacent threads hit adjacent LDS banks
    alignas(16) T ring[3 * SLayer<T, EJ, EI>::LW];
    alignas(16) T lo[EJ * EI * 6];
    alignas(16) T hi[EJ * EI * 6];
};

// The value of the lane one to the left within a 16-lane DPP row (row_shr:1);
// the row's first lane gets 0.
__device__ __forceinline__ float lane_left(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x111, 0xF, 0xF,
                                                                 false));
}
__device__ __forceinline__ double lane_left(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, int(unsigned(u)), 0x111, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, int(unsigned(u >> 32)), 0x111, 0xF, 0xF, false);
    return __builtin_bit_cast(double, (unsigned long long)(unsigned)lo | ((unsigned long long)(unsigned)hi << 32));
}

// Column tiles of the sector march for a grid: (EJ-1) x (EI-1) nodes, chunks of kd layers.
__host__ __device__ inline void sector_tiles(GridDev& g, int EJ, int EI, int kd) {
    g.tj = (g.NJ + EJ - 2) / (EJ - 1);
    g.ti = (g.NI + EI - 2) / (EI - 1);
    g.tk = (g.NK + kd - 1) / kd;
    g.ntiles = g.tj * g.ti * g.tk;
}

// The march of one block over `tile` (column tile + layer chunk) of grid A.g:
// stages layers, forms p (direction modes; written for owned nodes), evaluates the
// element products through the sectors and writes each owned node's outputs
// (node_output), accumulating the fused partials in `part`.
// TA: arithmetic of the element products (T, or double for FP32 data: the
// products' rounding, not the FP32 storage, is what slows the FP32 iterations'
// convergence); the per-node sums of the outputs are in T.
template <typename T, int EJ, int EI, int MODE, typename TA = T>
__device__ __forceinline__ void sector_march(const NodeArgs<T>& A, const TA* kh, int kd, int tile, T beta, int first,
                                             T* p_new, const T* p_old, T alpha_prev, int pending,
                                             SectorSmem<T, EJ, EI>& sm, double (&part)[2]) {
    constexpr int TJ = EJ - 1, TI = EI - 1;
    typedef SLayer<T, EJ, EI> L;
    constexpr bool OFFS = sizeof(TA) == 4;  // subtract the corner-0 value first (FP32 rounding)
    const GridDev& g = A.g;
    const int bj = tile % g.tj, bi = (tile / g.tj) % g.ti, bk = tile / (g.tj * g.ti);
    const int j0 = bj * TJ, i0 = bi * TI;
    const int kb = bk * kd, ke = min(g.NK, kb + kd);

    const int ej = threadIdx.x % EJ, ei = threadIdx.x / EJ;
    const int je = j0 - 1 + ej, ie = i0 - 1 + ei;          // element column of this thread
    const bool ecol = je >= 0 && je < g.ny && ie >= 0 && ie < g.nx;
    const unsigned sbytes = unsigned(g.nx) * unsigned(g.ny) * unsigned(g.nz) * sizeof(T);
    // owned node of this thread in the node phase: halo position (ej, ei), both >= 1
    const int nj = j0 - 1 + ej, ni = i0 - 1 + ei;
    const bool nown = ej >= 1 && ei >= 1 && nj < g.NJ && ni < g.NI;
    const int lbase = 3 * (ei * (EJ + 1) + ej);             // LDS word of corner 0 within a layer
    T* sRing = sm.ring;
    T* sLo = sm.lo;
    T* sHi = sm.hi;

    L lay;
    lay.init(g, j0, i0);
    // prologue: layers kb-1 and kb into ring slots 0 and 1
    lay.template load<MODE>(A, p_old, first, kb - 1, false);
    lay.template store<MODE>(A, beta, p_new, alpha_prev, false, kb - 1, false, sRing);
    lay.template load<MODE>(A, p_old, first, kb, pending && kb < ke);
    lay.template store<MODE>(A, beta, p_new, alpha_prev, pending && kb < ke, kb, kb < ke, sRing + L::LW);
    T se_next = Buf<T>::ld(buf_rsrc(A.s_e, sbytes),
                           (ecol && kb - 1 >= 0) ? int(sizeof(T)) * (((kb - 1) * g.nx + ie) * g.ny + je) : OOB);
    __syncthreads();

    T carry[3] = {T(0), T(0), T(0)};
    // the node operands of the fused outputs, one layer ahead: step s loads
    // layer s + 1's into ex_next, and ex holds layer s's (loaded in step s - 1)
    NodeEx<T> ex, ex_next;
    const int gn0 = 3 * (ni * g.NJ + nj);                   // DOF of the thread's node in layer 0
    const int lstride = 3 * g.NI * g.NJ;
    for (int s = kb - 1; s < ke; ++s) {
        const int q = s - (kb - 1);                          // step index
        const T* lo = sRing + (q % 3) * L::LW;               // layer s
        const T* up = sRing + ((q + 1) % 3) * L::LW;         // layer s + 1
        // prefetch layer s + 2 and the next slab's element scale
        const bool more = s + 2 <= ke;
        if (more) lay.template load<MODE>(A, p_old, first, s + 2, pending && s + 2 < ke);
        if constexpr (has_node_ex<MODE>())
            node_ex_load<T, MODE>(A, gn0 + (s + 1) * lstride, nown && s + 1 < ke, ex_next);
        const T se = se_next;
        se_next = Buf<T>::ld(buf_rsrc(A.s_e, sbytes),
                             (ecol && s + 1 < g.nz && s + 1 < ke) ? int(sizeof(T)) * (((s + 1) * g.nx + ie) * g.ny + je)
                                                                  : OOB);
        // ---- element phase: slab s, element (je, ie, s) (scale 0 outside the mesh)
        TA U[8][3];
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const T* src = OZ(a) ? up : lo;
#pragma unroll
            for (int c = 0; c < 3; ++c) U[a][c] = TA(src[lbase + 3 * (OX(a) * (EJ + 1) + OY(a)) + c]);
        }
        if constexpr (OFFS) {
#pragma unroll
            for (int a = 1; a < 8; ++a)
#pragma unroll
                for (int c = 0; c < 3; ++c) U[a][c] -= U[0][c];
#pragma unroll
            for (int c = 0; c < 3; ++c) U[0][c] = TA(0);
        }
        sector_product<TA>(kh, U, TA(se));
        // Corner a of this element is node (ej + OY(a), ei + OX(a)) of the halo grid.
        // Pairs along j are summed in registers: node (ej, ei) of a layer takes this
        // element's corner (OX, OY) = (0, 0) plus the (0, 1) corner of element ej - 1,
        // the lane to the left (threads are j-fastest; ej = 0 takes nothing), and
        // node (ej, ei + 1) the (1, 0) and (1, 1) corners likewise.  A node then
        // adds its own pair and the OX = 1 pair of thread (ej, ei - 1): the same
        // four corners in a fixed order (deterministic), from half the slots.
        // Corners on nodes beyond the tile (ej + 1 = EJ or ei + 1 = EI) are not used.
        constexpr int NT = EJ * EI;
#pragma unroll
        for (int z = 0; z < 2; ++z) {
            T* dst = z ? sHi : sLo;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                T l3 = lane_left(T(U[4 * z + 3][c])), l2 = lane_left(T(U[4 * z + 2][c]));
                if (ej == 0) {
                    l3 = T(0);
                    l2 = T(0);
                }
                dst[c * NT + threadIdx.x] = T(U[4 * z + 0][c]) + l3;
                dst[(3 + c) * NT + threadIdx.x] = T(U[4 * z + 1][c]) + l2;
            }
        }
        __syncthreads();
        // ---- node phase: layer s outputs, layer s + 1 carry
        {
            const int tw = threadIdx.x >= EJ ? int(threadIdx.x) - EJ : int(threadIdx.x);   // thread (ej, ei - 1)
            const T wmask = threadIdx.x >= EJ ? T(1) : T(0);
            T qv[3], ctr[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T sl = sLo[c * NT + threadIdx.x] + wmask * sLo[(3 + c) * NT + tw];
                const T sh = sHi[c * NT + threadIdx.x] + wmask * sHi[(3 + c) * NT + tw];
                qv[c] = carry[c] + sl;
                ctr[c] = lo[lbase + c];
                carry[c] = sh;
            }
            if constexpr (has_node_ex<MODE>()) {
                if (nown && s >= kb) node_output_ex<T, MODE>(A, gn0 + s * lstride, qv, ctr, ex, part);
                ex = ex_next;
            } else {
                if (nown && s >= kb) node_output<T, MODE>(A, 3 * ((s * g.NI + ni) * g.NJ + nj), qv, ctr, part);
            }
        }
        // ring refill with layer s + 2 (its slot is not read in this step)
        if (more)
            lay.template store<MODE>(A, beta, p_new, alpha_prev, pending && s + 2 < ke, s + 2, s + 2 < ke,
                                     sRing + ((q + 2) % 3) * L::LW);
        __syncthreads();
    }
}

// One block per (column tile, layer chunk); A.g.{tj, ti, tk} count the column
// tiles in j and i and the chunks of `kd` node layers.
template <typename T, int EJ, int EI, int MODE, typename TA = T>
__global__ void __launch_bounds__(EJ * EI) __attribute__((amdgpu_waves_per_eu(SectorWaves<TA>::value)))
k_sector(NodeArgs<T> A, const TA* kh, int kd) {
    static_assert((EJ * EI) % 64 == 0, "whole waves");
    __shared__ SectorSmem<T, EJ, EI> sm;
    __shared__ double sRed[16 * 4];
    T beta, alpha_prev;
    int first, pending;
    T* p_new;
    const T* p_old;
    if (!node_setup<T, MODE>(A, beta, first, p_new, p_old, alpha_prev, pending)) return;
    double part[2] = {0.0, 0.0};
    sector_march<T, EJ, EI, MODE, TA>(A, kh, kd, xcd_local_tile(blockIdx.x, gridDim.x), beta, first, p_new, p_old,
                                      alpha_prev, pending, sm, part);
    node_finish<T, MODE>(A, part, sRed);
}

}  // namespace h8
