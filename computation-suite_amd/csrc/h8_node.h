// h8_node.h -- node-centric matrix-free K*u for the structured H8 grid.
//
// One thread owns one node and computes its three rows of K*u by gathering
// over the (up to) eight elements that touch it:
//     q_n[c] = sum_a s_e(a) sum_b sum_d K0[3a+c][3b+d] * u[nbr(n,a,b)][d]
// where a is the corner the node occupies in element e(a) and b runs over that
// element's corners.  Each element's 24x24 product is therefore split across
// its 8 nodes with no redundant arithmetic (576 FMA per node) and no atomics:
// the result is deterministic.  A block owns a TJ x TI x TK brick of nodes,
// stages the brick plus a one-node halo (component-planar, so lane-adjacent
// reads are unit stride) and the (TJ+1)(TI+1)(TK+1) element scales in LDS,
// then each thread keeps its 27-node neighbourhood in registers.  K0 is read
// with uniform (scalar) loads; only its first three rows are used (the brick
// reflection structure gives every other block), so K0 must have that
// structure exactly (GridOperator imposes it).
//
// MODE_PLAIN : y = K u (every DOF).
// MODE_DIR   : PCG direction + product, fused: p = dinv*r (+ beta*p_old) is
//              formed while staging the halo (so no separate direction pass),
//              written for owned nodes, q = K p, and p.q is reduced; the last
//              block forms alpha = rz_old / p.q or flags p.q <= 0.
// MODE_RESID : FP64 true residual of the refined solve: r64 = mask(f - K u),
//              r32 = float(r64), reduces ||r64||^2 and r32.(dinv32*r32); the
//              last block decides between "converged" and "continue with the
//              replaced residual".
#pragma once
#include "h8_device.h"

namespace h8 {

enum { MODE_PLAIN = 0, MODE_DIR = 1, MODE_RESID = 2, MODE_DIRX = 3, MODE_DIRZ = 4, MODE_CHEB = 5, MODE_MGRES = 6,
       MODE_CHEB0 = 7 };
// MODE_DIRX = MODE_DIR that also applies the previous iteration's deferred
// x += alpha p_old to the owned nodes while staging p_old.
// MODE_DIRZ = MODE_DIR with a given preconditioned residual: p = z + beta p_old,
// z in NodeArgs::r (the multigrid-preconditioned loop; sector kernels only).
// MODE_CHEB = a Chebyshev smoothing step of the multigrid cycle on the grid:
// d = ca d + cb dinv (b - K x), x_out = x + d (x = NodeArgs::u, sector kernels
// only); with cheb_rz also the reduction r.z of x_out with b and the beta step.
// MODE_CHEB0 = the cycle's first two Chebyshev steps from x = 0 in one pass:
// x0 = d = c0 dinv b is formed while staging (NodeArgs::cb_b, dinv), then
// d = c1a d + c1b dinv (b - K x0), x_out = x0 + d.
// MODE_MGRES = the masked residual of the cycle, y = (dinv != 0) (b - K u),
// the input of the restriction (sector kernels only).
__host__ __device__ constexpr bool is_dir(int mode) { return mode == MODE_DIR || mode == MODE_DIRX || mode == MODE_DIRZ; }

template <typename T>
struct NodeArgs {
    GridDev g;
    const T* s_e;
    const T* K0;      // 24x24 row-major
    const T* u;       // PLAIN / RESID input
    T* y;             // PLAIN output
    const T* r;       // DIR
    const T* dinv;    // DIR (T); RESID mask (double)
    T* p0;            // DIR: direction buffers, iteration k writes p[k&1]
    T* p1;            //      and reads the previous direction from p[(k-1)&1]
    T* q;             // DIR
    T* x;             // DIR: iterate; x += alpha p_old is applied to owned nodes while staging
    const double* f;  // RESID
    float* r32;       // RESID
    const float* dinv32;  // RESID
    PcgState* st;
    double* partials;
    int only_on_replace;  // RESID: 1 = every block returns at once unless st asks for a replacement,
                          // 2 = unless the solve is running (the first residual of a captured solve)
    const float* x32;     // RESID (sector kernel, multigrid solve): the residual is formed at
                          // u + x32 (the FP32 iterate not yet folded into u); nullptr: at u
    const float* cheb;    // CHEB: the level's coefficients {c0, c1a, c1b, c2a, c2b, ...} (device)
    const T* cb_b;        // CHEB: right-hand side
    T* cb_d;              // CHEB: Chebyshev direction (updated in place)
    T* cb_out;            // CHEB: x_out (not u: other blocks still read u)
    int cheb_step;        // CHEB: step k of a pass (0: ca = 0, cb = c0; k: cheb[2k-1], cheb[2k])
    int cheb_rz;          // CHEB: reduce b.x_out and finish the iteration's beta
};

// Corner whose offsets are the XOR of corners a and b, and the sign of
// component c under the reflection that takes corner a to corner 0.
__host__ __device__ constexpr int corner_of(int ox, int oy, int oz) {
    return oz * 4 + (oy ? (ox ? 2 : 3) : (ox ? 1 : 0));
}
__host__ __device__ constexpr int corner_xor(int a, int b) {
    return corner_of(OX(a) ^ OX(b), OY(a) ^ OY(b), OZ(a) ^ OZ(b));
}
__host__ __device__ constexpr int refl_sign(int a, int c) {
    return ((c == 0 ? OX(a) : (c == 1 ? OY(a) : OZ(a))) != 0) ? -1 : 1;
}

// Last block of a direction kernel: alpha = rz_old / p.Ap, or stop on p.Ap <= 0
// (fe/iterative_solvers.py:449-454).
__device__ inline void dir_finish(PcgState* st, double pq) {
    st->pq = pq;
    if (!(pq > 0.0)) {
        st->bad_pq = pq;
        st->iters = st->k;
        st->status = ST_NOT_SPD;
        st->pending = 0;
    } else {
        st->alpha = st->rz_old / pq;
        st->pending = 1;       // x += alpha p is applied by the next direction kernel
    }
}

// End of the multigrid-preconditioned FP32 iteration, given rz = r.z with
// z the V-cycle's output: beta = 0 when the direction restarts (p = z), else
// rz / rz_old, and a replacement that makes beta >= 1/delta^2 restarts it.
__device__ inline void mg_rz_finish(PcgState* st, double rz) {
    double beta = 0.0;
    if (!st->first) {
        beta = rz / st->rz_old;
        if (!(beta >= 0.0) || (st->replaced && !(beta < 1.0 / (st->delta * st->delta)))) {
            st->first = 1;
            beta = 0.0;
        }
    }
    st->beta = beta;
    st->rz_old = rz;
    st->replaced = 0;
}

// Last block of an FP64 residual replacement: converged on the true residual,
// out of iterations, or continue with the replaced residual keeping the search
// direction (beta from the new r.z).
__device__ inline void resid_finish(PcgState* st, double rr, double rz) {
    const double rn = sqrt(rr);
    st->true_rel = rn / st->bnorm;
    st->outer += 1;
    st->pending = 0;           // the fold before this replacement applied it
    const bool progress = st->true_rel < 0.99 * st->best_true;   // false for NaN
    if (st->true_rel < st->tol) {
        st->status = ST_DONE;
        st->iters = st->k - 1;
        st->rel = st->true_rel;
    } else if (st->k - 1 >= st->maxiter) {
        st->status = ST_MAXITER;
        st->iters = st->k - 1;
    } else if (!progress && ++st->stall >= REFINE_STALL) {
        st->status = ST_MAXITER;       // stagnated (or not finite)
        st->stalled = 1;
        st->iters = st->k - 1;
    } else {
        if (progress) {
            st->best_true = st->true_rel;
            st->stall = 0;
        }
        if (st->k - st->k_rep > st->max_phase) st->max_phase = st->k - st->k_rep;
        st->k_rep = st->k;
        st->rmax = rn;
        st->rel = st->true_rel;
        st->status = ST_RUNNING;
        if (st->mg) {
            // beta and rz_old follow from the cycle's z of the replaced residual
            // (k_mg_rz); a replacement without progress restarts the direction;
            // the next update kernel folds x32 into u
            st->replaced = 1;
            st->fold = 1;
            if (!progress) st->first = 1;
            return;
        }
        const double b = st->first ? 0.0 : rz / st->rz_old;
        if (!st->first && (!progress || !(b < 1.0 / (st->delta * st->delta)))) {
            // no progress since the previous replacement: keeping the direction
            // would carry (and, with beta >= 1/delta^2, amplify) a direction that
            // no longer reduces the true residual, so restart it (p = z)
            st->first = 1;
            st->beta = 0.0;
        } else {
            st->beta = b;
        }
        st->rz_old = rz;
        st->rel = st->true_rel;
        st->status = ST_RUNNING;
    }
}

// Halo staging, split in two so a persistent block can issue the global loads
// of its next tile before computing the current one.
//
// The input vector over the block's node brick plus a one-node halo goes to
// LDS node-interleaved ([k][i][j][c], the global DOF order: a staged row is one
// contiguous LDS row and lane-adjacent nodes sit 3 words apart).  A halo row
// (fixed i, k) is 3*HJ contiguous words in global memory; each thread owns one
// word position w of a 32- or 64-lane row slot.  Rows are dealt to the RPP
// slots so that the first HK passes take row (ii = slot, kk = pass): the
// j and i parts of the address and their range tests are computed once per
// thread and the k part is uniform per pass.  The remaining (HI - RPP) * HK
// rows (ii >= RPP) follow in XP more passes.  Every access is a bounds-checked
// raw buffer access: a word outside the domain gets an out-of-range offset, so
// it loads 0 and needs no branch, and the direction p written for owned nodes
// uses the same trick for the nodes a block does not own.  MODE_DIR forms
// p = dinv*r + beta*p_old.  The (TJ+1)(TI+1)(TK+1) element scales (0 outside
// the mesh) are staged the same way.  Byte offsets are 32-bit (the library
// rejects grids with 2^28 or more DOFs).
template <typename T, int TJ, int TI, int TK, int NT>
struct Halo {
    static constexpr int HJ = TJ + 2, HI = TI + 2, HK = TK + 2;
    static constexpr int ROW = 3 * HJ;
    static constexpr int LW = ROW <= 32 ? 32 : 64;
    static constexpr int RPP = NT / LW;
    static_assert(ROW <= 64, "halo row longer than 64 words");
    // RPP <= HI: row (ii = slot, kk = pass) for the first HK passes, then the
    // XR rows with ii >= RPP.  RPP > HI: plain row = slot + RPP * pass.
    static constexpr bool SPLIT = RPP <= HI;
    static constexpr int XW = SPLIT ? HI - RPP : 0;      // extra rows per plane
    static constexpr int XR = XW * HK;                   // extra rows in all
    static constexpr int XP = (XR + RPP - 1) / RPP;      // extra passes
    static constexpr int PER = SPLIT ? HK + XP : (HI * HK + RPP - 1) / RPP;
    static constexpr int EJ = TJ + 1, EI = TI + 1, EK = TK + 1;
    static constexpr int NE = EJ * EI * EK;
    static constexpr int PES = (NE + NT - 1) / NT;
    T v0[PER], v1[PER], v2[PER], xv[PER];
    int own[PER];     // byte offset of an owned node's word, OOB otherwise
    T sv[PES];
};

// Halo row (ii, kk) of pass u for row slot `slot`.
template <int HI, int HK, int RPP>
__device__ __forceinline__ void halo_row(int u, int slot, int& ii, int& kk, bool& valid) {
    if constexpr (RPP > HI) {
        const int row = slot + RPP * u;
        ii = row % HI; kk = row / HI; valid = row < HI * HK;
        return;
    }
    constexpr int XW = RPP > HI ? 0 : HI - RPP;
    if (u < HK) {
        ii = slot; kk = u; valid = true;
    } else {
        const int e = (u - HK) * RPP + slot;
        if constexpr (XW > 0) {
            ii = RPP + e % XW; kk = e / XW; valid = e < XW * HK;
        } else {
            ii = 0; kk = 0; valid = false;
        }
    }
}

template <typename T, int TJ, int TI, int TK, int NT, int MODE>
__device__ __forceinline__ void halo_load(const NodeArgs<T>& A, int j0, int i0, int k0, int first, int pending,
                                          const T* p_old, Halo<T, TJ, TI, TK, NT>& H) {
    typedef Halo<T, TJ, TI, TK, NT> Hc;
    const GridDev& g = A.g;
    const unsigned vbytes = unsigned(g.ndof) * sizeof(T);
    const int w = threadIdx.x % Hc::LW, slot = threadIdx.x / Hc::LW;
    const int jj = w / 3, c = w - 3 * (w / 3);
    const int j = j0 - 1 + jj;
    const bool jok = (w < Hc::ROW) && j >= 0 && j < g.NJ;
    const bool jown = jj >= 1 && jj <= TJ;
    const int plane3 = 3 * g.NI * g.NJ;
    const int wofs = 3 * j + c;
#pragma unroll
    for (int u = 0; u < Hc::PER; ++u) {
        int ii, kk;
        bool rv;
        halo_row<Hc::HI, Hc::HK, Hc::RPP>(u, slot, ii, kk, rv);
        const int i = i0 - 1 + ii, k = k0 - 1 + kk;
        const bool in = jok && rv && i >= 0 && i < g.NI && k >= 0 && k < g.NK;
        const int off = in ? int(sizeof(T)) * (k * plane3 + 3 * i * g.NJ + wofs) : OOB;
        const bool owned = in && jown && ii >= 1 && ii <= TI && kk >= 1 && kk <= TK;
        H.own[u] = owned ? off : OOB;
        if constexpr (is_dir(MODE)) {
            H.v0[u] = Buf<T>::ld(buf_rsrc(A.r, vbytes), off);
            H.v1[u] = Buf<T>::ld(buf_rsrc(A.dinv, vbytes), off);
            H.v2[u] = first ? T(0) : Buf<T>::ld(buf_rsrc(p_old, vbytes), off);
            if constexpr (MODE == MODE_DIRX) H.xv[u] = Buf<T>::ld(buf_rsrc(A.x, vbytes), pending ? H.own[u] : OOB);
        } else {
            H.v0[u] = Buf<T>::ld(buf_rsrc(A.u, vbytes), off);
        }
    }
    const unsigned sbytes = unsigned(g.nx) * unsigned(g.ny) * unsigned(g.nz) * sizeof(T);
#pragma unroll
    for (int u = 0; u < Hc::PES; ++u) {
        const int idx = threadIdx.x + u * NT;
        const int ej = idx % Hc::EJ, ei = (idx / Hc::EJ) % Hc::EI, ek = idx / (Hc::EJ * Hc::EI);
        const int je = j0 - 1 + ej, ie = i0 - 1 + ei, ke = k0 - 1 + ek;
        const bool in = idx < Hc::NE && je >= 0 && je < g.ny && ie >= 0 && ie < g.nx && ke >= 0 && ke < g.nz;
        H.sv[u] = Buf<T>::ld(buf_rsrc(A.s_e, sbytes), in ? int(sizeof(T)) * ((ke * g.nx + ie) * g.ny + je) : OOB);
    }
}

template <typename T, int TJ, int TI, int TK, int NT, int MODE>
__device__ __forceinline__ void halo_store(const Halo<T, TJ, TI, TK, NT>& H, T beta, T* p_new, T alpha_prev,
                                           int pending, T* x, unsigned vbytes, T* sP, T* sS, T* sZero) {
    typedef Halo<T, TJ, TI, TK, NT> Hc;
    const int w = threadIdx.x % Hc::LW, slot = threadIdx.x / Hc::LW;
#pragma unroll
    for (int u = 0; u < Hc::PER; ++u) {
        int ii, kk;
        bool rv;
        halo_row<Hc::HI, Hc::HK, Hc::RPP>(u, slot, ii, kk, rv);
        T val;
        if constexpr (is_dir(MODE)) {
            val = H.v1[u] * H.v0[u] + beta * H.v2[u];
            Buf<T>::st(val, buf_rsrc(p_new, vbytes), H.own[u]);
            // the previous iteration's x += alpha p_old, for the owned nodes
            if constexpr (MODE == MODE_DIRX)
                Buf<T>::st(H.xv[u] + alpha_prev * H.v2[u], buf_rsrc(x, vbytes), pending ? H.own[u] : OOB);
        } else {
            val = H.v0[u];
        }
        if (w < Hc::ROW && rv) {
            const int li = (kk * Hc::HI + ii) * Hc::ROW + w;
            sP[li] = val;
            if (sZero) sZero[li] = T(0);
        }
    }
#pragma unroll
    for (int u = 0; u < Hc::PES; ++u) {
        const int idx = threadIdx.x + u * NT;
        if (idx < Hc::NE) sS[idx] = H.sv[u];
    }
}

// Tile range of a persistent block: the tiles are split into NXCD contiguous
// ranges, one per XCD (blocks are dispatched to the XCDs round-robin), so the
// blocks sharing an L2 work on neighbouring bricks and reuse each other's
// halos; inside a range the blocks of that XCD take every nloc-th tile.
constexpr int NXCD = 8;
__device__ __forceinline__ void tile_range(int ntiles, int& first_tile, int& end_tile, int& step) {
    const int nx = (gridDim.x % NXCD == 0) ? NXCD : 1;
    const int xcd = blockIdx.x % nx, loc = blockIdx.x / nx, nloc = gridDim.x / nx;
    const int b = int((long long)ntiles * xcd / nx), e = int((long long)ntiles * (xcd + 1) / nx);
    first_tile = b + loc;
    end_tile = e;
    step = nloc;
}

// Node-centric K*u over the tiles of a persistent block (grid <= tiles).  With
// PF the next brick's global loads are issued before the current brick is
// computed (more registers, fewer exposed latencies).
// Outputs of one node (global DOF gn of component 0): y, or q and p.q, or the
// FP64 true residual with its FP32 copy and norms.
template <typename T, int MODE>
__device__ __forceinline__ void node_output(const NodeArgs<T>& A, int gn, const T (&qv)[3], const T (&ctr)[3],
                                            double (&part)[2]) {
    const GridDev& g = A.g;
    const unsigned vbytes = unsigned(g.ndof) * sizeof(T);
    if constexpr (MODE == MODE_PLAIN) {
#pragma unroll
        for (int c = 0; c < 3; ++c) Buf<T>::st(qv[c], buf_rsrc(A.y, vbytes), int(sizeof(T)) * (gn + c));
    } else if constexpr (is_dir(MODE)) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            Buf<T>::st(qv[c], buf_rsrc(A.q, vbytes), int(sizeof(T)) * (gn + c));
            part[0] += double(ctr[c]) * double(qv[c]);
        }
    } else if constexpr (MODE == MODE_MGRES) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int q = gn + c;
            Buf<T>::st(A.dinv[q] != T(0) ? A.cb_b[q] - qv[c] : T(0), buf_rsrc(A.y, vbytes), int(sizeof(T)) * q);
        }
    } else if constexpr (MODE == MODE_CHEB0) {
        const T ca = T(A.cheb[1]), cb = T(A.cheb[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int q = gn + c;
            const T dn = ca * ctr[c] + cb * A.dinv[q] * (A.cb_b[q] - qv[c]);
            A.cb_d[q] = dn;
            A.cb_out[q] = ctr[c] + dn;
        }
    } else if constexpr (MODE == MODE_CHEB) {
        const T ca = A.cheb_step ? T(A.cheb[2 * A.cheb_step - 1]) : T(0), cb = T(A.cheb[2 * A.cheb_step]);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int q = gn + c;
            const T b = A.cb_b[q];
            const T dn = (A.cheb_step ? ca * A.cb_d[q] : T(0)) + cb * A.dinv[q] * (b - qv[c]);
            A.cb_d[q] = dn;
            const T xo = ctr[c] + dn;
            A.cb_out[q] = xo;
            part[0] += double(b) * double(xo);
        }
    } else {  // MODE_RESID (T = double)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double msk = A.dinv[gn + c];
            const double r64 = (msk != 0.0) ? (A.f[gn + c] - double(qv[c])) : 0.0;
            const float r32 = float(r64);
            A.r32[gn + c] = r32;
            part[0] += r64 * r64;
            part[1] += double(r32) * double(A.dinv32[gn + c] * r32);
        }
    }
}

// Per-node operands of the fused outputs beyond the product (b, d, dinv of a
// Chebyshev step; b and dinv of the masked residual; f, the mask and dinv32 of
// the FP64 residual), loaded ahead of their use by the marching sector kernel:
// a step's node phase then finds them in registers instead of waiting on a
// memory round trip.  Modes without such operands keep nothing.
template <int MODE> __host__ __device__ constexpr bool has_node_ex() {
    return MODE == MODE_CHEB || MODE == MODE_CHEB0 || MODE == MODE_MGRES || MODE == MODE_RESID;
}
template <typename T>
struct NodeEx {
    T a[3], b[3], c[3];
};

// Loads of node gn's operands (gn = DOF of component 0; off = false: out of
// range, the loads return 0 and are not used).  Branch-free buffer loads.
template <typename T, int MODE>
__device__ __forceinline__ void node_ex_load(const NodeArgs<T>& A, int gn, bool on, NodeEx<T>& e) {
    if constexpr (has_node_ex<MODE>()) {
        const unsigned vbytes = unsigned(A.g.ndof) * sizeof(T);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int o = on ? int(sizeof(T)) * (gn + c) : OOB;
            if constexpr (MODE == MODE_RESID) {
                // T = double: f, the mask (FP64 dinv), dinv32 (FP32, widened)
                e.a[c] = Buf<T>::ld(buf_rsrc(A.f, vbytes), o);
                e.c[c] = Buf<T>::ld(buf_rsrc(A.dinv, vbytes), o);
                e.b[c] = T(Buf<float>::ld(buf_rsrc(A.dinv32, unsigned(A.g.ndof) * sizeof(float)),
                                          on ? int(sizeof(float)) * (gn + c) : OOB));
            } else {
                e.a[c] = Buf<T>::ld(buf_rsrc(A.cb_b, vbytes), o);
                e.c[c] = Buf<T>::ld(buf_rsrc(A.dinv, vbytes), o);
                if constexpr (MODE == MODE_CHEB)
                    e.b[c] = Buf<T>::ld(buf_rsrc(A.cb_d, vbytes), A.cheb_step ? o : OOB);
                else
                    e.b[c] = T(0);
            }
        }
    }
}

// node_output with the operands of NodeEx (the same arithmetic, in the same
// order, as node_output's loads; other modes go to node_output).
template <typename T, int MODE>
__device__ __forceinline__ void node_output_ex(const NodeArgs<T>& A, int gn, const T (&qv)[3], const T (&ctr)[3],
                                               const NodeEx<T>& e, double (&part)[2]) {
    if constexpr (MODE == MODE_MGRES) {
        const unsigned vbytes = unsigned(A.g.ndof) * sizeof(T);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int q = gn + c;
            Buf<T>::st(e.c[c] != T(0) ? e.a[c] - qv[c] : T(0), buf_rsrc(A.y, vbytes), int(sizeof(T)) * q);
        }
    } else if constexpr (MODE == MODE_CHEB0) {
        const T ca = T(A.cheb[1]), cb = T(A.cheb[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int q = gn + c;
            const T dn = ca * ctr[c] + cb * e.c[c] * (e.a[c] - qv[c]);
            A.cb_d[q] = dn;
            A.cb_out[q] = ctr[c] + dn;
        }
    } else if constexpr (MODE == MODE_CHEB) {
        const T ca = A.cheb_step ? T(A.cheb[2 * A.cheb_step - 1]) : T(0), cb = T(A.cheb[2 * A.cheb_step]);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int q = gn + c;
            const T b = e.a[c];
            const T dn = (A.cheb_step ? ca * e.b[c] : T(0)) + cb * e.c[c] * (b - qv[c]);
            A.cb_d[q] = dn;
            const T xo = ctr[c] + dn;
            A.cb_out[q] = xo;
            part[0] += double(b) * double(xo);
        }
    } else if constexpr (MODE == MODE_RESID) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double msk = e.c[c];
            const double r64 = (msk != 0.0) ? (double(e.a[c]) - double(qv[c])) : 0.0;
            const float r32 = float(r64);
            A.r32[gn + c] = r32;
            part[0] += r64 * r64;
            part[1] += double(r32) * double(float(e.b[c]) * r32);
        }
    } else {
        node_output<T, MODE>(A, gn, qv, ctr, part);
    }
}

// One brick of the node-centric K*u: stage (unless PF staged it already),
// compute the owned nodes, write the outputs, accumulate the fused partials.
template <typename T, int TJ, int TI, int TK, int MODE>
__device__ __forceinline__ void node_brick(const NodeArgs<T>& A, int j0, int i0, int k0, const T* kp, T* sP,
                                           const T* sS, double (&part)[2]) {
    constexpr int NT = TJ * TI * TK;
    typedef Halo<T, TJ, TI, TK, NT> Hc;
    constexpr int HJ = Hc::HJ, HI = Hc::HI;
    constexpr int EJ = Hc::EJ, EI = Hc::EI;
    const GridDev& g = A.g;
    const int tj = threadIdx.x % TJ;
    const int ti = (threadIdx.x / TJ) % TI;
    const int tk = threadIdx.x / (TJ * TI);
    const int j = j0 + tj, i = i0 + ti, k = k0 + tk;
    if (!(j < g.NJ && i < g.NI && k < g.NK)) return;
    // Contract (u_b - u_n) instead of u_b: K0 annihilates rigid translations, so
    // the product is unchanged, but rounding now scales with the local variation
    // of u instead of its magnitude (matters in FP32).  Brick symmetry: K0 block
    // (a,b) = R_a K0 block (0, a^b) R_a with R_a the reflection taking corner a
    // to corner 0, so the 64 blocks come from the first three rows of K0 (72
    // values) with compile-time signs.  The input component d is the outer
    // loop: 27 neighbour values and 8 table values are live at a time.
    // hb: LDS word of the (-1,-1,-1) neighbour, so every neighbour offset is a
    // non-negative constant (folded into the LDS instruction's immediate)
    const int hb = 3 * (tk * HI * HJ + ti * HJ + tj);
    constexpr int CTR = 3 * (HI * HJ + HJ + 1);
    T ctr[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) ctr[d] = sP[hb + CTR + d];
    T acc[8][3];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[a][c] = T(0);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        T nb[3][3][3];
#pragma unroll
        for (int dz = 0; dz < 3; ++dz)
#pragma unroll
            for (int dx = 0; dx < 3; ++dx)
#pragma unroll
                for (int dy = 0; dy < 3; ++dy)
                    nb[dz][dx][dy] = sP[hb + 3 * (dz * HI * HJ + dx * HJ + dy) + d] - ctr[d];
        // the table slice of this d is loaded here, not hoisted: all 72 values of
        // an FP64 table would not fit in the scalar registers
        const T* kpd = kp;
        asm volatile("" : "+s"(kpd));
        const auto K0d = as_const(kpd);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T kt[8];
#pragma unroll
            for (int x = 0; x < 8; ++x) kt[x] = K0d[c * 24 + 3 * x + d];
#pragma unroll
            for (int a = 0; a < 8; ++a)
#pragma unroll
                for (int b = 0; b < 8; ++b) {
                    if (b == a) continue;   // u_n - u_n = 0: the self block contributes nothing
                    const int dz = 1 + OZ(b) - OZ(a), dx = 1 + OX(b) - OX(a), dy = 1 + OY(b) - OY(a);
                    const T kv = kt[corner_xor(a, b)];
                    if (refl_sign(a, c) * refl_sign(a, d) > 0)
                        acc[a][c] += kv * nb[dz][dx][dy];
                    else
                        acc[a][c] -= kv * nb[dz][dx][dy];
                }
        }
    }
    T qv[3] = {T(0), T(0), T(0)};
    const int sb = (tk * EI + ti) * EJ + tj;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const T s = sS[sb + ((1 - OZ(a)) * EI + (1 - OX(a))) * EJ + (1 - OY(a))];
#pragma unroll
        for (int c = 0; c < 3; ++c) qv[c] += s * acc[a][c];
    }
    node_output<T, MODE>(A, 3 * ((k * g.NI + i) * g.NJ + j), qv, ctr, part);
}

template <typename T, int MODE>
__device__ __forceinline__ void node_finish(const NodeArgs<T>& A, double (&part)[2], double* sRed) {
    if constexpr (is_dir(MODE)) {
        double v[1] = {part[0]}, tot[1];
        if (reduce_last_block<1>(v, A.partials, &A.st->cnt[0], tot, sRed)) {
            if (threadIdx.x == 0) dir_finish(A.st, tot[0]);
        }
    } else if constexpr (MODE == MODE_CHEB) {
        if (A.cheb_rz) {
            double v[1] = {part[0]}, tot[1];
            if (reduce_last_block<1>(v, A.partials, &A.st->cnt[3], tot, sRed)) {
                if (threadIdx.x == 0) mg_rz_finish(A.st, tot[0]);
            }
        }
    } else if constexpr (MODE == MODE_RESID) {
        double v[2] = {part[0], part[1]}, tot[2];
        if (reduce_last_block<2>(v, A.partials, &A.st->cnt[2], tot, sRed)) {
            if (threadIdx.x == 0) resid_finish(A.st, tot[0], tot[1]);
        }
    }
}

template <typename T, int MODE>
__device__ __forceinline__ bool node_setup(const NodeArgs<T>& A, T& beta, int& first, T*& p_new, const T*& p_old,
                                           T& alpha_prev, int& pending) {
    beta = T(0);
    first = 1;
    p_new = nullptr;
    p_old = nullptr;
    alpha_prev = T(0);
    pending = 0;
    if constexpr (MODE == MODE_RESID) {
        if (A.only_on_replace == 1 && A.st->status != ST_REPLACE) return false;
        if (A.only_on_replace == 2 && A.st->status != ST_RUNNING) return false;
    }
    if constexpr (MODE == MODE_PLAIN || MODE == MODE_CHEB || MODE == MODE_CHEB0 || MODE == MODE_MGRES) {
        // a product inside a solve's captured iterations (the multigrid cycle)
        // passes the solve state and does nothing once the solve has stopped
        if (A.st && A.st->status != ST_RUNNING) return false;
    }
    if constexpr (MODE == MODE_CHEB0) beta = T(A.cheb[0]);     // the staging scale c0
    if constexpr (is_dir(MODE)) {
        if (A.st->status != ST_RUNNING) return false;
        beta = T(A.st->beta);
        first = A.st->first;
        alpha_prev = T(A.st->alpha);
        pending = A.st->pending;
        const bool odd = (A.st->k & 1) != 0;
        p_new = odd ? A.p1 : A.p0;
        p_old = odd ? A.p0 : A.p1;
    }
    return true;
}

// Node-centric K*u, one brick per block (grid = bricks).
template <typename T, int TJ, int TI, int TK, int MODE>
__global__ void __launch_bounds__(TJ * TI * TK)
k_node(NodeArgs<T> A) {
    constexpr int NT = TJ * TI * TK;
    typedef Halo<T, TJ, TI, TK, NT> Hc;
    __shared__ T sP[3 * Hc::HJ * Hc::HI * Hc::HK];
    __shared__ T sS[Hc::NE];
    __shared__ double sRed[16 * 4];
    T beta, alpha_prev;
    int first, pending;
    T* p_new;
    const T* p_old;
    if (!node_setup<T, MODE>(A, beta, first, p_new, p_old, alpha_prev, pending)) return;
    int j0, i0, k0;
    tile_origin(A.g, xcd_local_tile(blockIdx.x, gridDim.x), TJ, TI, TK, j0, i0, k0);
    {
        Hc H;
        halo_load<T, TJ, TI, TK, NT, MODE>(A, j0, i0, k0, first, pending, p_old, H);
        halo_store<T, TJ, TI, TK, NT, MODE>(H, beta, p_new, alpha_prev, pending, A.x, unsigned(A.g.ndof) * sizeof(T),
                                            sP, sS, (T*)nullptr);
    }
    __syncthreads();
    double part[2] = {0.0, 0.0};
    node_brick<T, TJ, TI, TK, MODE>(A, j0, i0, k0, A.K0, sP, sS, part);
    node_finish<T, MODE>(A, part, sRed);
}

// Half of the node-centric product of one node: the four elements on one side
// of it in z (H = 0: the elements above the node, where it is a bottom corner,
// OZ(a) = 0; H = 1: the elements below it).  Same arithmetic as node_brick
// restricted to those elements; their corners lie in two halo planes, and in
// both cases the plane of corner b is OZ(b) planes above the first one.
// Returns the scaled contribution of the four elements in qv and the node's
// own value in ctr.
template <typename T, int TJ, int TI, int TK, int H>
__device__ __forceinline__ void node_half(const T* kp, const T* sP, const T* sS, int tj, int ti, int tk,
                                          T (&ctr)[3], T (&qv)[3]) {
    constexpr int HJ = TJ + 2, HI = TI + 2;
    constexpr int EJ = TJ + 1, EI = TI + 1;
    constexpr int PZ0 = H == 0 ? 1 : 0;      // first halo plane (counted from the node's -1 plane)
    const int hb = 3 * (tk * HI * HJ + ti * HJ + tj);
    constexpr int CTR = 3 * (HI * HJ + HJ + 1);
#pragma unroll
    for (int d = 0; d < 3; ++d) ctr[d] = sP[hb + CTR + d];
    T acc[4][3];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[a][c] = T(0);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        T nb[2][3][3];
#pragma unroll
        for (int pz = 0; pz < 2; ++pz)
#pragma unroll
            for (int dx = 0; dx < 3; ++dx)
#pragma unroll
                for (int dy = 0; dy < 3; ++dy)
                    nb[pz][dx][dy] = sP[hb + 3 * ((PZ0 + pz) * HI * HJ + dx * HJ + dy) + d] - ctr[d];
        // the table slice of this d is loaded here, not hoisted: all 72 values of
        // an FP64 table would not fit in the scalar registers
        const T* kpd = kp;
        asm volatile("" : "+s"(kpd));
        const auto K0d = as_const(kpd);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T kt[8];
#pragma unroll
            for (int x = 0; x < 8; ++x) kt[x] = K0d[c * 24 + 3 * x + d];
#pragma unroll
            for (int ai = 0; ai < 4; ++ai) {
                const int a = 4 * H + ai;
#pragma unroll
                for (int b = 0; b < 8; ++b) {
                    if (b == a) continue;
                    const int dx = 1 + OX(b) - OX(a), dy = 1 + OY(b) - OY(a);
                    const T kv = kt[corner_xor(a, b)];
                    if (refl_sign(a, c) * refl_sign(a, d) > 0)
                        acc[ai][c] += kv * nb[OZ(b)][dx][dy];
                    else
                        acc[ai][c] -= kv * nb[OZ(b)][dx][dy];
                }
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) qv[c] = T(0);
    const int sb = (tk * EI + ti) * EJ + tj;
#pragma unroll
    for (int ai = 0; ai < 4; ++ai) {
        const int a = 4 * H + ai;
        const T sc = sS[sb + ((1 - OZ(a)) * EI + (1 - OX(a))) * EJ + (1 - OY(a))];
#pragma unroll
        for (int c = 0; c < 3; ++c) qv[c] += sc * acc[ai][c];
    }
}

// Node-centric K*u with two threads per node, one brick per block: waves
// [0, NN/64) take the elements above their nodes and waves [NN/64, 2NN/64)
// the elements below (the half is wave-uniform, so the two code paths never
// diverge inside a wave); the lower half leaves its sums in LDS and the upper
// half adds them (fixed order: deterministic) and writes the outputs.  Half
// the registers and work per thread of k_node, twice the waves.
template <typename T> struct SplitWaves;
template <> struct SplitWaves<double> { static constexpr int value = 4; };
template <> struct SplitWaves<float> { static constexpr int value = 8; };

template <typename T, int TJ, int TI, int TK, int MODE>
__global__ void __launch_bounds__(2 * TJ * TI * TK) __attribute__((amdgpu_waves_per_eu(SplitWaves<T>::value)))
k_node2(NodeArgs<T> A) {
    constexpr int NN = TJ * TI * TK;
    constexpr int NT = 2 * NN;
    static_assert(NN % 64 == 0, "a half must be whole waves");
    typedef Halo<T, TJ, TI, TK, NT> Hc;
    __shared__ T sP[3 * Hc::HJ * Hc::HI * Hc::HK];
    __shared__ T sS[Hc::NE];
    __shared__ T sQ[3 * NN];
    __shared__ double sRed[16 * 4];
    T beta, alpha_prev;
    int first, pending;
    T* p_new;
    const T* p_old;
    if (!node_setup<T, MODE>(A, beta, first, p_new, p_old, alpha_prev, pending)) return;
    int j0, i0, k0;
    tile_origin(A.g, xcd_local_tile(blockIdx.x, gridDim.x), TJ, TI, TK, j0, i0, k0);
    {
        Hc H;
        halo_load<T, TJ, TI, TK, NT, MODE>(A, j0, i0, k0, first, pending, p_old, H);
        halo_store<T, TJ, TI, TK, NT, MODE>(H, beta, p_new, alpha_prev, pending, A.x, unsigned(A.g.ndof) * sizeof(T),
                                            sP, sS, (T*)nullptr);
    }
    __syncthreads();
    const int half = threadIdx.x / NN;
    const int nidx = threadIdx.x - half * NN;
    const int tj = nidx % TJ, ti = (nidx / TJ) % TI, tk = nidx / (TJ * TI);
    T ctr[3], qv[3];
    // nodes outside the grid compute on the zero halo (in-bounds LDS reads) and write nothing
    if (half) {
        node_half<T, TJ, TI, TK, 1>(A.K0, sP, sS, tj, ti, tk, ctr, qv);
#pragma unroll
        for (int c = 0; c < 3; ++c) sQ[3 * nidx + c] = qv[c];
    } else {
        node_half<T, TJ, TI, TK, 0>(A.K0, sP, sS, tj, ti, tk, ctr, qv);
    }
    __syncthreads();
    double part[2] = {0.0, 0.0};
    const GridDev& g = A.g;
    const int j = j0 + tj, i = i0 + ti, k = k0 + tk;
    if (!half && j < g.NJ && i < g.NI && k < g.NK) {
#pragma unroll
        for (int c = 0; c < 3; ++c) qv[c] += sQ[3 * nidx + c];
        node_output<T, MODE>(A, 3 * ((k * g.NI + i) * g.NJ + j), qv, ctr, part);
    }
    node_finish<T, MODE>(A, part, sRed);
}

// Node-centric K*u, persistent: a fixed grid whose blocks walk contiguous,
// XCD-local runs of bricks (one partial per block; neighbouring bricks share an
// L2).  WAVES bounds the registers so that many waves stay resident.
template <typename T> struct PersistWaves;
template <> struct PersistWaves<double> { static constexpr int value = 3; };
template <> struct PersistWaves<float> { static constexpr int value = 4; };

template <typename T, int TJ, int TI, int TK, int MODE>
__global__ void __launch_bounds__(TJ * TI * TK) __attribute__((amdgpu_waves_per_eu(PersistWaves<T>::value)))
k_node_persist(NodeArgs<T> A) {
    constexpr int NT = TJ * TI * TK;
    typedef Halo<T, TJ, TI, TK, NT> Hc;
    __shared__ T sP[3 * Hc::HJ * Hc::HI * Hc::HK];
    __shared__ T sS[Hc::NE];
    __shared__ double sRed[16 * 4];
    T beta, alpha_prev;
    int first, pending;
    T* p_new;
    const T* p_old;
    if (!node_setup<T, MODE>(A, beta, first, p_new, p_old, alpha_prev, pending)) return;
    int tile, tend, tstep;
    tile_range(A.g.ntiles, tile, tend, tstep);
    double part[2] = {0.0, 0.0};
    for (; tile < tend; tile += tstep) {
        int j0, i0, k0;
        tile_origin(A.g, tile, TJ, TI, TK, j0, i0, k0);
        {
            Hc H;
            halo_load<T, TJ, TI, TK, NT, MODE>(A, j0, i0, k0, first, pending, p_old, H);
            halo_store<T, TJ, TI, TK, NT, MODE>(H, beta, p_new, alpha_prev, pending, A.x,
                                                unsigned(A.g.ndof) * sizeof(T), sP, sS, (T*)nullptr);
        }
        __syncthreads();
        // opaque per-brick copy of the K0 pointer keeps the table reads in the loop
        const T* kp = A.K0;
        asm volatile("" : "+s"(kp));
        node_brick<T, TJ, TI, TK, MODE>(A, j0, i0, k0, kp, sP, sS, part);
        __syncthreads();
    }
    node_finish<T, MODE>(A, part, sRed);
}

// Element-batched kernels stage with the same layout (non-persistent form).
template <typename T, int TJ, int TI, int TK, int MODE, int NT>
__device__ __forceinline__ void stage_halo(const NodeArgs<T>& A, int j0, int i0, int k0, T beta, int first,
                                           const T* p_old, T* p_new, T alpha_prev, int pending, T* sP, T* sS,
                                           T* sZero) {
    Halo<T, TJ, TI, TK, NT> H;
    halo_load<T, TJ, TI, TK, NT, MODE>(A, j0, i0, k0, first, pending, p_old, H);
    halo_store<T, TJ, TI, TK, NT, MODE>(H, beta, p_new, alpha_prev, pending, A.x, unsigned(A.g.ndof) * sizeof(T), sP,
                                        sS, sZero);
}

}  // namespace h8
