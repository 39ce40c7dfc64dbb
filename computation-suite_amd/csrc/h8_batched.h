// h8_batched.h -- many structured instances x many right-hand sides in one
// launch per PCG phase.
//
// A "system" is one (instance, load case) pair; its vectors are contiguous
// segments of the batch vectors (instance i, case c at vec_off[i] + c*ndof_i).
// Every launch covers every system; a system whose status has left RUNNING is
// skipped (its per-system convergence mask), and the host polls one device
// counter of still-running systems per captured chunk of iterations.
//
//   k_binit   : flat chunk table over systems; r = mask(b), |b|, |r|, r.z; the
//               last block of each system applies the early-out test.
//   k_bsector : flat table of (system, column tile, layer chunk); a block
//               runs the sector march of the single-system kernel on its
//               system (p = dinv*r + beta*p_old formed while staging, q = K p,
//               p.q partial); the last block of each system forms alpha (or
//               flags p.q <= 0).
//   k_bupdate : flat chunk table over systems; r update, r.z and r.r
//               partials; the last block of each system applies the stopping
//               rule of the single-system solver and retires finished systems.
//               (x += alpha p is applied by the next k_bsector while it stages
//               p_old, and after the loop by k_bfinal.)
#pragma once
#include "h8_device.h"
#include "h8_node.h"
#include "h8_sector.h"

namespace h8 {

struct BInst {            // device copy of one instance
    int nx, ny, nz, nrhs;
    long long node_off;   // dinv at 3*node_off
    long long elem_off;   // s_e at elem_off
    long long vec_off;    // case c at vec_off + c*ndof
    int sys_off, k0_off;  // first system; k0_off: index of the instance's K0 (x H8_K0_WORDS)
};
struct BSTile { int sys, tile; };   // one block of the batched sector kernel
struct BChunk { int sys; int start; int len; int pad; };

struct BSys {             // per-system solver state
    double rz_old, alpha, beta, bnorm, rel, bad_pq;
    long long k, iters;
    int status, first;
    int pending;          // x lacks alpha * p of the last direction kernel
    unsigned cnt;         // last-block counter of this system's chunks
    int inst;             // owning instance
    long long off;        // vector offset
    int ndof, case_idx;
};

struct BArgs {
    const BInst* inst;
    const BChunk* chunks;
    BSys* sys;
    int* active;          // systems still RUNNING
    double* partials;     // per system: pstride doubles
    int pstride;
    double tol;
    long long maxiter;
};

template <typename T>
struct BVecs {
    const T* s_e;
    const T* K0;
    const T* dinv;
    const T* b;
    T* x;
    T* r;
    T* p0;
    T* p1;
    T* q;
};

__device__ inline GridDev inst_grid(const BInst& I, int TJ, int TI, int TK) {
    GridDev g;
    g.nx = I.nx; g.ny = I.ny; g.nz = I.nz;
    g.NJ = I.ny + 1; g.NI = I.nx + 1; g.NK = I.nz + 1;
    g.nnodes = (long long)g.NJ * g.NI * g.NK;
    g.ndof = 3 * g.nnodes;
    g.tj = (g.NJ + TJ - 1) / TJ;
    g.ti = (g.NI + TI - 1) / TI;
    g.tk = (g.NK + TK - 1) / TK;
    g.ntiles = g.tj * g.ti * g.tk;
    g.variant = 0;
    return g;
}

// Sums NV partials of one system written by `count` contributors (fixed order).
template <int NV>
__device__ void sum_partials(const double* part, int count, double (&tot)[NV], double* sh) {
    double acc[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = 0.0;
    for (int b = threadIdx.x; b < count; b += blockDim.x) {
#pragma unroll
        for (int i = 0; i < NV; ++i) acc[i] += ld_relaxed(&part[(size_t)b * NV + i]);
    }
    block_sum<NV>(acc, sh);
#pragma unroll
    for (int i = 0; i < NV; ++i) tot[i] = acc[i];
}

// Ticket on a counter after this block's device-scope partial stores; true in
// the last of `expected` blocks (which also resets the counter).
__device__ inline bool take_ticket(unsigned* counter, unsigned expected) {
    __shared__ int s_last;
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0);
        const unsigned t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (t == expected - 1);
        if (s_last) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    return s_last != 0;
}

__device__ inline void retire(BArgs& B, BSys& S, int status) {
    S.status = status;
    atomicSub(B.active, 1);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_binit(BArgs B, BVecs<T> V) {
    __shared__ double sRed[16 * 3];
    const BChunk C = B.chunks[blockIdx.x];
    BSys& S = B.sys[C.sys];
    const BInst& I = B.inst[S.inst];
    const T* dinv = V.dinv + 3 * I.node_off;
    double v[3] = {0.0, 0.0, 0.0};
    for (int e = C.start + threadIdx.x; e < C.start + C.len; e += blockDim.x) {
        const T d = dinv[e];
        const T bi = (d != T(0)) ? V.b[S.off + e] : T(0);
        V.r[S.off + e] = bi;
        V.x[S.off + e] = T(0);
        v[0] += double(bi) * double(bi);
        v[1] += double(bi) * double(bi);
        v[2] += double(bi) * double(d * bi);
    }
    block_sum<3>(v, sRed);
    const int nch = (S.ndof + 2047) / 2048;
    double* part = B.partials + (size_t)C.sys * B.pstride;
    if (threadIdx.x == 0) {
        const int ci = C.start / 2048;
#pragma unroll
        for (int i = 0; i < 3; ++i) st_relaxed(&part[3 * ci + i], v[i]);
    }
    if (!take_ticket(&S.cnt, nch)) return;
    double tot[3];
    sum_partials<3>(part, nch, tot, sRed);
    if (threadIdx.x == 0) {
        double bn = sqrt(tot[0]);
        if (bn == 0.0) bn = 1.0;
        S.bnorm = bn;
        S.rel = sqrt(tot[1]) / bn;
        S.k = 1;
        S.first = 1;
        S.beta = 0.0;
        S.alpha = 0.0;
        S.pending = 0;
        if (S.rel < B.tol) {
            S.iters = 0;
            retire(B, S, ST_DONE);
        } else if (B.maxiter == 0) {
            S.iters = 0;
            retire(B, S, ST_MAXITER);
        } else {
            S.rz_old = tot[2];
            S.status = ST_RUNNING;
        }
    }
}

// Batched direction + K*p through the reflection sectors: one block per
// (system, column tile, layer chunk) of a flat table; blocks of systems no
// longer running exit at once.  Each block runs the march of the single-system
// sector kernel on its system's vectors and instance grid, writes its p.q
// partial, and the last block of the system (ticket on the system's counter)
// forms alpha or retires the system on p.q <= 0.
template <typename T, int EJ, int EI>
__global__ void __launch_bounds__(EJ * EI) __attribute__((amdgpu_waves_per_eu(SectorWaves<T>::value)))
k_bsector(BArgs B, BVecs<T> V, const BSTile* tiles, int kd) {
    __shared__ SectorSmem<T, EJ, EI> sm;
    __shared__ double sRed[16 * 4];
    const BSTile Tl = tiles[blockIdx.x];
    BSys& S = B.sys[Tl.sys];
    if (S.status != ST_RUNNING) return;
    const BInst& I = B.inst[S.inst];
    NodeArgs<T> A{};
    A.g = inst_grid(I, EJ, EI, kd);
    sector_tiles(A.g, EJ, EI, kd);
    A.s_e = V.s_e + I.elem_off;
    A.K0 = V.K0 + (size_t)I.k0_off * 648;   // H8_K0_WORDS
    A.dinv = V.dinv + 3 * I.node_off;
    A.r = V.r + S.off;
    A.p0 = V.p0 + S.off;
    A.p1 = V.p1 + S.off;
    A.q = V.q + S.off;
    A.x = V.x + S.off;
    const bool odd = (S.k & 1) != 0;
    T* p_new = odd ? A.p1 : A.p0;
    const T* p_old = odd ? A.p0 : A.p1;
    double part[2] = {0.0, 0.0};
    sector_march<T, EJ, EI, MODE_DIRX>(A, A.K0 + 576, kd, Tl.tile, T(S.beta), S.first, p_new, p_old, T(S.alpha),
                                      S.pending, sm, part);
    double v[1] = {part[0]};
    block_sum<1>(v, sRed);
    double* pp = B.partials + (size_t)Tl.sys * B.pstride;
    if (threadIdx.x == 0) st_relaxed(&pp[Tl.tile], v[0]);
    if (!take_ticket(&S.cnt, A.g.ntiles)) return;
    double tot[1];
    sum_partials<1>(pp, A.g.ntiles, tot, sRed);
    if (threadIdx.x == 0) {
        const double pq = tot[0];
        if (!(pq > 0.0)) {
            S.bad_pq = pq;
            S.iters = S.k;
            S.pending = 0;
            retire(B, S, ST_NOT_SPD);
        } else {
            S.alpha = S.rz_old / pq;
            S.pending = 1;       // x += alpha p: next direction kernel, or k_bfinal
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_bupdate(BArgs B, BVecs<T> V) {
    __shared__ double sRed[16 * 2];
    const BChunk C = B.chunks[blockIdx.x];
    BSys& S = B.sys[C.sys];
    if (S.status != ST_RUNNING) return;
    const BInst& I = B.inst[S.inst];
    const T* dinv = V.dinv + 3 * I.node_off;
    T* r = V.r + S.off;
    const T* q = V.q + S.off;
    const T alpha = T(S.alpha);
    double v[2] = {0.0, 0.0};
    for (int e = C.start + threadIdx.x; e < C.start + C.len; e += blockDim.x) {
        const T d = dinv[e];
        const T ri = (d != T(0)) ? T(r[e] - alpha * q[e]) : T(0);
        r[e] = ri;
        v[0] += double(ri) * double(d * ri);
        v[1] += double(ri) * double(ri);
    }
    block_sum<2>(v, sRed);
    const int nch = (S.ndof + 2047) / 2048;
    double* part = B.partials + (size_t)C.sys * B.pstride;
    if (threadIdx.x == 0) {
        const int ci = C.start / 2048;
        st_relaxed(&part[2 * ci + 0], v[0]);
        st_relaxed(&part[2 * ci + 1], v[1]);
    }
    if (!take_ticket(&S.cnt, nch)) return;
    double tot[2];
    sum_partials<2>(part, nch, tot, sRed);
    if (threadIdx.x == 0) {
        const double rel = sqrt(tot[1]) / S.bnorm;
        S.rel = rel;
        const long long k = S.k;
        if (rel < B.tol) {
            S.iters = k;
            retire(B, S, ST_DONE);
        } else if (k >= B.maxiter) {
            S.iters = k;
            retire(B, S, ST_MAXITER);
        } else {
            S.beta = tot[0] / S.rz_old;
            S.rz_old = tot[0];
            S.first = 0;
            S.k = k + 1;
        }
    }
}

// Persistent batched solve: one launch for the whole batch.  Every workgroup
// takes whole systems from a queue (`order`, largest first; `queue` the next
// index) and runs each one's Jacobi-PCG to the end with the recurrence and
// stopping rule of k_binit / k_bsector / k_bupdate: the sector march of the
// direction kernel over the instance's column tiles (x += alpha p_old applied
// while staging, as MODE_DIRX), p.q, alpha, the update, r.z and r.r, the
// stopping test, beta.  Every reduction stays inside the workgroup (per-thread
// sums in a fixed node order, then block_sum), so no system waits on another
// workgroup and a system's result does not depend on which workgroup runs it.
// A workgroup ends when the queue is empty.
template <typename T, int EJ, int EI>
__global__ void __launch_bounds__(EJ * EI) __attribute__((amdgpu_waves_per_eu(SectorWaves<T>::value)))
k_bsolve(BArgs B, BVecs<T> V, const int* __restrict__ order, int n_sys, int* queue, int kd) {
    constexpr int NT = EJ * EI;
    __shared__ SectorSmem<T, EJ, EI> sm;
    __shared__ double sRed[16 * 4];
    __shared__ int sSys;
    for (;;) {
        __syncthreads();                      // everyone has read the previous sSys
        if (threadIdx.x == 0) {
            const int t = atomicAdd(queue, 1);
            sSys = t < n_sys ? order[t] : -1;
        }
        __syncthreads();
        const int sid = sSys;
        if (sid < 0) return;
        BSys& S = B.sys[sid];
        const BInst& I = B.inst[S.inst];
        NodeArgs<T> A{};
        A.g = inst_grid(I, EJ, EI, kd);
        sector_tiles(A.g, EJ, EI, kd);
        A.s_e = V.s_e + I.elem_off;
        A.K0 = V.K0 + (size_t)I.k0_off * 648;   // H8_K0_WORDS
        A.dinv = V.dinv + 3 * I.node_off;
        A.r = V.r + S.off;
        A.p0 = V.p0 + S.off;
        A.p1 = V.p1 + S.off;
        A.q = V.q + S.off;
        A.x = V.x + S.off;
        const T* b = V.b + S.off;
        const T* dinv = A.dinv;
        T* r = V.r + S.off;
        T* x = A.x;
        const int n = S.ndof;
        // init: r = mask(b), x = 0, |b|^2, |r|^2, r.z
        double v3[3] = {0.0, 0.0, 0.0};
        for (int e = threadIdx.x; e < n; e += NT) {
            const T d = dinv[e];
            const T bi = (d != T(0)) ? b[e] : T(0);
            r[e] = bi;
            x[e] = T(0);
            v3[0] += double(bi) * double(bi);
            v3[1] += double(bi) * double(bi);
            v3[2] += double(bi) * double(d * bi);
        }
        block_sum<3>(v3, sRed);
        double bn = sqrt(v3[0]);
        if (bn == 0.0) bn = 1.0;
        double rel = sqrt(v3[1]) / bn, rz_old = v3[2], alpha = 0.0, beta = 0.0;
        long long k = 1, iters = 0;
        int first = 1, pending = 0, status = ST_RUNNING;
        if (rel < B.tol) {
            status = ST_DONE;
        } else if (B.maxiter == 0) {
            status = ST_MAXITER;
        }
        while (status == ST_RUNNING) {
            const bool odd = (k & 1) != 0;
            T* p_new = odd ? A.p1 : A.p0;
            const T* p_old = odd ? A.p0 : A.p1;
            double part[2] = {0.0, 0.0};
            for (int tile = 0; tile < A.g.ntiles; ++tile) {
                sector_march<T, EJ, EI, MODE_DIRX>(A, A.K0 + 576, kd, tile, T(beta), first, p_new, p_old, T(alpha),
                                                  pending, sm, part);
                __syncthreads();
            }
            double v1[1] = {part[0]};
            block_sum<1>(v1, sRed);               // (its barriers also publish q, p and x to the workgroup)
            const double pq = v1[0];
            if (!(pq > 0.0)) {
                S.bad_pq = pq;
                iters = k;
                pending = 0;
                status = ST_NOT_SPD;
                break;
            }
            alpha = rz_old / pq;
            pending = 1;                          // x += alpha p: next march, or after the loop
            const T ta = T(alpha);
            double v2[2] = {0.0, 0.0};
            for (int e = threadIdx.x; e < n; e += NT) {
                const T d = dinv[e];
                const T ri = (d != T(0)) ? T(r[e] - ta * A.q[e]) : T(0);
                r[e] = ri;
                v2[0] += double(ri) * double(d * ri);
                v2[1] += double(ri) * double(ri);
            }
            block_sum<2>(v2, sRed);
            rel = sqrt(v2[1]) / bn;
            if (rel < B.tol) {
                iters = k;
                status = ST_DONE;
            } else if (k >= B.maxiter) {
                iters = k;
                status = ST_MAXITER;
            } else {
                beta = v2[0] / rz_old;
                rz_old = v2[0];
                first = 0;
                ++k;
            }
        }
        if (pending) {
            // the last direction's x += alpha p (k_bfinal of the launched form)
            const T* p = ((k & 1) ? A.p1 : A.p0);
            const T ta = T(alpha);
            for (int e = threadIdx.x; e < n; e += NT) x[e] += ta * p[e];
        }
        if (threadIdx.x == 0) {
            S.status = status;
            S.iters = iters;
            S.rel = rel;
            S.k = k;
            S.alpha = alpha;
            S.beta = beta;
            S.rz_old = rz_old;
            S.bnorm = bn;
            S.pending = 0;
        }
    }
}

// After the batched loop: x += alpha p of the last direction kernel for every
// system whose update is still pending (flat chunk table).
template <typename T>
__global__ void __launch_bounds__(256)
k_bfinal(BArgs B, BVecs<T> V) {
    const BChunk C = B.chunks[blockIdx.x];
    const BSys& S = B.sys[C.sys];
    if (!S.pending) return;
    const T* p = ((S.k & 1) ? V.p1 : V.p0) + S.off;
    T* x = V.x + S.off;
    const T alpha = T(S.alpha);
    for (int e = C.start + threadIdx.x; e < C.start + C.len; e += blockDim.x) x[e] += alpha * p[e];
}

}  // namespace h8
