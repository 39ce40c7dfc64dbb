// h8_stream.h -- streaming (one-pass, coalesced) kernels of the PCG recurrence
// and the per-element kernels (Jacobi diagonal, SIMP scale, bilinear forms).
#pragma once
#include "h8_device.h"

namespace h8 {

constexpr int STREAM_THREADS = 256;

// r = mask(b - Ax0) (Ax0 may be null), reduces b.b, r.r, r.(dinv r); the last
// block applies the early-out test of fe/iterative_solvers.py:434-437.
template <typename T>
__global__ void __launch_bounds__(STREAM_THREADS)
k_pcg_init(long long n, const T* __restrict__ b, const T* __restrict__ Ax0, const T* __restrict__ dinv,
           T* __restrict__ r, PcgState* st, double* partials) {
    __shared__ double sRed[16 * 3];
    double v[3] = {0.0, 0.0, 0.0};
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const T d = dinv[i];
        const T bi = (d != T(0)) ? b[i] : T(0);
        const T ri = (d != T(0)) ? (Ax0 ? T(b[i] - Ax0[i]) : b[i]) : T(0);
        r[i] = ri;
        v[0] += double(bi) * double(bi);
        v[1] += double(ri) * double(ri);
        v[2] += double(ri) * double(d * ri);
    }
    double tot[3];
    if (reduce_last_block<3>(v, partials, &st->cnt[1], tot, sRed) && threadIdx.x == 0) {
        double bn = sqrt(tot[0]);
        if (bn == 0.0) bn = 1.0;
        st->bnorm = bn;
        st->rel = sqrt(tot[1]) / bn;
        st->k = 1;
        st->first = 1;
        st->beta = 0.0;
        st->alpha = 0.0;
        st->outer = 0;
        st->pending = 0;
        if (st->rel < st->tol) {
            st->status = ST_DONE;
            st->iters = 0;
        } else {
            st->rz_old = tot[2];
            st->status = ST_RUNNING;
        }
    }
}

// End of an iteration, given rz = r.z and rr = r.r of the updated residual.
// REFINED = false: the stopping rule / beta of fe/iterative_solvers.py:455-466.
// REFINED = true : additionally stops the FP32 loop for an FP64 residual
//                  replacement when ||r|| < delta * rmax (or the recurrence
//                  claims convergence, the phase is too long, or maxiter is
//                  reached).
// MG: the multigrid-preconditioned refined loop, whose beta comes later from
//     the cycle's z (k_mg_rz): only the replacement decision is taken here.
template <bool REFINED, bool MG = false>
__device__ inline void update_finish(PcgState* st, double rz, double rr) {
    const double rn = sqrt(rr);
    const double rel = rn / st->bnorm;
    st->rel = rel;
    const long long k = st->k;
    if (!REFINED) {
        if (rel < st->tol) {
            st->status = ST_DONE;
            st->iters = k;
        } else if (k >= st->maxiter) {
            st->status = ST_MAXITER;
            st->iters = k;
        } else {
            st->beta = rz / st->rz_old;
            st->rz_old = rz;
            st->first = 0;
            st->k = k + 1;
        }
    } else {
        st->first = 0;
        st->k = k + 1;
        const long long phase_cap = REFINE_PHASE > 4 * st->max_phase ? REFINE_PHASE : 4 * st->max_phase;
        if (rel < st->tol || rn < st->delta * st->rmax || k - st->k_rep > phase_cap || k >= st->maxiter) {
            st->status = ST_REPLACE;
        } else if (!MG) {
            st->beta = rz / st->rz_old;
            st->rz_old = rz;
        }
    }
}

// r = mask(r - alpha q) ; reduces r.z and r.r with z = dinv r.
// XUPD = false: the iterate update x += alpha p is deferred to the next
//               direction kernel, which has p in registers while staging it
//               (PcgState::pending), so this pass streams four vectors (r, q,
//               dinv in; r out) -- the form for large grids.
// XUPD = true : x += alpha p here as well (seven vectors) and the pending flag
//               is cleared -- the form for small grids, where the direction
//               kernel is latency-bound and the extra staging costs more.
// REFINED: the stopping rule of update_finish.
// Vectors are read and written in 16-byte groups (2 doubles / 4 floats), four
// groups per thread per pass (two when x is updated too) with all loads issued
// first; the workspace and torch allocations are 256-byte aligned, the tail is
// scalar.
// MG (with XUPD, T = float): when PcgState::fold is set, the last FP64 residual
// was formed at u + x (u64 the FP64 solution, x the FP32 iterate): this kernel
// folds x into u (u += x) and restarts x from alpha p, instead of x += alpha p.
template <typename T, bool REFINED, bool XUPD, bool MG = false>
__global__ void __launch_bounds__(STREAM_THREADS)
k_pcg_update(long long n, T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p0,
             const T* __restrict__ p1, const T* __restrict__ q, const T* __restrict__ dinv, PcgState* st,
             double* partials, double* __restrict__ u64) {
    constexpr int V = 16 / sizeof(T);
    constexpr int G = XUPD ? 2 : 4;
    typedef T vec __attribute__((ext_vector_type(V)));
    typedef double dvec2 __attribute__((ext_vector_type(2)));
    __shared__ double sRed[16 * 2];
    if (st->status != ST_RUNNING) return;
    const T alpha = T(st->alpha);
    const T* __restrict__ p = (st->k & 1) ? p1 : p0;
    const bool fold = MG && XUPD && st->fold;     // (uniform)
    double v[2] = {0.0, 0.0};
    const long long nvec = n / V;
    const long long gstride = (long long)gridDim.x * blockDim.x;
    for (long long iv0 = (long long)blockIdx.x * blockDim.x + threadIdx.x; iv0 < nvec; iv0 += G * gstride) {
        vec dv[G], rv[G], qv[G], xv[G], pv[G];
        bool ok[G];
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const long long iv = iv0 + u * gstride;
            ok[u] = iv < nvec;
            const long long j = ok[u] ? iv : iv0;
            dv[u] = reinterpret_cast<const vec*>(dinv)[j];
            rv[u] = reinterpret_cast<const vec*>(r)[j];
            qv[u] = reinterpret_cast<const vec*>(q)[j];
            if constexpr (XUPD) {
                xv[u] = reinterpret_cast<const vec*>(x)[j];
                pv[u] = reinterpret_cast<const vec*>(p)[j];
            }
        }
#pragma unroll
        for (int u = 0; u < G; ++u) {
            if (!ok[u]) continue;
            if constexpr (MG && XUPD && sizeof(T) == 4) {
                if (fold) {
                    // u64 += x (4 doubles per FP32 group, two 16-byte halves), x = alpha p
                    dvec2* ud = reinterpret_cast<dvec2*>(u64) + 2 * (iv0 + u * gstride);
                    dvec2 a = ud[0], b = ud[1];
                    a[0] += double(xv[u][0]); a[1] += double(xv[u][1]);
                    b[0] += double(xv[u][2]); b[1] += double(xv[u][3]);
                    ud[0] = a;
                    ud[1] = b;
#pragma unroll
                    for (int c = 0; c < V; ++c) xv[u][c] = T(0);
                }
            }
#pragma unroll
            for (int c = 0; c < V; ++c) {
                if constexpr (XUPD) xv[u][c] += alpha * pv[u][c];
                const T ri = (dv[u][c] != T(0)) ? T(rv[u][c] - alpha * qv[u][c]) : T(0);
                rv[u][c] = ri;
                v[0] += double(ri) * double(dv[u][c] * ri);
                v[1] += double(ri) * double(ri);
            }
            reinterpret_cast<vec*>(r)[iv0 + u * gstride] = rv[u];
            if constexpr (XUPD) reinterpret_cast<vec*>(x)[iv0 + u * gstride] = xv[u];
        }
    }
    for (long long i = nvec * V + (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gstride) {
        const T d = dinv[i];
        if constexpr (MG && XUPD) {
            if (fold) {
                u64[i] += double(x[i]);
                x[i] = T(0);
            }
        }
        if constexpr (XUPD) x[i] += alpha * p[i];
        const T ri = (d != T(0)) ? T(r[i] - alpha * q[i]) : T(0);
        r[i] = ri;
        v[0] += double(ri) * double(d * ri);
        v[1] += double(ri) * double(ri);
    }
    double tot[2];
    if (reduce_last_block<2>(v, partials, &st->cnt[1], tot, sRed) && threadIdx.x == 0) {
        if (XUPD) st->pending = 0;   // applied above
        if (MG && XUPD) st->fold = 0;   // every block has read it (they all finished)
        update_finish<REFINED, MG>(st, tot[0], tot[1]);
    }
}

// When the FP32 loop stopped for a residual replacement: u64 += double(x32)
// (+ alpha p32 when the last direction's update is still pending, p32 the
// direction of iteration k - 1); x32 = 0: fold the FP32 iterate into the FP64
// solution.  In any other state it changes nothing.
__global__ void __launch_bounds__(STREAM_THREADS)
k_fold(long long n, double* __restrict__ u, float* __restrict__ x32, const float* __restrict__ p0,
       const float* __restrict__ p1, const PcgState* st) {
    if (st->status != ST_REPLACE) return;
    const float* __restrict__ p32 = ((st->k - 1) & 1) ? p1 : p0;
    const double alpha = st->pending ? st->alpha : 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        u[i] += double(x32[i]) + alpha * double(p32[i]);
        x32[i] = 0.0f;
    }
}

// out = u64 + double(x32); x32 = 0 (the end of the multigrid-preconditioned
// refined solve: the FP32 iterate not yet folded by an update kernel, and the
// solution copied from the workspace to the caller's vector).
__global__ void __launch_bounds__(STREAM_THREADS)
k_fold_final(long long n, const double* __restrict__ u, float* __restrict__ x32, double* __restrict__ out) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        out[i] = u[i] + double(x32[i]);
        x32[i] = 0.0f;
    }
}

// x += alpha p when the last direction's update is still pending (after the
// last iteration of a solve).
template <typename T>
__global__ void __launch_bounds__(STREAM_THREADS)
k_pending_axpy(long long n, T* __restrict__ x, const T* __restrict__ p, const PcgState* st) {
    if (!st->pending) return;
    const T alpha = T(st->alpha);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        x[i] += alpha * p[i];
}

__global__ void __launch_bounds__(STREAM_THREADS)
k_cast_f64_f32(long long n, const double* __restrict__ x, float* __restrict__ y) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        y[i] = float(x[i]);
}

__global__ void __launch_bounds__(STREAM_THREADS)
k_masked_sub(long long n, const double* __restrict__ a, const double* __restrict__ b,
             const double* __restrict__ dinv, double* __restrict__ out) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        out[i] = (dinv[i] != 0.0) ? (b ? a[i] - b[i] : a[i]) : 0.0;
}

// The same for ncols columns of length n stored one after another (a, out);
// b (length n, may be NULL) and the mask are shared by every column.
__global__ void __launch_bounds__(STREAM_THREADS)
k_masked_sub_cols(long long n, long long ncols, const double* __restrict__ a, const double* __restrict__ b,
                  const double* __restrict__ dinv, double* __restrict__ out) {
    const long long total = n * ncols;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const long long i = t % n;
        out[t] = (dinv[i] != 0.0) ? (b ? a[t] - b[i] : a[t]) : 0.0;
    }
}

template <typename T>
__global__ void __launch_bounds__(STREAM_THREADS)
k_fill(long long n, T* __restrict__ x, T v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        x[i] = v;
}

// Reflection-sector table of K0 (72 entries after the 576 matrix entries):
// Kh[psi][c][c'] = (1/8) sum_x (-1)^popcount((psi ^ 2^c') & mask(x)) K0[c][3x+c'],
// mask(x) = the corner's offset bits (x, y, z).  One thread per entry.
__global__ void __launch_bounds__(128) k_sector_table(double* K0) {
    const int t = threadIdx.x;
    if (t >= 72) return;
    const int psi = t / 9, c = (t / 3) % 3, c2 = t % 3;
    double acc = 0.0;
#pragma unroll
    for (int x = 0; x < 8; ++x) {
        const int m = OX(x) | (OY(x) << 1) | (OZ(x) << 2);
        const double v = K0[c * 24 + 3 * x + c2];
        acc += (__builtin_popcount((psi ^ (1 << c2)) & m) & 1) ? -v : v;
    }
    K0[576 + t] = 0.125 * acc;
}

// Jacobi diagonal: K_ii = sum over the node's elements of s_e K0[3a+c][3a+c].
__global__ void __launch_bounds__(STREAM_THREADS)
k_diag(GridDev g, const double* __restrict__ s_e, const double* __restrict__ K0,
       const uint8_t* __restrict__ free_mask, double* __restrict__ dinv, int* __restrict__ flag) {
    for (long long nd = (long long)blockIdx.x * blockDim.x + threadIdx.x; nd < g.nnodes;
         nd += (long long)gridDim.x * blockDim.x) {
        const int j = int(nd % g.NJ);
        const int i = int((nd / g.NJ) % g.NI);
        const int k = int(nd / ((long long)g.NJ * g.NI));
        double dg[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < 8; ++a) {
            const int je = j - OY(a), ie = i - OX(a), ke = k - OZ(a);
            if (je < 0 || je >= g.ny || ie < 0 || ie >= g.nx || ke < 0 || ke >= g.nz) continue;
            const double s = s_e[(long long)ke * g.nx * g.ny + (long long)ie * g.ny + je];
            for (int c = 0; c < 3; ++c) dg[c] += s * K0[(3 * a + c) * 25];
        }
        for (int c = 0; c < 3; ++c) {
            const long long di = 3 * nd + c;
            const bool fr = free_mask[di] != 0;
            if (fr && dg[c] == 0.0) atomicOr(flag, 1);
            dinv[di] = (fr && dg[c] != 0.0) ? 1.0 / dg[c] : 0.0;
        }
    }
}

// s = rho_min + (1-rho_min) rho^p ; ds = p (1-rho_min) rho^(p-1)
__global__ void __launch_bounds__(STREAM_THREADS)
k_simp(long long m, const double* __restrict__ rho, double p, double rho_min,
       double* __restrict__ s, double* __restrict__ ds) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < m;
         e += (long long)gridDim.x * blockDim.x) {
        const double r = rho[e];
        s[e] = rho_min + (1.0 - rho_min) * pow(r, p);
        if (ds) ds[e] = p * (1.0 - rho_min) * pow(r, p - 1.0);
    }
}

// General-geometry element stiffness, one wavefront per element.  Lanes 0..7
// each take one Gauss point (xi outer, eta middle, zeta inner, weight 1): they
// form J = dN X, detJ and the physical gradients J^-1 dN into LDS; then every
// lane owns 9 of the 576 (i,j) entries of sum_gp B^T D B detJ w.
// BBAR: the B-bar (mean-dilatation) form (fe/elements/solids.py:225-272): at
// every Gauss point the volumetric part (m m^T / 3) B of B is replaced by its
// element volume average, i.e. column (a, c) of B gets (gbar_c(a) - g_c(a)) / 3
// added to its three normal-strain rows, g_c(a) = dN_a/dx_c at the point and
// gbar_c(a) = sum_q g_c(a)(q) detJ_q / sum_q detJ_q.
template <bool BBAR>
__global__ void __launch_bounds__(256)
k_ke(long long m, const double* __restrict__ X, const double* __restrict__ D, long long d_stride,
     double* __restrict__ ke) {
    __shared__ double sG[4][8][3][8];
    __shared__ double sDet[4][8];
    __shared__ double sGb[4][3][8];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long long e = (long long)blockIdx.x * 4 + w;
    const bool live = e < m;
    if (live && lane < 8) {
        const int q = lane;
        const double gp = 0.57735026918962576451;  // 1/sqrt(3)
        const double xi = (q & 4) ? gp : -gp;
        const double et = (q & 2) ? gp : -gp;
        const double ze = (q & 1) ? gp : -gp;
        double dN[3][8];
        for (int a = 0; a < 8; ++a) {
            const double xa = OX(a) ? 1.0 : -1.0, ya = OY(a) ? 1.0 : -1.0, za = OZ(a) ? 1.0 : -1.0;
            dN[0][a] = 0.125 * xa * (1 + et * ya) * (1 + ze * za);
            dN[1][a] = 0.125 * ya * (1 + xi * xa) * (1 + ze * za);
            dN[2][a] = 0.125 * za * (1 + xi * xa) * (1 + et * ya);
        }
        double J[3][3];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double acc = 0.0;
                for (int a = 0; a < 8; ++a) acc += dN[r][a] * X[e * 24 + 3 * a + c];
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
        for (int r = 0; r < 3; ++r)
            for (int a = 0; a < 8; ++a)
                sG[w][q][r][a] = Ji[r][0] * dN[0][a] + Ji[r][1] * dN[1][a] + Ji[r][2] * dN[2][a];
        sDet[w][q] = det;
    }
    __syncthreads();
    if constexpr (BBAR) {
        if (live && lane < 24) {
            const int r = lane / 8, a = lane % 8;
            double num = 0.0, vol = 0.0;
            for (int q = 0; q < 8; ++q) {
                num += sG[w][q][r][a] * sDet[w][q];
                vol += sDet[w][q];
            }
            sGb[w][r][a] = num / vol;
        }
        __syncthreads();
    }
    if (!live) return;
    const double* De = D + e * d_stride;
    double Dm[6][6];
#pragma unroll
    for (int s = 0; s < 6; ++s)
#pragma unroll
        for (int t = 0; t < 6; ++t) Dm[s][t] = De[6 * s + t];
#pragma unroll 1
    for (int t = 0; t < 9; ++t) {
        const int idx = lane + 64 * t;
        const int I = idx / 24, Jc = idx % 24;
        const int a = I / 3, c = I % 3, b = Jc / 3, d = Jc % 3;
        double acc = 0.0;
#pragma unroll 1
        for (int q = 0; q < 8; ++q) {
            const double ga0 = sG[w][q][0][a], ga1 = sG[w][q][1][a], ga2 = sG[w][q][2][a];
            const double gb0 = sG[w][q][0][b], gb1 = sG[w][q][1][b], gb2 = sG[w][q][2][b];
            // column 3a+c of B over rows (xx, yy, zz, xy, yz, zx)
            double Bi[6] = {0, 0, 0, 0, 0, 0}, Bj[6] = {0, 0, 0, 0, 0, 0};
            if (c == 0) { Bi[0] = ga0; Bi[3] = ga1; Bi[5] = ga2; }
            else if (c == 1) { Bi[1] = ga1; Bi[3] = ga0; Bi[4] = ga2; }
            else { Bi[2] = ga2; Bi[4] = ga1; Bi[5] = ga0; }
            if (d == 0) { Bj[0] = gb0; Bj[3] = gb1; Bj[5] = gb2; }
            else if (d == 1) { Bj[1] = gb1; Bj[3] = gb0; Bj[4] = gb2; }
            else { Bj[2] = gb2; Bj[4] = gb1; Bj[5] = gb0; }
            if constexpr (BBAR) {
                const double vi = (sGb[w][c][a] - (c == 0 ? ga0 : (c == 1 ? ga1 : ga2))) / 3.0;
                const double vj = (sGb[w][d][b] - (d == 0 ? gb0 : (d == 1 ? gb1 : gb2))) / 3.0;
#pragma unroll
                for (int s = 0; s < 3; ++s) {
                    Bi[s] += vi;
                    Bj[s] += vj;
                }
            }
            double sum = 0.0;
#pragma unroll
            for (int s = 0; s < 6; ++s) {
                double dbj = 0.0;
#pragma unroll
                for (int tt = 0; tt < 6; ++tt) dbj += Dm[s][tt] * Bj[tt];
                sum += Bi[s] * dbj;
            }
            acc += sum * sDet[w][q];
        }
        ke[e * 576 + idx] = acc;
    }
}

}  // namespace h8
