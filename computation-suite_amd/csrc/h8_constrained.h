// h8_constrained.h -- Jacobi-PCG with linear multi-point constraints eliminated
// (fe/solver.py:1341-1479: add_constraint / tie / form_linear_system).
//
// The host reduces the constraints to slave DOFs u_s = g_s + sum_m T_sm u_m
// over kept (master) DOFs, as fe/boundary.py:98-166 does.  The eliminated
// system is T^T K T x = T^T (F - K w), with w = the prescribed values at fixed
// DOFs and g at slaves, and u = T x + w.  On full-length vectors that are zero
// at fixed and slave DOFs (the masked form of the other solves), one product
// T^T K T p is: expand p to the slaves (p_s = sum T_sm p_m), q = K p with the
// grid kernels, contract q onto the masters (q_m += sum_s T_sm q_s) and zero
// the slave entries of p again; the slave entries of q are left as they are
// (the update kernel masks them, and p is zero there in p.q).  The recurrence
// and stopping rule are those of h8_pcg_f64 (fe/iterative_solvers.py:434-469)
// with the Jacobi diagonal of T^T K T, which at a master m is
//     K_mm + 2 sum_s T_sm K_sm + sum_s sum_s' T_sm T_s'm K_ss'
// with the K entries evaluated from K0 and the element scales.
#pragma once
#include "h8_device.h"
#include "h8_node.h"
#include "h8_stream.h"

namespace h8 {

// The reduced constraint map in device memory (CSR both ways).
struct CMapDev {
    long long n_slave, n_master;
    const int64_t* slave;       // [n_slave] slave DOFs
    const int32_t* s_ptr;       // [n_slave + 1] into s_master / s_coef
    const int64_t* s_master;    // masters of each slave
    const double* s_coef;       // T_sm
    const int64_t* master;      // [n_master] DOFs that some slave depends on
    const int32_t* m_ptr;       // [n_master + 1] into m_slave / m_coef
    const int64_t* m_slave;     // slaves of each master
    const double* m_coef;       // T_sm
};

// K_ij of the assembled grid operator (sum over the elements holding both
// nodes of s_e K0[3a + c][3b + d], a and b the nodes' corners in the element).
__device__ inline double k_entry(const GridDev& g, const double* s_e, const double* K0, long long i, long long j) {
    const long long na = i / 3, nb = j / 3;
    const int c = int(i - 3 * na), d = int(j - 3 * nb);
    const long long pl = (long long)g.NI * g.NJ;
    const int ka = int(na / pl), ia = int((na % pl) / g.NJ), ja = int(na % g.NJ);
    const int kb = int(nb / pl), ib = int((nb % pl) / g.NJ), jb = int(nb % g.NJ);
    if (abs(ia - ib) > 1 || abs(ja - jb) > 1 || abs(ka - kb) > 1) return 0.0;
    double acc = 0.0;
    for (int ke = max(max(ka, kb) - 1, 0); ke <= min(min(ka, kb), g.nz - 1); ++ke)
        for (int ie = max(max(ia, ib) - 1, 0); ie <= min(min(ia, ib), g.nx - 1); ++ie)
            for (int je = max(max(ja, jb) - 1, 0); je <= min(min(ja, jb), g.ny - 1); ++je) {
                const int la = corner_of(ia - ie, ja - je, ka - ke), lb = corner_of(ib - ie, jb - je, kb - ke);
                acc += s_e[((long long)ke * g.nx + ie) * g.ny + je] * K0[(3 * la + c) * 24 + 3 * lb + d];
            }
    return acc;
}

// dinv at the masters: 1 / diag(T^T K T) (0 if not positive: the host checks).
__global__ void __launch_bounds__(256) k_cdiag(GridDev g, const double* s_e, const double* K0, CMapDev c,
                                               double* dinv) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= c.n_master) return;
    const long long m = c.master[t];
    double a = k_entry(g, s_e, K0, m, m);
    for (int p = c.m_ptr[t]; p < c.m_ptr[t + 1]; ++p) {
        const long long s = c.m_slave[p];
        const double ts = c.m_coef[p];
        a += 2.0 * ts * k_entry(g, s_e, K0, s, m);
        for (int q = c.m_ptr[t]; q < c.m_ptr[t + 1]; ++q) a += ts * c.m_coef[q] * k_entry(g, s_e, K0, s, c.m_slave[q]);
    }
    dinv[m] = a > 0.0 ? 1.0 / a : 0.0;
}

// p = dinv r (first direction) or dinv r + beta p.
__global__ void __launch_bounds__(STREAM_THREADS) k_cdir(long long n, const double* r, const double* dinv, double* p,
                                                        const PcgState* st) {
    if (st->status != ST_RUNNING) return;
    const bool first = st->first != 0;
    const double beta = st->beta;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double z = dinv[i] * r[i];
        p[i] = first ? z : fma(beta, p[i], z);
    }
}

// v_s = sum_m T_sm v_m at every slave (v zero at the slaves before; the
// masters are never slaves, so in place is safe).  st NULL: always runs.
__global__ void __launch_bounds__(256) k_cexpand(CMapDev c, double* v, const PcgState* st) {
    if (st && st->status != ST_RUNNING) return;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= c.n_slave) return;
    double acc = 0.0;
    for (int k = c.s_ptr[t]; k < c.s_ptr[t + 1]; ++k) acc += c.s_coef[k] * v[c.s_master[k]];
    v[c.slave[t]] = acc;
}

// Threads [0, n_master): q_m += sum_s T_sm q_s (reads slave entries, writes
// master entries).  Threads [n_master, n_master + n_slave): p_s = 0 (when p).
__global__ void __launch_bounds__(256) k_cfix(CMapDev c, double* p, double* q, const PcgState* st) {
    if (st && st->status != ST_RUNNING) return;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < c.n_master) {
        double acc = q[c.master[t]];
        for (int k = c.m_ptr[t]; k < c.m_ptr[t + 1]; ++k) acc += c.m_coef[k] * q[c.m_slave[k]];
        q[c.master[t]] = acc;
    } else if (p && t < c.n_master + c.n_slave) {
        p[c.slave[t - c.n_master]] = 0.0;
    }
}

// p.q over every DOF (p is zero at fixed and slave DOFs) -> alpha (dir_finish).
__global__ void __launch_bounds__(STREAM_THREADS) k_cpq(long long n, const double* p, const double* q, PcgState* st,
                                                       double* partials) {
    __shared__ double sRed[16];
    if (st->status != ST_RUNNING) return;
    double v[1] = {0.0};
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) v[0] += p[i] * q[i];
    double tot[1];
    if (reduce_last_block<1>(v, partials, &st->cnt[0], tot, sRed) && threadIdx.x == 0) dir_finish(st, tot[0]);
}

}  // namespace h8
