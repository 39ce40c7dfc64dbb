// h8_topopt.h -- the SIMP loop around the compliance gradient: Sigmund's
// sensitivity filter on the structured grid and the optimality-criteria (OC)
// density update with its bisection on the volume multiplier.
#pragma once
#include "h8_device.h"
#include "h8_stream.h"

namespace h8 {

// Sensitivity filter (fe/topopt.py:235-252): with the linear-decay weights
// H_ej = max(0, r - |c_e - c_j|) between element centroids,
//     out_e = sum_j H_ej rho_j dc_j / max(max(rho_e, eps) sum_j H_ej, eps).
// On the grid H is a stencil: offsets (di, dj, dk) with |di| <= Ri etc.
// (R = ceil(r / h)), weight r - sqrt((di hx)^2 + (dj hy)^2 + (dk hz)^2) where
// positive, restricted to elements inside the grid.  Each block first
// compacts the offsets of positive weight into LDS (wave 0, 64 offsets per
// ballot; box order dk, di, dj), so every element visits only those; boxes
// larger than FILT_MAX offsets are walked directly.  One thread per element,
// offsets in the same fixed order either way (deterministic, identical sums).
constexpr int FILT_MAX = 1024;

__global__ void __launch_bounds__(256) k_sens_filter(GridDev g, double hx, double hy, double hz, double r, int Ri,
                                                     int Rj, int Rk, const double* rho, const double* dc, double eps,
                                                     double* out) {
    __shared__ double sW[FILT_MAX];
    __shared__ int sD[FILT_MAX];          // offset packed as ((dk + 64) << 16) | ((di + 64) << 8) | (dj + 64)
    __shared__ int sN;
    const int bj = 2 * Rj + 1, bi = 2 * Ri + 1;
    const int box = bj * bi * (2 * Rk + 1);
    const bool table = box <= FILT_MAX;
    if (table) {
        if (threadIdx.x < 64) {
            const int lane = threadIdx.x;
            int base = 0;
            for (int t0 = 0; t0 < box; t0 += 64) {
                const int t = t0 + lane;
                const int dj = t % bj - Rj, di = (t / bj) % bi - Ri, dk = t / (bj * bi) - Rk;
                const double x = di * hx, y = dj * hy, z = dk * hz;
                const double w = r - sqrt(x * x + y * y + z * z);
                const bool keep = t < box && w > 0.0;
                const unsigned long long mask = __ballot(keep);
                const int pos = base + __popcll(mask & ((1ull << lane) - 1ull));
                if (keep) {
                    sW[pos] = w;
                    sD[pos] = ((dk + 64) << 16) | ((di + 64) << 8) | (dj + 64);
                }
                base += __popcll(mask);
            }
            if (lane == 0) sN = base;
        }
        __syncthreads();
    }
    const long long m = (long long)g.nx * g.ny * g.nz;
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    const int je = int(e % g.ny), ie = int((e / g.ny) % g.nx), ke = int(e / ((long long)g.ny * g.nx));
    double num = 0.0, hs = 0.0;
    if (table) {
        const int n = sN;
        for (int q = 0; q < n; ++q) {
            const int pk = sD[q];
            const int k = ke + (pk >> 16) - 64, i = ie + ((pk >> 8) & 255) - 64, j = je + (pk & 255) - 64;
            if (k < 0 || k >= g.nz || i < 0 || i >= g.nx || j < 0 || j >= g.ny) continue;
            const double w = sW[q];
            const long long f = ((long long)k * g.nx + i) * g.ny + j;
            num += w * (rho[f] * dc[f]);
            hs += w;
        }
    } else {
        for (int dk = -Rk; dk <= Rk; ++dk) {
            const int k = ke + dk;
            if (k < 0 || k >= g.nz) continue;
            for (int di = -Ri; di <= Ri; ++di) {
                const int i = ie + di;
                if (i < 0 || i >= g.nx) continue;
                for (int dj = -Rj; dj <= Rj; ++dj) {
                    const int j = je + dj;
                    if (j < 0 || j >= g.ny) continue;
                    const double x = di * hx, y = dj * hy, z = dk * hz;
                    const double w = r - sqrt(x * x + y * y + z * z);
                    if (!(w > 0.0)) continue;
                    const long long f = ((long long)k * g.nx + i) * g.ny + j;
                    num += w * (rho[f] * dc[f]);
                    hs += w;
                }
            }
        }
    }
    out[e] = num / fmax(fmax(rho[e], eps) * hs, eps);
}

// Bisection state of the OC update (device memory).
struct OcState {
    double l1, l2;          // multiplier bracket
    int phase;              // 0: widening the bracket, 1: bisecting, 2: done
    int k_exp, k_bis;       // passes taken in each phase
    unsigned cnt;           // ticket counter of the pass reduction
    double mean;            // mean density of the last pass
};
struct OcParams {
    double vf, move, lo, hi, tol;
    int max_bis;
};
constexpr int OC_MAX_EXPAND = 60;    // bracket doublings (fe/topopt.py:280)

// One pass of the OC update (fe/topopt.py:258-296).  With the trial
// multiplier lam (the bracket's upper end while widening, its midpoint while
// bisecting), rho_new_e = clip(clip(rho_e sqrt(-min(dc_e, -1e-12) / lam),
// rho_e - move, rho_e + move), lo, hi); the pass sums rho_new and the last
// block applies the reference's bracket rule to the mean.  Bisecting passes
// write rho_new, so the last one leaves the result.  A pass after the state
// reached `done` returns at once.
__global__ void __launch_bounds__(STREAM_THREADS) k_oc_pass(long long m, const double* rho, const double* dc,
                                                           OcParams P, OcState* st, double* partials,
                                                           double* out) {
    __shared__ double sRed[16];
    const int phase = st->phase;
    if (phase == 2) return;
    const double lam = phase == 0 ? st->l2 : 0.5 * (st->l1 + st->l2);
    double v[1] = {0.0};
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += stride) {
        const double re = rho[e];
        const double d = fmin(dc[e], -1e-12);
        double x = re * sqrt(-d / lam);
        x = fmin(fmax(x, re - P.move), re + P.move);
        x = fmin(fmax(x, P.lo), P.hi);
        v[0] += x;
        if (phase == 1) out[e] = x;
    }
    double tot[1];
    if (reduce_last_block<1>(v, partials, &st->cnt, tot, sRed) && threadIdx.x == 0) {
        const double mean = tot[0] / double(m);
        st->mean = mean;
        if (phase == 0) {
            if (mean <= P.vf) {
                st->phase = 1;
            } else {
                st->l2 = 2.0 * st->l2;
                if (++st->k_exp >= OC_MAX_EXPAND) st->phase = 1;
            }
        } else {
            if (mean > P.vf) st->l1 = lam;
            else st->l2 = lam;
            ++st->k_bis;
            if (st->l2 - st->l1 < P.tol * fmax(st->l2, 1.0) || st->k_bis >= P.max_bis) st->phase = 2;
        }
    }
}

// max_e |a_e - b_e| into *out (the bits of a non-negative double order like
// the value, so an unsigned atomic max is exact and order-independent);
// *out must be 0 before the launch.
__global__ void __launch_bounds__(STREAM_THREADS) k_absdiff_max(long long n, const double* a, const double* b,
                                                               unsigned long long* out) {
    __shared__ double sm[STREAM_THREADS / 64];
    double v = 0.0;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        v = fmax(v, fabs(a[i] - b[i]));
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < int(blockDim.x >> 6); ++w) v = fmax(v, sm[w]);
        atomicMax(out, __builtin_bit_cast(unsigned long long, v));
    }
}

}  // namespace h8
