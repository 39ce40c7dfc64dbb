// h8_device.h -- device-side building blocks shared by the h8pcg kernels (gfx950).
//
// Structured H8 grid (box mesh): node(i,j,k) = k*NI*NJ + i*NJ + j with
// NJ = ny+1 (y fastest), NI = nx+1, NK = nz+1; element e(i,j,k) = k*nx*ny +
// i*ny + j; element corner a sits at node offset (OX[a], OY[a], OZ[a]) from the
// element origin, DOF = 3*node + c.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace h8 {

// Corner offsets of the Hex8 natural ordering (bottom face CCW, then top).
__host__ __device__ constexpr int OX(int a) { return (a == 1 || a == 2 || a == 5 || a == 6) ? 1 : 0; }
__host__ __device__ constexpr int OY(int a) { return (a == 2 || a == 3 || a == 6 || a == 7) ? 1 : 0; }
__host__ __device__ constexpr int OZ(int a) { return a >= 4 ? 1 : 0; }

struct GridDev {
    int nx, ny, nz;        // elements
    int NJ, NI, NK;        // nodes per axis
    long long nnodes;      // NJ*NI*NK
    long long ndof;        // 3*nnodes
    int tj, ti, tk;        // tiles per axis (node tiles)
    int ntiles;            // tj*ti*tk
    int variant;           // H8_VARIANT_*
    int layers;            // node layers per block of the sector kernel (0 = choose)
    int arith;             // element-product arithmetic of FP32 sector/resident kernels (0 = 32, 32, 64)
    int mg_degree;         // Chebyshev degree of the multigrid level-0 smoother (0 = library choice)
};

inline void set_tiles(GridDev& d, int TJ, int TI, int TK) {
    d.tj = (d.NJ + TJ - 1) / TJ;
    d.ti = (d.NI + TI - 1) / TI;
    d.tk = (d.NK + TK - 1) / TK;
    d.ntiles = d.tj * d.ti * d.tk;
}

// Solver status held in device memory.
enum : int { ST_RUNNING = 0, ST_DONE = 1, ST_NOT_SPD = 2, ST_MAXITER = 3, ST_REPLACE = 4 };

struct PcgState {
    double rz_old;     // r.z of the previous iteration
    double alpha, beta;
    double pq;         // p.Ap of the current iteration
    double bnorm;      // ||b|| (1 if 0) -- for the refined solve, ||f||
    double tol;
    double rel;        // recurrence ||r||/bnorm
    double rmax;       // refined: ||r64|| at the last replacement
    double delta;      // refined: replacement threshold
    double true_rel;   // refined: FP64 ||f-Ku||/||f|| at the last replacement
    double bad_pq;     // p.Ap when status = NOT_SPD
    long long k;       // current iteration (1-based)
    long long maxiter;
    long long iters;   // iteration count reported at exit
    long long outer;   // refined: replacements done
    int status;
    int first;         // next direction kernel uses p = z
    unsigned cnt[4];   // last-block counters (reset by the last block)
    double best_true;  // refined: smallest true relative residual so far
    long long stall;   // refined: replacements since it last fell by 1%
    int stalled;       // refined: stopped because the true residual stagnated
    long long k_rep;   // refined: iteration index at the last replacement
    long long max_phase;  // refined: longest FP32 phase (iterations between replacements)
    int pending;       // x still lacks alpha * p of the last direction kernel: the next
                       // direction kernel (or the final/fold kernel) applies it
    int mg;            // refined solve preconditioned by the multigrid V-cycle: beta and
                       // rz_old come from r.z of the cycle's z (k_mg_rz), not from dinv r
    int replaced;      // multigrid: the residual was replaced since the last r.z
    int fold;          // multigrid: the FP64 residual was formed at u + x32 and u has not
                       // taken x32 yet -- the next update kernel folds it (u += x32)
};

// Refined solve: replacements without a 1% drop of the true residual after
// which the solve stops (the FP32 iterations cannot reduce it further).
constexpr long long REFINE_STALL = 16;
// An FP32 phase longer than max(REFINE_PHASE, 4 x the longest phase so far) is
// ended by a replacement: at the FP32 rounding floor the FP32 iterations can
// drift away instead of reaching the next replacement threshold.
constexpr long long REFINE_PHASE = 500;

// ---------------------------------------------------------------------------
// Reductions: wave shuffle tree + ordered LDS sum, then "last block" finish.
// Both stages use a fixed order, so results are bitwise reproducible.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* sh /* >= 16*NV */) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) sh[wid * NV + i] = v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double s = 0.0;
        for (int w = 0; w < nw; ++w) s += sh[w * NV + i];
        v[i] = s;
    }
    __syncthreads();
}

__device__ __forceinline__ void st_relaxed(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_relaxed(const double* p) {
    return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Every block contributes v[NV]; returns true in the last block to finish, in
// which `tot` holds the sum over blocks (ordered by block index).
//
// Cross-block visibility without L2 maintenance: the partials are written and
// read with device-scope (sc1) accesses, which bypass the per-XCD L2, and the
// writer waits for its stores to complete (workgroup-scope release = wait for
// the outstanding stores, no cache write-back) before taking a ticket with a
// device-scope atomic.  A device-scope __threadfence() here would write back
// and invalidate the whole L2 of the XCD in every block.
template <int NV>
__device__ bool reduce_last_block(double (&v)[NV], double* partials, unsigned* counter,
                                  double (&tot)[NV], double* sh) {
    __shared__ int s_last;
    block_sum<NV>(v, sh);
    const unsigned nb = gridDim.x;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) st_relaxed(&partials[(size_t)blockIdx.x * NV + i], v[i]);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0);
        const unsigned t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (t == nb - 1);
    }
    __syncthreads();
    if (!s_last) return false;
    // issue the device-scope loads in groups of 8 per thread before summing,
    // so their latencies overlap (fixed order: b = t, t+nt, ...)
    double acc[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = 0.0;
    const unsigned nt = blockDim.x;
    for (unsigned b0 = threadIdx.x; b0 < nb; b0 += 8 * nt) {
        double buf[8][NV];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const unsigned b = b0 + u * nt;
#pragma unroll
            for (int i = 0; i < NV; ++i) buf[u][i] = b < nb ? ld_relaxed(&partials[(size_t)b * NV + i]) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int i = 0; i < NV; ++i) acc[i] += buf[u][i];
    }
    block_sum<NV>(acc, sh);
#pragma unroll
    for (int i = 0; i < NV; ++i) tot[i] = acc[i];
    if (threadIdx.x == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// Read-only, wave-uniform tables (the 24x24 K0): accessed through the constant
// address space so the compiler emits scalar (s_load) reads shared by the wave
// instead of one vector load per lane.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* as_const(const T* p) {
    return (const __attribute__((address_space(4))) T*)p;
}

__device__ __forceinline__ void tile_origin(const GridDev& g, int tile, int TJ, int TI, int TK,
                                            int& j0, int& i0, int& k0) {
    const int bj = tile % g.tj;
    const int rest = tile / g.tj;
    const int bi = rest % g.ti;
    const int bk = rest / g.ti;
    j0 = bj * TJ; i0 = bi * TI; k0 = bk * TK;
}

// Brick of a block in a one-brick-per-block launch.  Workgroups are dealt to
// the 8 XCDs round-robin (block b runs on XCD b % 8), so consecutive block ids
// would put neighbouring bricks on different XCDs and every halo would be
// fetched into several L2s.  This bijection gives XCD x the contiguous brick
// range [x*q + min(x, r), ...) (q = n / 8, r = n % 8), which has exactly as
// many bricks as XCD x has blocks.
__device__ __forceinline__ int xcd_local_tile(int b, int n) {
    const int x = b & 7, q = n >> 3, r = n & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// Bounds-checked raw buffer access.  A buffer resource covers `bytes` bytes
// from `p`; a load at an offset >= bytes returns 0 and a store there is
// dropped (the hardware range check of raw buffers), so out-of-domain halo
// nodes need no branch: their offset is set to OOB.
constexpr int OOB = int(0x80000000u);
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, int(bytes), 0x00020000);
}
template <typename T> struct Buf;
template <> struct Buf<float> {
    static __device__ __forceinline__ float ld(__amdgpu_buffer_rsrc_t r, int off) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
    }
    static __device__ __forceinline__ void st(float v, __amdgpu_buffer_rsrc_t r, int off) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
    }
};
template <> struct Buf<double> {
    static __device__ __forceinline__ double ld(__amdgpu_buffer_rsrc_t r, int off) {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
    }
    static __device__ __forceinline__ void st(double v, __amdgpu_buffer_rsrc_t r, int off) {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, off, 0, 0);
    }
};

}  // namespace h8
