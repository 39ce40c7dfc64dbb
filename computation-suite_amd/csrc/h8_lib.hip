// h8_lib.hip -- C ABI of libh8pcg.so: launches, workspace carving, the PCG
// driver (hipGraph-captured iteration chunks, device-side convergence flag,
// one host poll per chunk) and the mixed-precision refined solve.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/h8pcg.h"
#include "h8_device.h"
#include "h8_node.h"
#include "h8_stream.h"
#include "h8_mfma.h"
#include "h8_hex20.h"
#include "h8_batched.h"
#include "h8_sector.h"
#include "h8_resident.h"
#include "h8_recovery.h"
#include "h8_topopt.h"
#include "h8_constrained.h"
#include "h8_mg.h"

using namespace h8;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;
static thread_local double g_batched_ms = 0.0;
static thread_local long long g_batched_launches = 0;

static int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(H8_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                           \
    } while (0)

#define LAUNCHCHK()                                                                    \
    do {                                                                               \
        hipError_t e_ = hipGetLastError();                                             \
        if (e_ != hipSuccess)                                                          \
            return fail(H8_ERR_HIP, "kernel launch failed: %s (%s:%d)",                \
                        hipGetErrorString(e_), __FILE__, __LINE__);                    \
    } while (0)

// A HIP event destroyed with its scope, so every early return (HIPCHK,
// argument errors) releases it.
struct Ev {
    hipEvent_t e = nullptr;
    Ev() = default;
    Ev(const Ev&) = delete;
    Ev& operator=(const Ev&) = delete;
    ~Ev() {
        if (e) (void)hipEventDestroy(e);
    }
    hipError_t create(unsigned flags = hipEventDefault) { return hipEventCreateWithFlags(&e, flags); }
    operator hipEvent_t() const { return e; }
};

extern "C" const char* h8_last_error(void) { return g_err.c_str(); }
extern "C" int h8_version(void) { return 1; }

extern "C" int h8_device_info(char* buf, int buflen) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, dev));
    snprintf(buf, buflen, "%s|%s|cu=%d|l2=%d|clock_khz=%d", p.name, p.gcnArchName,
             p.multiProcessorCount, p.l2CacheSize, p.clockRate);
    return H8_OK;
}

extern "C" int h8_sector_table_f64(double* K0, void* stream) {
    if (!K0) return fail(H8_ERR_ARG, "NULL pointer");
    hipLaunchKernelGGL(k_sector_table, dim3(1), dim3(128), 0, (hipStream_t)stream, K0);
    LAUNCHCHK();
    return H8_OK;
}

// ---------------------------------------------------------------------------
// tiling
// ---------------------------------------------------------------------------
// Node brick owned by one block of the node-centric kernel: TJ (y, unit
// stride) x TI (x) x TK (z) = 256 threads.
constexpr int NTJ = 8, NTI = 8, NTK = 4;
// Node brick of the MFMA kernel (256 threads, larger bricks amortise the
// one-element halo the block recomputes): FP32 16x8x8, FP64 8x8x8 (LDS).
template <typename T> struct MfmaTile;
template <> struct MfmaTile<float> { static constexpr int J = 16, I = 8, K = 8; };
template <> struct MfmaTile<double> { static constexpr int J = 8, I = 8, K = 8; };

// Library default per precision and size (measured on MI355X): the sector
// kernel in FP32; in FP64 the node-centric kernel below ~1M DOF (the sector
// kernel's heavier FP64 register state costs more than its arithmetic saves
// there) and the sector kernel above.
template <typename T> static int default_variant(const GridDev& d);
template <> int default_variant<float>(const GridDev&) { return H8_VARIANT_SECTOR; }
template <> int default_variant<double>(const GridDev& d) {
    return d.ndof >= 1000000 ? H8_VARIANT_SECTOR : H8_VARIANT_NODE;
}

static int make_grid(const h8_grid* g, GridDev& d) {
    if (!g) return fail(H8_ERR_ARG, "grid is NULL");
    if (g->nx < 1 || g->ny < 1 || g->nz < 1)
        return fail(H8_ERR_ARG, "grid needs nx, ny, nz >= 1 (got %d, %d, %d)", g->nx, g->ny, g->nz);
    d.nx = g->nx; d.ny = g->ny; d.nz = g->nz;
    d.NJ = g->ny + 1; d.NI = g->nx + 1; d.NK = g->nz + 1;
    d.nnodes = (long long)d.NJ * d.NI * d.NK;
    d.ndof = 3 * d.nnodes;
    d.tj = (d.NJ + NTJ - 1) / NTJ;
    d.ti = (d.NI + NTI - 1) / NTI;
    d.tk = (d.NK + NTK - 1) / NTK;
    long long nt = (long long)d.tj * d.ti * d.tk;
    // byte offsets of FP64 vectors must fit the 31-bit range of buffer offsets
    if (nt > (1LL << 30) || d.ndof >= (1LL << 28))
        return fail(H8_ERR_ARG, "grid too large: %lld DOFs (the kernels address vectors with 32-bit byte offsets; "
                    "at most 2^28 - 1 DOFs)", d.ndof);
    d.ntiles = int(nt);
    if (g->variant < H8_VARIANT_AUTO || g->variant > H8_VARIANT_RESIDENT)
        return fail(H8_ERR_ARG, "unknown kernel variant %d", g->variant);
    if (g->layers < 0) return fail(H8_ERR_ARG, "layers must be >= 0 (0 = library choice), got %d", g->layers);
    if (g->arith != 0 && g->arith != 32 && g->arith != 64)
        return fail(H8_ERR_ARG, "arith must be 0 (library choice), 32 or 64, got %d", g->arith);
    if (g->mg_degree != 0 && (g->mg_degree < 2 || g->mg_degree > MG_CHEB_MAXDEG))
        return fail(H8_ERR_ARG, "mg_degree must be 0 (library choice) or 2..%d, got %d", MG_CHEB_MAXDEG,
                    g->mg_degree);
    d.variant = g->variant;
    d.layers = g->layers;
    d.arith = g->arith;
    d.mg_degree = g->mg_degree;
    return H8_OK;
}

// Persistent grid for a tile kernel: (resident blocks per CU) x CUs, capped at
// the tile count and rounded down to a multiple of the XCD count so every XCD
// gets the same number of blocks.
static std::mutex g_occ_mu;
static std::vector<std::pair<const void*, int>> g_occ;
static int g_cus[64] = {};

// Compute units of the current device (256 if the query fails).
static int device_cus() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    std::lock_guard<std::mutex> lk(g_occ_mu);
    if (!g_cus[dev]) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, dev) != hipSuccess) return 256;
        g_cus[dev] = p.multiProcessorCount;
    }
    return g_cus[dev];
}

static int persistent_grid(const void* kernel, int threads, int ntiles) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return ntiles;
    std::lock_guard<std::mutex> lk(g_occ_mu);
    if (!g_cus[dev]) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, dev) != hipSuccess) return ntiles;
        g_cus[dev] = p.multiProcessorCount;
    }
    int per = 0;
    for (auto& e : g_occ)
        if (e.first == kernel) per = e.second;
    if (!per) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess || per < 1) per = 1;
        g_occ.push_back({kernel, per});
    }
    long long gsz = std::min<long long>((long long)g_cus[dev] * per, ntiles);
    if (gsz >= NXCD) gsz -= gsz % NXCD;
    return int(std::max(1LL, gsz));
}

// FP32 data with FP64 element products (arith 64).  The library's choice is
// FP32 products: the FP32 iterations need 1.6-2x the FP64 count on
// ill-conditioned systems, but that comes from rounding the recurrence
// vectors (p, r, q, z) to FP32 -- any one of them -- not from the products;
// FP64 products leave the count unchanged and cost registers.
template <typename T> static bool fp64_products(const GridDev& d) {
    return sizeof(T) == 4 && d.arith == 64;
}
// The FP64 sector table of a K0 buffer: at word 576 of an FP64 buffer, at
// word 648 (bit-copied doubles, 8-byte aligned) of an FP32 one.
static const double* k0_table64(const double* K0) { return K0 + 576; }
static const double* k0_table64(const float* K0) { return reinterpret_cast<const double*>(K0 + H8_K0_WORDS); }

// Sector kernel: column tiles of (SEJ-1) x (SEI-1) nodes, chunks of `layers`
// node layers (library choice: about 4 blocks per CU, 2..16 layers).  The
// partials buffer of the fused modes holds max(brick tiles, 2048) blocks, so
// the launch is refused if it would need more.
template <typename T, int MODE, int EJ, int EI>
static int launch_sector(const GridDev& d0, NodeArgs<T> A, hipStream_t s) {
    GridDev d = d0;
    sector_tiles(d, EJ, EI, 1);
    int kd = d.layers;
    if (kd <= 0) {
        const long long cols = (long long)d.tj * d.ti;
        const long long want = 4LL * device_cus();
        kd = int(std::max(2LL, std::min(16LL, (d.NK * cols + want - 1) / want)));
    }
    sector_tiles(d, EJ, EI, kd);
    const long long nt = (long long)d.tj * d.ti * d.tk;
    if (MODE != MODE_PLAIN && nt > std::max<long long>(d0.ntiles, 2048))
        return fail(H8_ERR_ARG, "sector kernel: %lld blocks exceed the partials capacity (use more layers)", nt);
    d.ntiles = int(nt);
    A.g = d;
    if (fp64_products<T>(d)) {
        hipLaunchKernelGGL((k_sector<T, EJ, EI, MODE, double>), dim3(d.ntiles), dim3(EJ * EI), 0, s, A,
                           k0_table64(A.K0), kd);
    } else {
        hipLaunchKernelGGL((k_sector<T, EJ, EI, MODE>), dim3(d.ntiles), dim3(EJ * EI), 0, s, A, A.K0 + 576, kd);
    }
    LAUNCHCHK();
    return H8_OK;
}

// Launch one structured operator kernel (mode MODE) with the variant recorded
// in d (AUTO -> the precision's default).
template <typename T, int MODE>
static int launch_op(const GridDev& d0, NodeArgs<T> A, hipStream_t s) {
    GridDev d = d0;
    // the resident solver's products outside its loop use the streaming default
    const int v = (d.variant == H8_VARIANT_AUTO || d.variant == H8_VARIANT_RESIDENT) ? default_variant<T>(d)
                                                                                      : d.variant;
    if constexpr (MODE == MODE_DIRX || MODE == MODE_DIRZ || MODE == MODE_CHEB || MODE == MODE_CHEB0 ||
                  MODE == MODE_MGRES) {   // sector kernels only
        if (v == H8_VARIANT_SECTOR) return launch_sector<T, MODE, 16, 16>(d0, A, s);
        if (v == H8_VARIANT_SECTOR_NARROW) return launch_sector<T, MODE, 8, 16>(d0, A, s);
        return fail(H8_ERR_ARG, "direction mode %d needs a sector kernel (variant %d)", MODE, v);
    } else if (v == H8_VARIANT_MFMA) {
        constexpr int TJ = MfmaTile<T>::J, TI = MfmaTile<T>::I, TK = MfmaTile<T>::K;
        set_tiles(d, TJ, TI, TK);
        A.g = d;
        hipLaunchKernelGGL((k_mfma<T, TJ, TI, TK, MODE>), dim3(d.ntiles), dim3(256), 0, s, A);
    } else {
        set_tiles(d, NTJ, NTI, NTK);
        A.g = d;
        if (v == H8_VARIANT_SECTOR) {
            return launch_sector<T, MODE, 16, 16>(d0, A, s);    // 16 x 16 element columns
        } else if (v == H8_VARIANT_SECTOR_NARROW) {
            return launch_sector<T, MODE, 8, 16>(d0, A, s);     // 8 x 16: half the LDS per block
        } else if (v == H8_VARIANT_NODE_SPLIT) {
            hipLaunchKernelGGL((k_node2<T, NTJ, NTI, NTK, MODE>), dim3(d.ntiles), dim3(2 * NTJ * NTI * NTK), 0, s, A);
        } else if (v == H8_VARIANT_NODE_PERSIST) {
            const void* kfn = (const void*)k_node_persist<T, NTJ, NTI, NTK, MODE>;
            const int grid = persistent_grid(kfn, NTJ * NTI * NTK, d.ntiles);
            hipLaunchKernelGGL((k_node_persist<T, NTJ, NTI, NTK, MODE>), dim3(grid), dim3(NTJ * NTI * NTK), 0, s,
                               A);
        } else {
            hipLaunchKernelGGL((k_node<T, NTJ, NTI, NTK, MODE>), dim3(d.ntiles), dim3(NTJ * NTI * NTK), 0, s, A);
        }
    }
    LAUNCHCHK();
    return H8_OK;
}

// Streaming grid: ~8 elements per thread (two 16-byte groups in FP32, four in
// FP64), at most 2048 blocks (the partials array holds one slot per block).
static int stream_blocks(long long n, int per_thread = 8) {
    long long b = (n + STREAM_THREADS * per_thread - 1) / (STREAM_THREADS * per_thread);
    // keep every CU busy: at least 2 blocks per CU when there is work for them
    b = std::max(b, std::min(512LL, (n + STREAM_THREADS - 1) / STREAM_THREADS));
    return int(std::max(1LL, std::min(b, 2048LL)));
}

static size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

// ---------------------------------------------------------------------------
// element stiffness
// ---------------------------------------------------------------------------
extern "C" int h8_ke_batched_f64(const double* X, const double* D, int64_t d_stride, int64_t m,
                                 double* ke, void* stream) {
    if (m < 0) return fail(H8_ERR_ARG, "m < 0");
    if (m == 0) return H8_OK;
    if (!X || !D || !ke) return fail(H8_ERR_ARG, "NULL pointer");
    if (d_stride != 0 && d_stride != 36) return fail(H8_ERR_ARG, "d_stride must be 0 or 36");
    hipStream_t s = (hipStream_t)stream;
    const long long blocks = (m + 3) / 4;
    hipLaunchKernelGGL(k_ke<false>, dim3((unsigned)blocks), dim3(256), 0, s, (long long)m, X, D,
                       (long long)d_stride, ke);
    LAUNCHCHK();
    return H8_OK;
}

extern "C" int h8_ke_bbar_batched_f64(const double* X, const double* D, int64_t d_stride, int64_t m,
                                      double* ke, void* stream) {
    if (m < 0) return fail(H8_ERR_ARG, "m < 0");
    if (m == 0) return H8_OK;
    if (!X || !D || !ke) return fail(H8_ERR_ARG, "NULL pointer");
    if (d_stride != 0 && d_stride != 36) return fail(H8_ERR_ARG, "d_stride must be 0 or 36");
    const long long blocks = (m + 3) / 4;
    hipLaunchKernelGGL(k_ke<true>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, (long long)m, X, D,
                       (long long)d_stride, ke);
    LAUNCHCHK();
    return H8_OK;
}

// ---------------------------------------------------------------------------
// operator
// ---------------------------------------------------------------------------
template <typename T>
static int launch_node_plain(const GridDev& d, const T* s_e, const T* K0, const T* u, T* y,
                             hipStream_t s) {
    NodeArgs<T> A{};
    A.g = d; A.s_e = s_e; A.K0 = K0; A.u = u; A.y = y;
    return launch_op<T, MODE_PLAIN>(d, A, s);
}

extern "C" int h8_spmv_f64(const h8_grid* g, const double* s_e, const double* K0,
                           const double* u, double* y, void* stream) {
    GridDev d;
    int rc = make_grid(g, d);
    if (rc) return rc;
    if (!s_e || !K0 || !u || !y) return fail(H8_ERR_ARG, "NULL pointer");
    return launch_node_plain<double>(d, s_e, K0, u, y, (hipStream_t)stream);
}

extern "C" int h8_spmv_f32(const h8_grid* g, const float* s_e, const float* K0,
                           const float* u, float* y, void* stream) {
    GridDev d;
    int rc = make_grid(g, d);
    if (rc) return rc;
    if (!s_e || !K0 || !u || !y) return fail(H8_ERR_ARG, "NULL pointer");
    return launch_node_plain<float>(d, s_e, K0, u, y, (hipStream_t)stream);
}

static int diag_launch(const h8_grid* g, const double* s_e, const double* K0, const uint8_t* free_mask,
                       double* dinv, int32_t* d_flag, hipStream_t s, bool clear) {
    GridDev d;
    int rc = make_grid(g, d);
    if (rc) return rc;
    if (!s_e || !K0 || !free_mask || !dinv || !d_flag) return fail(H8_ERR_ARG, "NULL pointer");
    if (clear) {
        hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(STREAM_THREADS), 0, s, 1LL, (int*)d_flag, 0);
        LAUNCHCHK();
    }
    hipLaunchKernelGGL(k_diag, dim3(stream_blocks(d.nnodes)), dim3(STREAM_THREADS), 0, s, d, s_e, K0,
                       free_mask, dinv, (int*)d_flag);
    LAUNCHCHK();
    return H8_OK;
}

extern "C" int h8_diag_async_f64(const h8_grid* g, const double* s_e, const double* K0,
                                 const uint8_t* free_mask, double* dinv, int32_t* d_flag, void* stream) {
    return diag_launch(g, s_e, K0, free_mask, dinv, d_flag, (hipStream_t)stream, false);
}

extern "C" int h8_diag_f64(const h8_grid* g, const double* s_e, const double* K0,
                           const uint8_t* free_mask, double* dinv, int32_t* d_flag, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    int rc = diag_launch(g, s_e, K0, free_mask, dinv, d_flag, s, true);
    if (rc) return rc;
    int32_t flag = 0;
    HIPCHK(hipMemcpyAsync(&flag, d_flag, sizeof flag, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (flag) return fail(H8_ZERO_DIAG, "jacobi_preconditioner: A has a zero diagonal entry.");
    return H8_OK;
}

extern "C" int h8_simp_scale_f64(const double* rho, double p, double rho_min, int64_t m,
                                 double* s, double* ds, void* stream) {
    if (m <= 0) return m == 0 ? H8_OK : fail(H8_ERR_ARG, "m < 0");
    if (!rho || !s) return fail(H8_ERR_ARG, "NULL pointer");
    hipLaunchKernelGGL(k_simp, dim3(stream_blocks(m)), dim3(STREAM_THREADS), 0, (hipStream_t)stream,
                       (long long)m, rho, p, rho_min, s, ds);
    LAUNCHCHK();
    return H8_OK;
}

extern "C" int h8_elem_bilinear_f64(const h8_grid* g, const double* K0, const double* u,
                                    const double* v, const double* w, double* out, void* stream) {
    GridDev d;
    int rc = make_grid(g, d);
    if (rc) return rc;
    if (!K0 || !u || !v || !w || !out) return fail(H8_ERR_ARG, "NULL pointer");
    // column tiles x chunks of kd slabs: about 1024 blocks, at least 4 slabs per chunk
    const int tiles = ((d.ny + BIL_BJ - 1) / BIL_BJ) * ((d.nx + BIL_BI - 1) / BIL_BI);
    const int want = std::max(1, (1024 + tiles - 1) / tiles);
    const int kd = std::max(4, (d.nz + want - 1) / want);
    const int chunks = (d.nz + kd - 1) / kd;
    const long long blocks = (long long)tiles * chunks;
    if (blocks >= (1LL << 31)) return fail(H8_ERR_ARG, "grid too large for the element kernel");
    if (u == v)
        hipLaunchKernelGGL(k_elem_bilinear<true>, dim3(unsigned(blocks)), dim3(BIL_BJ * BIL_BI), 0,
                           (hipStream_t)stream, d, K0, u, v, w, out, kd);
    else
        hipLaunchKernelGGL(k_elem_bilinear<false>, dim3(unsigned(blocks)), dim3(BIL_BJ * BIL_BI), 0,
                           (hipStream_t)stream, d, K0, u, v, w, out, kd);
    LAUNCHCHK();
    return H8_OK;
}

// ---------------------------------------------------------------------------
// stress / strain recovery
// ---------------------------------------------------------------------------
extern "C" int h8_recovery_table_f64(double hx, double hy, double hz, const double* D, double* tab,
                                     void* stream) {
    if (!(hx > 0.0) || !(hy > 0.0) || !(hz > 0.0))
        return fail(H8_ERR_ARG, "recovery table: spacings must be > 0 (got %g, %g, %g)", hx, hy, hz);
    if (!tab) return fail(H8_ERR_ARG, "NULL pointer");
    static_assert(H8_RECOVERY_WORDS == REC_WORDS, "table size");
    hipLaunchKernelGGL(k_recovery_table, dim3((REC_WORDS + 255) / 256), dim3(256), 0, (hipStream_t)stream, hx, hy,
                       hz, D, tab);
    LAUNCHCHK();
    return H8_OK;
}

extern "C" int h8_recover_f64(const h8_grid* g, const double* tab, const double* s_e, const double* u,
                              double* nodal, double* vm_nodal, double* elem, double* vm_elem, void* stream) {
    GridDev d;
    int rc = make_grid(g, d);
    if (rc) return rc;
    if (!tab || !u) return fail(H8_ERR_ARG, "NULL pointer");
    hipStream_t s = (hipStream_t)stream;
    if (nodal || vm_nodal) {
        constexpr int TJ = 8, TI = 8, TK = 4;
        set_tiles(d, TJ, TI, TK);
        hipLaunchKernelGGL((k_recover<TJ, TI, TK>), dim3(d.ntiles), dim3(TJ * TI * TK), 0, s, d, tab, s_e, u, nodal,
                           vm_nodal);
        LAUNCHCHK();
    }
    if (elem || vm_elem) {
        const long long m = (long long)d.nx * d.ny * d.nz;
        hipLaunchKernelGGL(k_recover_elem, dim3(unsigned((m + 255) / 256)), dim3(256), 0, s, d, tab, s_e, u, elem,
                           vm_elem);
        LAUNCHCHK();
    }
    return H8_OK;
}

// ---------------------------------------------------------------------------
// SIMP loop: sensitivity filter, OC update
// ---------------------------------------------------------------------------
extern "C" int h8_sens_filter_f64(const h8_grid* g, double hx, double hy, double hz, double radius,
                                  const double* rho, const double* dc, double eps, double* out, void* stream) {
    GridDev d;
    int rc = make_grid(g, d);
    if (rc) return rc;
    if (!rho || !dc || !out) return fail(H8_ERR_ARG, "NULL pointer");
    if (!(hx > 0.0) || !(hy > 0.0) || !(hz > 0.0)) return fail(H8_ERR_ARG, "spacings must be > 0");
    if (!(radius > 0.0) || !std::isfinite(radius)) return fail(H8_ERR_ARG, "filter radius must be > 0 and finite");
    const double rmax = std::max(std::max(radius / hx, radius / hy), radius / hz);
    if (rmax > 64.0)
        return fail(H8_ERR_ARG, "filter radius %g spans more than 64 elements (the stencil is evaluated per element)",
                    radius);
    const int Ri = int(std::ceil(radius / hx)), Rj = int(std::ceil(radius / hy)), Rk = int(std::ceil(radius / hz));
    const long long m = (long long)d.nx * d.ny * d.nz;
    hipLaunchKernelGGL(k_sens_filter, dim3(unsigned((m + 255) / 256)), dim3(256), 0, (hipStream_t)stream, d, hx, hy,
                       hz, radius, Ri, Rj, Rk, rho, dc, eps, out);
    LAUNCHCHK();
    return H8_OK;
}

extern "C" int h8_cast_f64_f32(const double* x, float* y, int64_t n, void* stream) {
    if (n <= 0) return n == 0 ? H8_OK : fail(H8_ERR_ARG, "n < 0");
    if (!x || !y) return fail(H8_ERR_ARG, "NULL pointer");
    hipLaunchKernelGGL(k_cast_f64_f32, dim3(stream_blocks(n)), dim3(STREAM_THREADS), 0,
                       (hipStream_t)stream, (long long)n, x, y);
    LAUNCHCHK();
    return H8_OK;
}

extern "C" int h8_masked_sub_cols_f64(const double* a, const double* b, const double* dinv, double* out,
                                      int64_t n, int64_t ncols, void* stream) {
    if (n < 0 || ncols < 0) return fail(H8_ERR_ARG, "n or ncols < 0");
    if (n == 0 || ncols == 0) return H8_OK;
    if (!a || !dinv || !out) return fail(H8_ERR_ARG, "NULL pointer");
    hipLaunchKernelGGL(k_masked_sub_cols, dim3(stream_blocks(n * ncols)), dim3(STREAM_THREADS), 0,
                       (hipStream_t)stream, (long long)n, (long long)ncols, a, b, dinv, out);
    LAUNCHCHK();
    return H8_OK;
}

extern "C" int h8_masked_sub_f64(const double* a, const double* b, const double* dinv, double* out,
                                 int64_t n, void* stream) {
    if (n <= 0) return n == 0 ? H8_OK : fail(H8_ERR_ARG, "n < 0");
    if (!a || !dinv || !out) return fail(H8_ERR_ARG, "NULL pointer");
    hipLaunchKernelGGL(k_masked_sub, dim3(stream_blocks(n)), dim3(STREAM_THREADS), 0,
                       (hipStream_t)stream, (long long)n, a, b, dinv, out);
    LAUNCHCHK();
    return H8_OK;
}

// ---------------------------------------------------------------------------
// workspace layout
// ---------------------------------------------------------------------------
struct Ws64 {
    double *r, *p0, *p1, *q, *partials;
    PcgState* st;
};
struct WsRef {
    float *r, *x, *p0, *p1, *q, *dinv32;
    double* partials;
    PcgState* st;
};

static size_t partial_slots(const GridDev& d) {
    // four per block of the streaming kernels; the resident solver's sum rings and error word
    return size_t(std::max<long long>(std::max<long long>(d.ntiles, 2048) * 4, RES_RING_DOUBLES + 64));
}

static size_t ws_bytes_64(const GridDev& d) {
    // r, p0, p1, q, partials, state, one spare vector (the timing helper's x),
    // and the resident solver's two further z-exchange vectors (q is the third)
    const size_t nv = align_up(sizeof(double) * d.ndof);
    return 7 * nv + align_up(sizeof(double) * partial_slots(d)) + align_up(sizeof(PcgState));
}
static size_t ws_bytes_ref(const GridDev& d) {
    // r, x, p0, p1, q, dinv32, partials, state, the two z-exchange vectors, and
    // the multigrid form's FP64 copies of f and u (its graphs then hold only
    // pointers that do not change between solves)
    const size_t nv = align_up(sizeof(float) * d.ndof);
    return 8 * nv + align_up(sizeof(double) * partial_slots(d)) + align_up(sizeof(PcgState)) +
           2 * align_up(sizeof(double) * d.ndof);
}

extern "C" size_t h8_pcg_workspace_bytes(const h8_grid* g, int precision) {
    GridDev d;
    if (make_grid(g, d)) return 0;
    if (precision != 32 && precision != 64) {
        fail(H8_ERR_ARG, "precision must be 64 or 32, got %d", precision);
        return 0;
    }
    return precision == 32 ? ws_bytes_ref(d) : ws_bytes_64(d);
}

static char* carve(char*& p, size_t bytes) {
    char* q = p;
    p += align_up(bytes);
    return q;
}

// The resident solver's z-exchange ring: q and two more vectors of the workspace.
template <typename T>
struct ZRing {
    T* z[3];
};
template <typename T>
static ZRing<T> carve_ring(char*& p, const GridDev& d, T* q) {
    ZRing<T> zr;
    zr.z[0] = q;
    zr.z[1] = (T*)carve(p, sizeof(T) * d.ndof);
    zr.z[2] = (T*)carve(p, sizeof(T) * d.ndof);
    return zr;
}

// ---------------------------------------------------------------------------
// graph cache: one captured chunk of iterations per (kind, pointers, grid)
// ---------------------------------------------------------------------------
struct GraphKey {
    int kind;
    int variant;
    int layers;
    int arith;
    int dev;
    long long nx, ny, nz;
    const void* ptr[12];
    // values a captured launch reads from its by-value arguments (bit patterns:
    // tolerances, iteration limits, table sizes) -- a graph replays them as
    // captured, so two calls that differ in any of them need different graphs
    unsigned long long extra[6];
    int chunk;
    bool operator==(const GraphKey& o) const { return std::memcmp(this, &o, sizeof(GraphKey)) == 0; }
};

// A captured graph, destroyed when the last holder lets go: the cache holds
// one reference, and a solve that replays it holds another for as long as it
// launches it, so an eviction by another host thread never destroys an exec
// that is still being launched.
struct GraphExec {
    hipGraphExec_t exec = nullptr;
    ~GraphExec() {
        if (exec) (void)hipGraphExecDestroy(exec);
    }
};
typedef std::shared_ptr<GraphExec> GraphRef;
struct GraphEntry {
    GraphKey key;
    GraphRef exec;
    unsigned long long last_use;
};

// The cache and the capture streams live on the heap and are released only by
// h8_release() (never by static destructors, which would run after the HIP
// runtime has been torn down at process exit).
static std::mutex g_graph_mu;
static std::vector<GraphEntry>* g_graphs = new std::vector<GraphEntry>();
static unsigned long long g_clock = 0;
static hipStream_t g_capture_stream[64] = {};

static int capture_stream(hipStream_t* out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(H8_ERR_HIP, "device index out of range");
    std::lock_guard<std::mutex> lk(g_graph_mu);
    if (!g_capture_stream[dev]) HIPCHK(hipStreamCreateWithFlags(&g_capture_stream[dev], hipStreamNonBlocking));
    *out = g_capture_stream[dev];
    return H8_OK;
}

// Captures body(capture stream) into an instantiated graph.  Captures share the
// device's capture stream, one at a time.
static std::mutex g_capture_mu;
template <typename Body>
static int capture_graph(Body body, GraphExec& out) {
    hipStream_t cs;
    int rc = capture_stream(&cs);
    if (rc) return rc;
    std::lock_guard<std::mutex> ck(g_capture_mu);
    HIPCHK(hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed));
    rc = body(cs);
    hipGraph_t graph;
    const hipError_t ee = hipStreamEndCapture(cs, &graph);
    if (rc) {
        if (ee == hipSuccess) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (ee != hipSuccess) return fail(H8_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ee));
    const hipError_t ie = hipGraphInstantiate(&out.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
        out.exec = nullptr;
        return fail(H8_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ie));
    }
    return H8_OK;
}

// The graph of `key` from the cache (16 entries, least recently used evicted),
// or captured from body(capture stream) and cached.
template <typename Body>
static int get_graph(const GraphKey& key, Body body, GraphRef* out) {
    {
        std::lock_guard<std::mutex> lk(g_graph_mu);
        for (auto& e : *g_graphs)
            if (e.key == key) {
                e.last_use = ++g_clock;
                *out = e.exec;
                return H8_OK;
            }
    }
    auto ref = std::make_shared<GraphExec>();
    const int rc = capture_graph(body, *ref);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_graph_mu);
    if (g_graphs->size() >= 16) {
        auto it = std::min_element(g_graphs->begin(), g_graphs->end(),
                                   [](const GraphEntry& a, const GraphEntry& b) { return a.last_use < b.last_use; });
        g_graphs->erase(it);
    }
    g_graphs->push_back(GraphEntry{key, ref, ++g_clock});
    *out = ref;
    return H8_OK;
}

static unsigned long long dbits(double v) { return __builtin_bit_cast(unsigned long long, v); }

// Destroys the cached graphs and the capture streams (the only HIP objects the
// library keeps between calls).  The host binding calls it at interpreter
// exit, before the HIP runtime is torn down; any later call re-creates what it
// needs.
extern "C" int h8_release(void) {
    std::vector<GraphEntry> drop;
    hipStream_t streams[64];
    {
        std::lock_guard<std::mutex> lk(g_graph_mu);
        drop.swap(*g_graphs);
        for (int i = 0; i < 64; ++i) {
            streams[i] = g_capture_stream[i];
            g_capture_stream[i] = nullptr;
        }
    }
    int dev = 0;
    (void)hipGetDevice(&dev);
    drop.clear();
    for (int i = 0; i < 64; ++i)
        if (streams[i]) {
            (void)hipSetDevice(i);
            (void)hipStreamSynchronize(streams[i]);
            (void)hipStreamDestroy(streams[i]);
        }
    (void)hipSetDevice(dev);
    return H8_OK;
}

// Blocks of the update kernel: about 32 elements of each vector per thread
// (two rounds of the kernel's 4-group unroll in FP32), at least one block per
// CU when there is work, at most 2048 (one partial slot per block).  Fewer,
// longer-running blocks stream better and leave the last block fewer partials:
// on MI355X at 3.26M DOF, 398 blocks take 13.8 us (FP32) and 29.9 us (FP64)
// against 18.7 us and 37.1 us with 16 floats or 8 doubles per thread.
template <typename T>
static int update_blocks(long long n) {
    long long b = (n + STREAM_THREADS * 32LL - 1) / (STREAM_THREADS * 32LL);
    b = std::max(b, std::min(256LL, (n + STREAM_THREADS - 1) / STREAM_THREADS));
    return int(std::max(1LL, std::min(b, 2048LL)));
}

// FP32 iterations on grids of at least this many DOFs with a sector kernel
// defer x += alpha p to the next direction kernel (MODE_DIRX), which has p_old
// in registers while staging it: the update kernel then streams 4 vectors
// instead of 7.  Measured at 3.3M DOF: direction +1.5 us, update -6 us per
// iteration in FP32; in FP64 the direction kernel loses what the update gains,
// and on small grids the latency-bound direction kernel gains nothing.
constexpr long long DEFER_X_MIN_DOF = 1000000;
template <typename T>
static bool defer_x(const GridDev& d) {
    const int v = d.variant == H8_VARIANT_AUTO ? default_variant<T>(d) : d.variant;
    return sizeof(T) == 4 && d.ndof >= DEFER_X_MIN_DOF && (v == H8_VARIANT_SECTOR || v == H8_VARIANT_SECTOR_NARROW);
}

// The direction kernel of one iteration, in the x mode defer_x picks.
template <typename T>
static int launch_dir(const GridDev& d, const NodeArgs<T>& A, hipStream_t s) {
    if constexpr (sizeof(T) == 4)
        if (defer_x<T>(d)) return launch_op<T, MODE_DIRX>(d, A, s);
    return launch_op<T, MODE_DIR>(d, A, s);
}

template <typename T, bool REFINED>
static void launch_update(const GridDev& d, T* x, T* r, T* p0, T* p1, T* q, const T* dinv, PcgState* st,
                          double* partials, hipStream_t s) {
    if (defer_x<T>(d))
        hipLaunchKernelGGL((k_pcg_update<T, REFINED, false>), dim3(update_blocks<T>(d.ndof)), dim3(STREAM_THREADS),
                           0, s, d.ndof, x, r, (const T*)p0, (const T*)p1, (const T*)q, dinv, st, partials,
                           (double*)nullptr);
    else
        hipLaunchKernelGGL((k_pcg_update<T, REFINED, true>), dim3(update_blocks<T>(d.ndof)), dim3(STREAM_THREADS),
                           0, s, d.ndof, x, r, (const T*)p0, (const T*)p1, (const T*)q, dinv, st, partials,
                           (double*)nullptr);
}

// One PCG iteration = direction+product kernel (node-centric) + update kernel.
template <typename T, bool REFINED>
static int launch_iteration(const GridDev& d, const T* s_e, const T* K0, const T* dinv, T* x, T* r,
                            T* p0, T* p1, T* q, PcgState* st, double* partials, hipStream_t s) {
    NodeArgs<T> A{};
    A.g = d; A.s_e = s_e; A.K0 = K0; A.r = r; A.dinv = dinv; A.p0 = p0; A.p1 = p1; A.q = q; A.x = x;
    A.st = st; A.partials = partials;
    int rc = launch_dir<T>(d, A, s);
    if (rc) return rc;
    launch_update<T, REFINED>(d, x, r, p0, p1, q, dinv, st, partials, s);
    LAUNCHCHK();
    return H8_OK;
}

// Iterations captured per graph: once the device flag leaves RUNNING the rest
// of the chunk returns at kernel entry, so a chunk bounds the wasted launches.
static int chunk_for(long long maxiter) {
    return int(std::max(1LL, std::min<long long>(maxiter, 32)));
}

// ---------------------------------------------------------------------------
// Resident solver (one cooperative launch per solve / refined phase)
// ---------------------------------------------------------------------------
// run_resident's answer when the device cannot hold all blocks at once and the
// variant was the library's choice (not an error: the streaming loop runs)
constexpr int RES_DECLINED = -1;

// Threads per block of k_resident by (storage, element-product arithmetic).
template <typename T, typename TA> struct ResThreads;
template <> struct ResThreads<float, float> { static constexpr int value = 1024; };
template <> struct ResThreads<float, double> { static constexpr int value = 1024; };
template <> struct ResThreads<double, double> { static constexpr int value = 512; };
template <typename T> static int res_threads(const GridDev& d) {
    return fp64_products<T>(d) ? ResThreads<T, double>::value : ResThreads<T, T>::value;
}

// Most elements touched by one part of the split a N / P of N nodes: its nodes
// plus one, less one on each side that lies on the grid boundary.
static int res_axis_elems(int N, int P) {
    int m = 0;
    for (int a = 0; a < P; ++a) {
        const int s = int((long long)(a + 1) * N / P - (long long)a * N / P);
        m = std::max(m, s + 1 - (a == 0) - (a == P - 1));
    }
    return m;
}

// Brick split for the resident solver: at most one block per CU, LDS within
// 158 KiB (the kernel's static LDS stays below 2 KiB of the CU's 160).  Plans
// are ranked by the element batches (elements / threads, rounded up) of the
// busiest block, then by W = its elements + 1.5 x its shell nodes on faces
// across j, then by the fewest blocks.  j is the fastest axis of the node
// numbering, so a face across j is one scattered 12- or 24-byte word group per
// node in the z exchange, where faces across i or k are runs along j.  On
// MI355X at 128x32x32 (FP32, 100-iteration launches, s_memrealtime stamps):
// 32x1x8 bricks (W = 800) 12.1 us per iteration, 16x2x8 (881) 12.4,
// 26x3x3 (1137) 13.3, 16x4x4 (1092) 13.6; FP64 (two element passes):
// 32x2x4 (881) 17.1, 16x2x8 (881) 17.0, 22x1x11 (896) 17.4, 26x3x3 17.5.
template <typename T>
static bool res_plan_search(const GridDev& d, int nt, int max_blocks, ResGeom& best) {
    constexpr long long LDS_MAX = 158 * 1024;
    bool found = false;
    long long bcost = 0;
    int bnb = 0;
    for (int PK = 1; PK <= std::min(d.NK, max_blocks); ++PK)
        for (int PI = 1; PI <= std::min(d.NI, max_blocks / PK); ++PI)
            for (int PJ = 1; PJ <= std::min(d.NJ, max_blocks / (PK * PI)); ++PJ) {
                ResGeom c{};
                c.PI = PI; c.PJ = PJ; c.PK = PK;
                c.BI = (d.NI + PI - 1) / PI;
                c.BJ = (d.NJ + PJ - 1) / PJ;
                c.BK = (d.NK + PK - 1) / PK;
                c.HI = c.BI + 2;
                c.HJ = c.BJ + 2;
                c.nown = c.BI * c.BJ * c.BK;
                c.nhalo = c.HI * c.HJ * (c.BK + 2);
                c.nelem = (c.BI + 1) * (c.BJ + 1) * (c.BK + 1);
                long long off = 0;
                auto take = [&](long long bytes) {
                    const long long o = off;
                    off += (bytes + 15) & ~15LL;
                    return int(o);
                };
                c.o_p = take(3LL * c.nhalo * sizeof(T));
                c.o_z = take(3LL * c.nhalo * sizeof(T));
                c.o_r = take(3LL * c.nown * sizeof(T));
                c.o_x = take(3LL * c.nown * sizeof(T));
                c.o_d = take(3LL * c.nown * sizeof(T));
                c.o_q = take(3LL * c.nown * sizeof(T));
                c.o_s = take((long long)c.nelem * sizeof(T));
                if (off > LDS_MAX || c.nhalo >= (1 << 15)) continue;   // (index decoding needs < 2^16)
                c.lds_bytes = int(off);
                const int nb = PI * PJ * PK;
                const long long emax = (long long)res_axis_elems(d.NI, PI) * res_axis_elems(d.NJ, PJ) *
                                       res_axis_elems(d.NK, PK);
                const long long jface = (long long)std::min(PJ - 1, 2) * c.HI * (c.BK + 2);
                const long long cost = ((emax + nt - 1) / nt) * (1LL << 32) + 2 * emax + 3 * jface;   // 2 W
                if (!found || cost < bcost || (cost == bcost && nb < bnb)) {
                    found = true;
                    bcost = cost;
                    bnb = nb;
                    best = c;
                }
            }
    return found;
}

// res_plan_search, remembered per (grid, threads, block limit, word size): the
// search visits thousands of splits and runs before every resident launch (one
// per FP32 phase of a refined solve).
template <typename T>
static bool res_plan(const GridDev& d, int nt, int max_blocks, ResGeom& best) {
    struct Entry {
        int NI, NJ, NK, nt, max_blocks;
        bool found;
        ResGeom geom;
    };
    static std::mutex mu;
    static std::vector<Entry> memo;
    std::lock_guard<std::mutex> lock(mu);
    for (const Entry& e : memo)
        if (e.NI == d.NI && e.NJ == d.NJ && e.NK == d.NK && e.nt == nt && e.max_blocks == max_blocks) {
            best = e.geom;
            return e.found;
        }
    Entry e{d.NI, d.NJ, d.NK, nt, max_blocks, false, ResGeom{}};
    e.found = res_plan_search<T>(d, nt, max_blocks, e.geom);
    if (memo.size() >= 64) memo.erase(memo.begin());
    memo.push_back(e);
    best = e.geom;
    return e.found;
}

// The resident solver runs the iterations when asked for (H8_VARIANT_RESIDENT)
// or, with H8_VARIANT_AUTO, whenever the grid fits it: it then keeps every
// vector on chip and replaces two launches per iteration by two grid sums.
template <typename T>
static bool use_resident(const GridDev& d) {
    if (d.variant == H8_VARIANT_RESIDENT) return true;
    if (d.variant != H8_VARIANT_AUTO) return false;
    ResGeom G;
    return res_plan<T>(d, res_threads<T>(d), std::min(device_cus(), RES_MAX_BLOCKS), G);
}

// Runs the iterations until the device-side status leaves RUNNING in one
// cooperative launch of k_resident (q serves as the z exchange vector).
// ev0/ev1 (optional): recorded just before and just after the launch.
template <typename T, typename TA, bool REFINED>
static int run_resident_k(const GridDev& d, const T* s_e, const T* K0, const T* dinv, T* x, T* r, T* p0, T* p1,
                          const ZRing<T>& zr, PcgState* st, double* partials, hipStream_t s, PcgState& hst,
                          long long& launches, hipEvent_t ev0, hipEvent_t ev1, bool sync) {
    constexpr int NT = ResThreads<T, TA>::value;
    const void* fn = (const void*)k_resident<T, NT, REFINED, TA>;
    ResGeom G;
    if (!res_plan<T>(d, NT, std::min(device_cus(), RES_MAX_BLOCKS), G))
        return fail(H8_ERR_ARG, "resident solver: a %dx%dx%d grid does not fit in LDS (one block per CU)", d.nx,
                    d.ny, d.nz);
    const int nb = G.PI * G.PJ * G.PK;
    int coop = 0, dev = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    if (!coop) return fail(H8_ERR_HIP, "resident solver: the device has no cooperative launch");
    HIPCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, G.lds_bytes));
    int per = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, NT, G.lds_bytes));
    if (per < 1 || nb > per * device_cus())
        return fail(H8_ERR_HIP, "resident solver: %d blocks of %d threads and %d bytes of LDS cannot all be resident",
                    nb, NT, G.lds_bytes);
    ResArgs<T> A{};
    A.g = d; A.G = G; A.s_e = s_e; A.kh = K0 + 576; A.kh64 = k0_table64(K0); A.dinv = dinv; A.x = x; A.r = r;
    A.p0 = p0; A.p1 = p1;
    for (int e = 0; e < 3; ++e) A.zr[e] = zr.z[e];
    A.ring = partials; A.err = reinterpret_cast<unsigned*>(partials + RES_RING_DOUBLES); A.st = st;
    hipLaunchKernelGGL(k_ring_reset<T>, dim3(stream_blocks(3 * d.ndof)), dim3(256), 0, s,
                       reinterpret_cast<unsigned long long*>(partials), RES_RING_DOUBLES, A.err, A);
    LAUNCHCHK();
    void* args[] = {&A};
    if (ev0) HIPCHK(hipEventRecord(ev0, s));
    const hipError_t le = hipLaunchCooperativeKernel(fn, dim3(nb), dim3(NT), args, unsigned(G.lds_bytes), s);
    if (le == hipErrorCooperativeLaunchTooLarge && d.variant == H8_VARIANT_AUTO) {
        (void)hipGetLastError();
        return RES_DECLINED;      // the caller runs the streaming kernels instead
    }
    HIPCHK(le);
    if (ev1) HIPCHK(hipEventRecord(ev1, s));
    ++launches;
    if (!sync) return H8_OK;      // the caller reads the state and the error word
    unsigned err = 0;
    HIPCHK(hipMemcpyAsync(&hst, st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&err, A.err, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (err) return fail(H8_ERR_HIP, "resident solver: a grid-wide sum timed out");
    return H8_OK;
}

template <typename T, bool REFINED>
static int run_resident(const GridDev& d, const T* s_e, const T* K0, const T* dinv, T* x, T* r, T* p0, T* p1,
                        const ZRing<T>& zr, PcgState* st, double* partials, hipStream_t s, PcgState& hst,
                        long long& launches, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr,
                        bool sync = true) {
    if (fp64_products<T>(d))
        return run_resident_k<T, double, REFINED>(d, s_e, K0, dinv, x, r, p0, p1, zr, st, partials, s, hst, launches,
                                                  ev0, ev1, sync);
    return run_resident_k<T, T, REFINED>(d, s_e, K0, dinv, x, r, p0, p1, zr, st, partials, s, hst, launches, ev0,
                                         ev1, sync);
}

// Runs captured chunks until the device-side status leaves RUNNING.
template <typename T, bool REFINED>
static int run_chunks(const GraphKey& key, const GridDev& d, const T* s_e, const T* K0, const T* dinv,
                      T* x, T* r, T* p0, T* p1, T* q, const ZRing<T>& zr, PcgState* st, double* partials,
                      hipStream_t s, PcgState& hst, long long& launches) {
    if (use_resident<T>(d)) {
        const int rr = run_resident<T, REFINED>(d, s_e, K0, dinv, x, r, p0, p1, zr, st, partials, s, hst, launches);
        if (rr != RES_DECLINED) return rr;
    }
    const int chunk = key.chunk;
    GraphRef exec;
    int rc = get_graph(key, [&](hipStream_t cs) -> int {
        for (int it = 0; it < chunk; ++it) {
            int r1 = launch_iteration<T, REFINED>(d, s_e, K0, dinv, x, r, p0, p1, q, st, partials, cs);
            if (r1) return r1;
        }
        return H8_OK;
    }, &exec);
    if (rc) return rc;
    for (;;) {
        HIPCHK(hipGraphLaunch(exec->exec, s));
        ++launches;
        HIPCHK(hipMemcpyAsync(&hst, st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (hst.status != ST_RUNNING) break;
    }
    return H8_OK;
}

static void fill_stats(h8_pcg_stats* stats, const PcgState& h, float ms, long long launches) {
    if (!stats) return;
    stats->iterations = h.iters;
    stats->status = h.status == ST_DONE ? H8_OK : (h.status == ST_NOT_SPD ? H8_NOT_SPD : H8_MAXITER);
    stats->outer = 0;
    stats->rel_residual = h.rel;
    stats->true_rel_residual = -1.0;
    stats->bnorm = h.bnorm;
    stats->pAp = h.bad_pq;
    stats->elapsed_ms = ms;
    stats->inner_iterations = 0;
    stats->launches = launches;
}

// OC update (the SIMP loop's density step)
static size_t oc_partials(long long m) { return size_t(stream_blocks(m)); }

extern "C" size_t h8_oc_workspace_bytes(int64_t m) {
    if (m < 1) return 0;
    return align_up(sizeof(OcState)) + align_up(sizeof(double) * oc_partials(m)) + align_up(sizeof(unsigned long long));
}

extern "C" int h8_oc_update_f64(const double* rho, const double* dc, int64_t m, double volume_fraction, double move,
                                double rho_min_bound, double rho_max_bound, double bisection_tol,
                                int32_t max_bisection_iter, double* rho_new, double* change, void* ws,
                                size_t ws_bytes, void* stream) {
    if (m < 1) return fail(H8_ERR_ARG, "m must be >= 1");
    if (!rho || !dc || !rho_new || !ws) return fail(H8_ERR_ARG, "NULL pointer");
    if (ws_bytes < h8_oc_workspace_bytes(m)) return fail(H8_ERR_WORKSPACE, "workspace too small");
    if (max_bisection_iter < 1) return fail(H8_ERR_ARG, "max_bisection_iter must be >= 1");
    hipStream_t s = (hipStream_t)stream;
    char* p = (char*)ws;
    OcState* st = (OcState*)carve(p, sizeof(OcState));
    double* partials = (double*)carve(p, sizeof(double) * oc_partials(m));
    unsigned long long* cmax = (unsigned long long*)carve(p, sizeof(unsigned long long));
    OcState h{};
    h.l1 = 1e-12;
    h.l2 = 1.0;
    HIPCHK(hipMemcpyAsync(st, &h, sizeof h, hipMemcpyHostToDevice, s));
    OcParams P{volume_fraction, move, rho_min_bound, rho_max_bound, bisection_tol, max_bisection_iter};
    const int nb = stream_blocks(m);
    // every pass either widens the bracket (at most 60) or bisects it (at most
    // max_bisection_iter): launch in chunks, poll the phase between chunks
    const long long max_pass = OC_MAX_EXPAND + (long long)max_bisection_iter;
    for (long long done = 0; done < max_pass;) {
        const long long chunk = std::min<long long>(16, max_pass - done);
        for (long long i = 0; i < chunk; ++i)
            hipLaunchKernelGGL(k_oc_pass, dim3(nb), dim3(STREAM_THREADS), 0, s, (long long)m, rho, dc, P, st, partials,
                               rho_new);
        LAUNCHCHK();
        done += chunk;
        HIPCHK(hipMemcpyAsync(&h, st, sizeof h, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (h.phase == 2) break;
    }
    if (h.phase != 2) return fail(H8_ERR_HIP, "OC update: bisection did not finish in %lld passes", max_pass);
    if (change) {
        unsigned long long bits = 0;
        HIPCHK(hipMemcpyAsync(cmax, &bits, sizeof bits, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_absdiff_max, dim3(nb), dim3(STREAM_THREADS), 0, s, (long long)m, rho_new, rho, cmax);
        LAUNCHCHK();
        HIPCHK(hipMemcpyAsync(&bits, cmax, sizeof bits, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        *change = __builtin_bit_cast(double, bits);
    }
    return H8_OK;
}

// ---------------------------------------------------------------------------
// FP64 Jacobi-PCG
// ---------------------------------------------------------------------------
// The iterations of run_chunks launched directly on the stream, `chunk` per
// host poll, with no captured graph (nothing instantiated or allocated).
template <typename T, bool REFINED>
static int run_direct(int chunk, const GridDev& d, const T* s_e, const T* K0, const T* dinv, T* x, T* r, T* p0,
                      T* p1, T* q, const ZRing<T>& zr, PcgState* st, double* partials, hipStream_t s, PcgState& hst,
                      long long& launches) {
    if (use_resident<T>(d)) {
        const int rr = run_resident<T, REFINED>(d, s_e, K0, dinv, x, r, p0, p1, zr, st, partials, s, hst, launches);
        if (rr != RES_DECLINED) return rr;
    }
    for (;;) {
        for (int it = 0; it < chunk; ++it) {
            const int r1 = launch_iteration<T, REFINED>(d, s_e, K0, dinv, x, r, p0, p1, q, st, partials, s);
            if (r1) return r1;
        }
        ++launches;
        HIPCHK(hipMemcpyAsync(&hst, st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (hst.status != ST_RUNNING) break;
    }
    return H8_OK;
}

// Jacobi-PCG in FP64 on a workspace of ws_bytes_64(d) bytes at ws.  noalloc:
// no events (elapsed_ms = -1) and no captured graphs, so the call creates,
// instantiates and allocates nothing.
static int pcg_f64_impl(const GridDev& d, const double* s_e, const double* K0, const double* dinv, const double* b,
                        double* x, int use_x0, double tol, int64_t maxiter, void* ws, h8_pcg_stats* stats,
                        hipStream_t s, bool noalloc) {
    int rc;
    char* p = (char*)ws;
    Ws64 w;
    w.r = (double*)carve(p, sizeof(double) * d.ndof);
    w.p0 = (double*)carve(p, sizeof(double) * d.ndof);
    w.p1 = (double*)carve(p, sizeof(double) * d.ndof);
    w.q = (double*)carve(p, sizeof(double) * d.ndof);
    w.partials = (double*)carve(p, sizeof(double) * partial_slots(d));
    w.st = (PcgState*)carve(p, sizeof(PcgState));
    (void)carve(p, sizeof(double) * d.ndof);                 // the timing helper's x
    const ZRing<double> zr = carve_ring<double>(p, d, w.q);

    Ev e0, e1;
    if (!noalloc) {
        HIPCHK(e0.create());
        HIPCHK(e1.create());
        HIPCHK(hipEventRecord(e0, s));
    }

    PcgState hst{};
    hst.tol = tol;
    hst.maxiter = maxiter;
    hst.status = ST_RUNNING;
    HIPCHK(hipMemcpyAsync(w.st, &hst, sizeof(PcgState), hipMemcpyHostToDevice, s));
    const double* Ax0 = nullptr;
    if (use_x0) {
        rc = launch_node_plain<double>(d, s_e, K0, x, w.q, s);
        if (rc) return rc;
        Ax0 = w.q;
    } else {
        hipLaunchKernelGGL(k_fill<double>, dim3(stream_blocks(d.ndof)), dim3(STREAM_THREADS), 0, s,
                           d.ndof, x, 0.0);
        LAUNCHCHK();
    }
    hipLaunchKernelGGL(k_pcg_init<double>, dim3(stream_blocks(d.ndof)), dim3(STREAM_THREADS), 0, s,
                       d.ndof, b, Ax0, dinv, w.r, w.st, w.partials);
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(&hst, w.st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    long long launches = 0;
    if (hst.status == ST_RUNNING && maxiter > 0) {
        if (noalloc) {
            rc = run_direct<double, false>(chunk_for(maxiter), d, s_e, K0, dinv, x, w.r, w.p0, w.p1, w.q, zr, w.st,
                                           w.partials, s, hst, launches);
        } else {
            GraphKey key{};
            key.kind = 1;
            key.variant = d.variant;
            key.layers = d.layers;
            key.arith = d.arith;
            HIPCHK(hipGetDevice(&key.dev));
            key.nx = d.nx; key.ny = d.ny; key.nz = d.nz;
            const void* ptrs[] = {s_e, K0, dinv, x, w.r, w.p0, w.p1, w.q, w.st, w.partials};
            std::memcpy(key.ptr, ptrs, sizeof ptrs);
            key.chunk = chunk_for(maxiter);
            rc = run_chunks<double, false>(key, d, s_e, K0, dinv, x, w.r, w.p0, w.p1, w.q, zr, w.st, w.partials, s,
                                           hst, launches);
        }
        if (rc) return rc;
        // the last iteration's x += alpha p (the update kernel defers it to the next direction kernel)
        if (hst.pending) {
            hipLaunchKernelGGL(k_pending_axpy<double>, dim3(stream_blocks(d.ndof)), dim3(STREAM_THREADS), 0, s,
                               d.ndof, x, (hst.k & 1) ? (const double*)w.p1 : (const double*)w.p0, w.st);
            LAUNCHCHK();
        }
    } else if (hst.status == ST_RUNNING) {
        hst.status = ST_MAXITER;
        hst.iters = 0;
    }
    float ms = -1.f;
    if (!noalloc) {
        HIPCHK(hipEventRecord(e1, s));
        HIPCHK(hipEventSynchronize(e1));
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    } else {
        HIPCHK(hipStreamSynchronize(s));
    }
    fill_stats(stats, hst, ms, launches);
    if (hst.status == ST_NOT_SPD)
        return fail(H8_NOT_SPD, "preconditioned_cg: p^T A p = %.3e <= 0 at iteration %lld -- A is not SPD.",
                    hst.bad_pq, hst.iters);
    if (hst.status == ST_MAXITER)
        return fail(H8_MAXITER,
                    "preconditioned_cg: did not converge in %lld iterations (final relative residual "
                    "%.3e >= tol=%.1e).",
                    (long long)maxiter, hst.rel, tol);
    return H8_OK;
}

static int pcg_f64_args(const h8_grid* g, GridDev& d, const double* s_e, const double* K0, const double* dinv,
                        const double* b, double* x, double tol, int64_t maxiter) {
    int rc = make_grid(g, d);
    if (rc) return rc;
    if (!s_e || !K0 || !dinv || !b || !x) return fail(H8_ERR_ARG, "NULL pointer");
    if (!(tol >= 0.0)) return fail(H8_ERR_ARG, "tol must be >= 0");
    if (maxiter < 0) return fail(H8_ERR_ARG, "maxiter must be >= 0");
    return H8_OK;
}

extern "C" int h8_pcg_f64(const h8_grid* g, const double* s_e, const double* K0, const double* dinv,
                          const double* b, double* x, int use_x0, double tol, int64_t maxiter,
                          void* ws, size_t ws_bytes, h8_pcg_stats* stats, void* stream) {
    GridDev d;
    int rc = pcg_f64_args(g, d, s_e, K0, dinv, b, x, tol, maxiter);
    if (rc) return rc;
    if (!ws) return fail(H8_ERR_ARG, "NULL pointer");
    if (ws_bytes < ws_bytes_64(d))
        return fail(H8_ERR_WORKSPACE, "workspace too small: %zu < %zu", ws_bytes, ws_bytes_64(d));
    return pcg_f64_impl(d, s_e, K0, dinv, b, x, use_x0, tol, maxiter, ws, stats, (hipStream_t)stream, false);
}

// ---------------------------------------------------------------------------
// FP64 Jacobi-PCG that allocates nothing: the workspace of h8_pcg_f64 in a
// file-scope device array (each GPU holds its own copy until the process ends)
// ---------------------------------------------------------------------------
static_assert(H8_SCRATCH_BYTES % 256 == 0, "scratch size");
static __device__ __align__(256) unsigned char scratch[H8_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64] = {};

// The device address of `scratch` on the current device, looked up once per
// device with that device current (host code never passes the symbol itself).
static int scratch_address(void** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(H8_ERR_HIP, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(H8_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return H8_OK;
}

extern "C" int h8_pcg_f64_noalloc(const h8_grid* g, const double* s_e, const double* K0, const double* dinv,
                                  const double* b, double* x, int use_x0, double tol, int64_t maxiter,
                                  h8_pcg_stats* stats, void* stream) {
    GridDev d;
    int rc = pcg_f64_args(g, d, s_e, K0, dinv, b, x, tol, maxiter);
    if (rc) return rc;
    const size_t need = ws_bytes_64(d);
    if (need > H8_SCRATCH_BYTES)
        return fail(H8_ERR_WORKSPACE,
                    "h8_pcg_f64_noalloc: a %dx%dx%d grid needs %zu bytes of workspace, more than the library's "
                    "static scratch of %zu bytes; use h8_pcg_f64 with a workspace",
                    d.nx, d.ny, d.nz, need, (size_t)H8_SCRATCH_BYTES);
    void* ws = nullptr;
    if ((rc = scratch_address(&ws))) return rc;
    return pcg_f64_impl(d, s_e, K0, dinv, b, x, use_x0, tol, maxiter, ws, stats, (hipStream_t)stream, true);
}

// ---------------------------------------------------------------------------
// Mixed precision: FP32 PCG with FP64 reliable residual replacement
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Linear multi-point constraints (eliminated; Jacobi-PCG on T^T K T)
// ---------------------------------------------------------------------------
static int cmap_dev(const h8_cmap* c, CMapDev& m) {
    if (!c) return fail(H8_ERR_ARG, "constraint map is NULL");
    if (c->n_slave < 0 || c->n_master < 0) return fail(H8_ERR_ARG, "negative constraint counts");
    if (c->n_slave > 0 && (!c->slave || !c->s_ptr || !c->s_master || !c->s_coef))
        return fail(H8_ERR_ARG, "constraint map: NULL slave arrays");
    if (c->n_master > 0 && (!c->master || !c->m_ptr || !c->m_slave || !c->m_coef))
        return fail(H8_ERR_ARG, "constraint map: NULL master arrays");
    if (c->n_slave + c->n_master >= (1LL << 31)) return fail(H8_ERR_ARG, "constraint map too large");
    m = CMapDev{c->n_slave, c->n_master, c->slave, c->s_ptr, c->s_master, c->s_coef,
                c->master, c->m_ptr, c->m_slave, c->m_coef};
    return H8_OK;
}

extern "C" int h8_constraint_expand_f64(const h8_cmap* c, double* v, void* stream) {
    CMapDev m;
    int rc = cmap_dev(c, m);
    if (rc) return rc;
    if (!v) return fail(H8_ERR_ARG, "NULL pointer");
    if (m.n_slave == 0) return H8_OK;
    hipLaunchKernelGGL(k_cexpand, dim3(unsigned((m.n_slave + 255) / 256)), dim3(256), 0, (hipStream_t)stream, m, v,
                       (const PcgState*)nullptr);
    LAUNCHCHK();
    return H8_OK;
}

extern "C" int h8_constraint_contract_f64(const h8_cmap* c, double* v, void* stream) {
    CMapDev m;
    int rc = cmap_dev(c, m);
    if (rc) return rc;
    if (!v) return fail(H8_ERR_ARG, "NULL pointer");
    if (m.n_master == 0) return H8_OK;
    hipLaunchKernelGGL(k_cfix, dim3(unsigned((m.n_master + 255) / 256)), dim3(256), 0, (hipStream_t)stream, m,
                       (double*)nullptr, v, (const PcgState*)nullptr);
    LAUNCHCHK();
    return H8_OK;
}

extern "C" int h8_constraint_diag_f64(const h8_grid* g, const double* s_e, const double* K0, const h8_cmap* c,
                                      double* dinv, void* stream) {
    GridDev d;
    int rc = make_grid(g, d);
    if (rc) return rc;
    CMapDev m;
    rc = cmap_dev(c, m);
    if (rc) return rc;
    if (!s_e || !K0 || !dinv) return fail(H8_ERR_ARG, "NULL pointer");
    if (m.n_master == 0) return H8_OK;
    hipLaunchKernelGGL(k_cdiag, dim3(unsigned((m.n_master + 255) / 256)), dim3(256), 0, (hipStream_t)stream, d, s_e,
                       K0, m, dinv);
    LAUNCHCHK();
    return H8_OK;
}

extern "C" int h8_pcg_constrained_f64(const h8_grid* g, const double* s_e, const double* K0, const double* dinv,
                                      const h8_cmap* c, const double* b, double* x, double tol, int64_t maxiter,
                                      void* ws, size_t ws_bytes, h8_pcg_stats* stats, void* stream) {
    GridDev d;
    int rc = make_grid(g, d);
    if (rc) return rc;
    CMapDev m;
    rc = cmap_dev(c, m);
    if (rc) return rc;
    if (!s_e || !K0 || !dinv || !b || !x || !ws) return fail(H8_ERR_ARG, "NULL pointer");
    if (ws_bytes < ws_bytes_64(d))
        return fail(H8_ERR_WORKSPACE, "workspace too small: %zu < %zu", ws_bytes, ws_bytes_64(d));
    if (!(tol >= 0.0)) return fail(H8_ERR_ARG, "tol must be >= 0");
    if (maxiter < 0) return fail(H8_ERR_ARG, "maxiter must be >= 0");
    hipStream_t s = (hipStream_t)stream;
    char* p = (char*)ws;
    double* r = (double*)carve(p, sizeof(double) * d.ndof);
    double* pv = (double*)carve(p, sizeof(double) * d.ndof);
    (void)carve(p, sizeof(double) * d.ndof);
    double* q = (double*)carve(p, sizeof(double) * d.ndof);
    double* partials = (double*)carve(p, sizeof(double) * partial_slots(d));
    PcgState* st = (PcgState*)carve(p, sizeof(PcgState));

    Ev e0, e1;
    HIPCHK(e0.create());
    HIPCHK(e1.create());
    HIPCHK(hipEventRecord(e0, s));
    const int sb = stream_blocks(d.ndof);
    hipLaunchKernelGGL(k_fill<double>, dim3(sb), dim3(STREAM_THREADS), 0, s, d.ndof, x, 0.0);
    hipLaunchKernelGGL(k_fill<double>, dim3(sb), dim3(STREAM_THREADS), 0, s, d.ndof, pv, 0.0);
    LAUNCHCHK();
    PcgState hst{};
    hst.tol = tol;
    hst.maxiter = maxiter;
    hst.status = ST_RUNNING;
    HIPCHK(hipMemcpyAsync(st, &hst, sizeof(PcgState), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_pcg_init<double>, dim3(sb), dim3(STREAM_THREADS), 0, s, d.ndof, b, (const double*)nullptr,
                       dinv, r, st, partials);
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(&hst, st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    long long launches = 0;
    if (hst.status == ST_RUNNING && maxiter > 0) {
        // one iteration: p = dinv r + beta p; expand p to the slaves; q = K p;
        // contract q to the masters and zero p at the slaves; p.q -> alpha;
        // x += alpha p, r -= alpha q, r.z, r.r -> beta / stop
        NodeArgs<double> A{};
        A.g = d; A.s_e = s_e; A.K0 = K0; A.u = pv; A.y = q;
        const unsigned gs = unsigned(std::max<long long>(1, (m.n_slave + 255) / 256));
        const unsigned gf = unsigned(std::max<long long>(1, (m.n_master + m.n_slave + 255) / 256));
        const int chunk = chunk_for(maxiter);
        GraphExec exec;       // destroyed on every return
        rc = capture_graph([&](hipStream_t cs) -> int {
            for (int it = 0; it < chunk; ++it) {
                hipLaunchKernelGGL(k_cdir, dim3(sb), dim3(STREAM_THREADS), 0, cs, d.ndof, r, dinv, pv, st);
                if (m.n_slave > 0) hipLaunchKernelGGL(k_cexpand, dim3(gs), dim3(256), 0, cs, m, pv, st);
                const int r1 = launch_op<double, MODE_PLAIN>(d, A, cs);
                if (r1) return r1;
                if (m.n_slave + m.n_master > 0)
                    hipLaunchKernelGGL(k_cfix, dim3(gf), dim3(256), 0, cs, m, pv, q, st);
                hipLaunchKernelGGL(k_cpq, dim3(sb), dim3(STREAM_THREADS), 0, cs, d.ndof, pv, q, st, partials);
                hipLaunchKernelGGL((k_pcg_update<double, false, true>), dim3(update_blocks<double>(d.ndof)),
                                   dim3(STREAM_THREADS), 0, cs, d.ndof, x, r, pv, pv, q, dinv, st, partials,
                                   (double*)nullptr);
            }
            LAUNCHCHK();
            return H8_OK;
        }, exec);
        if (rc) return rc;
        for (;;) {
            HIPCHK(hipGraphLaunch(exec.exec, s));
            ++launches;
            HIPCHK(hipMemcpyAsync(&hst, st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            if (hst.status != ST_RUNNING) break;
        }
    } else if (hst.status == ST_RUNNING) {
        hst.status = ST_MAXITER;
        hst.iters = 0;
    }
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    fill_stats(stats, hst, ms, launches);
    if (hst.status == ST_NOT_SPD)
        return fail(H8_NOT_SPD, "preconditioned_cg: p^T A p = %.3e <= 0 at iteration %lld -- A is not SPD.",
                    hst.bad_pq, hst.iters);
    if (hst.status == ST_MAXITER)
        return fail(H8_MAXITER, "preconditioned_cg: did not converge in %lld iterations (rel residual %.3e).",
                    (long long)maxiter, hst.rel);
    return H8_OK;
}

// FP32 phases of the refined solve on the resident solver.  A phase is one
// resident launch with the fold and the FP64 residual queued behind it (both
// act only when the phase stopped for a replacement); its state and error word
// are copied to a pinned host slot and an event is recorded.  Phase j + 1 is
// queued before the host waits for phase j, so the host's reading and
// queueing overlap the GPU's work; a phase queued after the solve ended does
// nothing (the resident kernel returns at once unless the state is RUNNING).
struct RefineSlot {
    PcgState st;
    unsigned err;
};

static int queue_refine_phase(const GridDev& d, const float* s_e32, const float* K032, const WsRef& w,
                              const ZRing<float>& zr, double* u, const NodeArgs<double>& R, hipStream_t s,
                              PcgState& hst, long long& launches, RefineSlot* slot, hipEvent_t ev) {
    int rc = run_resident<float, true>(d, s_e32, K032, w.dinv32, w.x, w.r, w.p0, w.p1, zr, w.st, w.partials, s,
                                       hst, launches, nullptr, nullptr, false);
    if (rc) return rc;
    hipLaunchKernelGGL(k_fold, dim3(stream_blocks(d.ndof)), dim3(STREAM_THREADS), 0, s, d.ndof, u, w.x,
                       (const float*)w.p0, (const float*)w.p1, (const PcgState*)w.st);
    LAUNCHCHK();
    NodeArgs<double> Rq = R;
    Rq.only_on_replace = 1;
    rc = launch_op<double, MODE_RESID>(d, Rq, s);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(&slot->st, w.st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&slot->err, reinterpret_cast<unsigned*>(w.partials + RES_RING_DOUBLES), sizeof(unsigned),
                          hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(ev, s));
    return H8_OK;
}

static int refine_resident_phases(const GridDev& d, const float* s_e32, const float* K032, const WsRef& w,
                                  const ZRing<float>& zr, double* u, const NodeArgs<double>& R, hipStream_t s,
                                  PcgState& hst, long long& launches, RefineSlot* hs, hipEvent_t* ev) {
    int cur = 0;
    int rc = queue_refine_phase(d, s_e32, K032, w, zr, u, R, s, hst, launches, &hs[0], ev[0]);
    if (rc) return rc;                                   // RES_DECLINED: nothing ran
    for (;;) {
        rc = queue_refine_phase(d, s_e32, K032, w, zr, u, R, s, hst, launches, &hs[cur ^ 1], ev[cur ^ 1]);
        if (rc == RES_DECLINED) return fail(H8_ERR_HIP, "resident solver: a cooperative launch was refused");
        if (rc) return rc;
        HIPCHK(hipEventSynchronize(ev[cur]));
        hst = hs[cur].st;
        if (hs[cur].err) return fail(H8_ERR_HIP, "resident solver: a grid-wide sum timed out");
        if (hst.status != ST_RUNNING) {
            --launches;                                  // the phase queued last did nothing
            return H8_OK;
        }
        cur ^= 1;
    }
}

static int refine_resident(const GridDev& d, const float* s_e32, const float* K032, const WsRef& w,
                           const ZRing<float>& zr, double* u, const NodeArgs<double>& R, hipStream_t s,
                           PcgState& hst, long long& launches) {
    RefineSlot* hs = nullptr;
    HIPCHK(hipHostMalloc((void**)&hs, 2 * sizeof(RefineSlot), hipHostMallocDefault));
    hipEvent_t ev[2] = {nullptr, nullptr};
    int rc = H8_OK;
    for (int i = 0; i < 2 && rc == H8_OK; ++i)
        if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess)
            rc = fail(H8_ERR_HIP, "hipEventCreateWithFlags failed");
    if (rc == H8_OK) rc = refine_resident_phases(d, s_e32, K032, w, zr, u, R, s, hst, launches, hs, ev);
    (void)hipStreamSynchronize(s);       // the phase queued last (and, on an error, all queued work)
    for (int i = 0; i < 2; ++i)
        if (ev[i]) (void)hipEventDestroy(ev[i]);
    (void)hipHostFree(hs);
    return rc;
}

static int refine_solve(const h8_grid* g, const double* s_e64, const double* K064, const float* s_e32,
                        const float* K032, const double* dinv64, const double* f, double* u, double tol,
                        int64_t maxiter, double delta, void* ws, size_t ws_bytes, h8_pcg_stats* stats, void* stream,
                        bool warm) {
    GridDev d;
    int rc = make_grid(g, d);
    if (rc) return rc;
    if (!s_e64 || !K064 || !s_e32 || !K032 || !dinv64 || !f || !u || !ws)
        return fail(H8_ERR_ARG, "NULL pointer");
    if (ws_bytes < ws_bytes_ref(d))
        return fail(H8_ERR_WORKSPACE, "workspace too small: %zu < %zu", ws_bytes, ws_bytes_ref(d));
    if (!(tol > 0.0)) return fail(H8_ERR_ARG, "tol must be > 0");
    if (!(delta > 0.0 && delta < 1.0)) return fail(H8_ERR_ARG, "delta must be in (0, 1)");
    if (maxiter < 0) return fail(H8_ERR_ARG, "maxiter must be >= 0");
    hipStream_t s = (hipStream_t)stream;
    char* p = (char*)ws;
    WsRef w;
    w.r = (float*)carve(p, sizeof(float) * d.ndof);
    w.x = (float*)carve(p, sizeof(float) * d.ndof);
    w.p0 = (float*)carve(p, sizeof(float) * d.ndof);
    w.p1 = (float*)carve(p, sizeof(float) * d.ndof);
    w.q = (float*)carve(p, sizeof(float) * d.ndof);
    w.dinv32 = (float*)carve(p, sizeof(float) * d.ndof);
    w.partials = (double*)carve(p, sizeof(double) * partial_slots(d));
    w.st = (PcgState*)carve(p, sizeof(PcgState));
    const ZRing<float> zr = carve_ring<float>(p, d, w.q);

    Ev e0, e1;
    HIPCHK(e0.create());
    HIPCHK(e1.create());
    HIPCHK(hipEventRecord(e0, s));
    const int sb = stream_blocks(d.ndof);
    hipLaunchKernelGGL(k_cast_f64_f32, dim3(sb), dim3(STREAM_THREADS), 0, s, d.ndof, dinv64, w.dinv32);
    LAUNCHCHK();
    if (warm)      // the initial guess, zero at the fixed DOFs (dinv64 = 0 there)
        hipLaunchKernelGGL(k_masked_sub, dim3(sb), dim3(STREAM_THREADS), 0, s, d.ndof, (const double*)u,
                           (const double*)nullptr, dinv64, u);
    else
        hipLaunchKernelGGL(k_fill<double>, dim3(sb), dim3(STREAM_THREADS), 0, s, d.ndof, u, 0.0);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_fill<float>, dim3(sb), dim3(STREAM_THREADS), 0, s, d.ndof, w.x, 0.0f);
    LAUNCHCHK();

    // ||f|| over the free DOFs and the early-out test: the init kernel on a
    // zero guess, with p0|p1 (two adjacent float vectors) as FP64 scratch.
    PcgState hst{};
    hst.tol = tol;
    hst.maxiter = maxiter;
    hst.delta = delta;
    hst.best_true = INFINITY;
    hst.status = ST_RUNNING;
    HIPCHK(hipMemcpyAsync(w.st, &hst, sizeof(PcgState), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_pcg_init<double>, dim3(sb), dim3(STREAM_THREADS), 0, s, d.ndof, f,
                       (const double*)nullptr, dinv64, (double*)w.p0, w.st, w.partials);
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(&hst, w.st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (warm && hst.status == ST_DONE) {
        // the early out above assumes a zero guess (||f|| below tol): with an
        // initial guess the first true residual decides
        hst.status = ST_RUNNING;
        HIPCHK(hipMemcpyAsync(w.st, &hst, sizeof(PcgState), hipMemcpyHostToDevice, s));
    }
    if (hst.status == ST_RUNNING && maxiter == 0) {
        hst.status = ST_MAXITER;
        hst.iters = 0;
    }
    NodeArgs<double> R{};
    R.g = d; R.s_e = s_e64; R.K0 = K064; R.u = u; R.dinv = dinv64; R.f = f; R.r32 = w.r;
    R.dinv32 = w.dinv32; R.st = w.st; R.partials = w.partials;
    long long launches = 0;
    long long resid_calls = 0;
    hst.true_rel = hst.rel;
    if (hst.status == ST_RUNNING) {
        // first residual (u = 0): r32 = float(mask f), rz_old, rmax
        rc = launch_op<double, MODE_RESID>(d, R, s);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(&hst, w.st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    GraphKey key{};
    key.kind = 2;
    key.variant = d.variant;
    key.layers = d.layers;
    key.arith = d.arith;       // FP32 or FP64 element products (a different kernel)
    HIPCHK(hipGetDevice(&key.dev));
    key.nx = d.nx; key.ny = d.ny; key.nz = d.nz;
    const void* ptrs[] = {s_e32, K032, w.dinv32, w.x, w.r, w.p0, w.p1, w.q, w.st, w.partials};
    std::memcpy(key.ptr, ptrs, sizeof ptrs);
    key.chunk = chunk_for(maxiter);
    if (hst.status == ST_RUNNING && use_resident<float>(d)) {
        rc = refine_resident(d, s_e32, K032, w, zr, u, R, s, hst, launches);
        if (rc && rc != RES_DECLINED) return rc;     // declined: the streaming loop below runs
    }
    while (hst.status == ST_RUNNING) {
        rc = run_chunks<float, true>(key, d, s_e32, K032, w.dinv32, w.x, w.r, w.p0, w.p1, w.q, zr, w.st,
                                     w.partials, s, hst, launches);
        if (rc) return rc;
        if (hst.status != ST_REPLACE) break;
        // fold the FP32 iterate (with the pending x += alpha p of the last
        // direction, p = p[k-1], the update kernel having advanced k) into u,
        // then replace the residual in FP64
        hipLaunchKernelGGL(k_fold, dim3(sb), dim3(STREAM_THREADS), 0, s, d.ndof, u, w.x, (const float*)w.p0,
                           (const float*)w.p1, (const PcgState*)w.st);
        LAUNCHCHK();
        rc = launch_op<double, MODE_RESID>(d, R, s);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(&hst, w.st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    resid_calls = hst.outer;      // residual evaluations (the first one and every replacement)
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    fill_stats(stats, hst, ms, launches);
    if (stats) {
        stats->outer = int(std::max(0LL, resid_calls - 1));
        stats->true_rel_residual = hst.true_rel;
        stats->inner_iterations = hst.iters;
        stats->bnorm = hst.bnorm;
    }
    if (hst.status == ST_NOT_SPD)
        return fail(H8_NOT_SPD, "preconditioned_cg: p^T A p = %.3e <= 0 at iteration %lld -- A is not SPD.",
                    hst.bad_pq, hst.iters);
    if (hst.status == ST_MAXITER && hst.stalled)
        return fail(H8_MAXITER,
                    "refined solve: the true relative residual stagnated at %.3e (tol=%.1e) after %lld FP32 "
                    "iterations -- FP32 iterations cannot reduce it further; use the FP64 solve or a larger tol.",
                    hst.true_rel, tol, (long long)hst.iters);
    if (hst.status == ST_MAXITER)
        return fail(H8_MAXITER,
                    "refined solve: did not converge in %lld iterations (true relative residual %.3e >= "
                    "tol=%.1e).",
                    (long long)maxiter, hst.true_rel, tol);
    return H8_OK;
}

// ---------------------------------------------------------------------------
// Multigrid-preconditioned refined solve
// ---------------------------------------------------------------------------
// The hierarchy of a grid: level 0 is the grid itself; every further level
// halves the axes whose element count is even, until a level has at most
// MG_COARSE_DOF DOFs or no axis can be halved.  Usable when there are at least
// two levels and the coarsest has at most MG_DENSE_MAX DOFs.
struct MgPlan {
    int nlev = 0;
    int deg0 = 2;                 // Chebyshev steps per level-0 smoothing pass
    MgLevel L[MG_MAX_LEVELS];
    MgConst* cst = nullptr;       // [MG_MAX_LEVELS]
    double* M = nullptr;          // the 8 child matrices of level 1, [8][576]
    unsigned* emask = nullptr;    // free-DOF bits of every fine element
    int* flag = nullptr;          // singular coarsest operator
    double* A = nullptr;          // coarsest level, dense (inverted in place)
    float* Ainv = nullptr;
};

static bool mg_plan(const GridDev& d, MgPlan& P) {
    int nx = d.nx, ny = d.ny, nz = d.nz;
    P.nlev = 0;
    for (;;) {
        MgLevel& L = P.L[P.nlev++];
        L = MgLevel{};
        L.nx = nx; L.ny = ny; L.nz = nz;
        L.NJ = ny + 1; L.NI = nx + 1; L.NK = nz + 1;
        L.nnodes = (long long)L.NJ * L.NI * L.NK;
        L.ndof = 3 * L.nnodes;
        const int rx = (nx % 2 == 0) ? 2 : 1, ry = (ny % 2 == 0) ? 2 : 1, rz = (nz % 2 == 0) ? 2 : 1;
        if (L.ndof <= MG_COARSE_DOF || (rx == 1 && ry == 1 && rz == 1) || P.nlev == MG_MAX_LEVELS) break;
        L.rx = rx; L.ry = ry; L.rz = rz;
        nx /= rx; ny /= ry; nz /= rz;
    }
    P.deg0 = d.mg_degree ? d.mg_degree : mg_deg0_default(d.ndof);
    // the MG kernels index stencil levels in 32 bits
    if (P.nlev >= 2 && 243LL * P.L[1].nnodes >= (1LL << 31)) return false;
    return P.nlev >= 2 && P.L[P.nlev - 1].ndof <= MG_DENSE_MAX;
}

// Carves the hierarchy's buffers from `base` (nullptr: sizes only) and
// returns the bytes needed.  Level 0 borrows the solve's z (its x1), r32 and
// dinv32 (set by the caller); its product scratch y, direction d and the
// other iterate x0 are carved here.
static size_t mg_layout(MgPlan& P, char* base) {
    size_t off = 0;
    auto take = [&](size_t bytes) -> char* {
        char* q = base ? base + off : nullptr;
        off += align_up(bytes);
        return q;
    };
    P.cst = (MgConst*)take(sizeof(MgConst) * MG_MAX_LEVELS);
    P.M = (double*)take(sizeof(double) * 8 * 576);
    P.emask = (unsigned*)take(sizeof(unsigned) * (size_t)P.L[0].nx * P.L[0].ny * P.L[0].nz);
    P.flag = (int*)take(sizeof(int));
    P.L[0].y = (float*)take(sizeof(float) * P.L[0].ndof);
    P.L[0].d = (float*)take(sizeof(float) * P.L[0].ndof);
    P.L[0].x0 = (float*)take(sizeof(float) * P.L[0].ndof);
    for (int l = 1; l < P.nlev; ++l) {
        MgLevel& L = P.L[l];
        const size_t nv = sizeof(float) * L.ndof;
        L.S = (float*)take(sizeof(float) * MG_STENCIL * L.nnodes);
        L.dinv = (float*)take(nv);
        L.x0 = (float*)take(nv);
        L.x1 = (float*)take(nv);
        L.b = (float*)take(nv);
        L.d = (float*)take(nv);
        L.y = (float*)take(nv);
    }
    const size_t nc = P.L[P.nlev - 1].ndof;
    P.A = (double*)take(sizeof(double) * nc * nc);
    P.Ainv = (float*)take(sizeof(float) * nc * nc);
    return off;
}

static unsigned mg_blocks(long long n) { return unsigned(std::max(1LL, std::min((n + 255) / 256, 4096LL))); }

extern "C" size_t h8_mg_workspace_bytes(const h8_grid* g) {
    GridDev d;
    if (make_grid(g, d)) return 0;
    MgPlan P;
    if (!mg_plan(d, P)) return 0;
    return mg_layout(P, nullptr);
}

// First level of the cycle that runs inside k_mg_bottom: the trailing levels
// of at most MG_BOTTOM_NODES nodes (never level 0, the matrix-free grid).
static int mg_bottom_level(const MgPlan& P) {
    int lb = P.nlev - 1;
    while (lb > 1 && P.L[lb - 1].nnodes <= MG_BOTTOM_NODES) --lb;
    return lb;
}

// Arguments of k_mg_bottom (the LDS layout of its staged form, or none
// when the small levels do not fit).
static MgBottom mg_bottom_args(const MgPlan& P) {
    MgBottom B{};
    for (int l = 0; l < P.nlev; ++l) B.L[l] = P.L[l];
    B.lb = mg_bottom_level(P);
    B.nl = P.nlev;
    B.cst = P.cst;
    B.Ainv = P.Ainv;
    const long long w = mg_bottom_layout(B.L, B.lb, B.nl, B.lds_off);
    B.lds_words = (w * (long long)sizeof(float) <= MG_BOTTOM_LDS_BYTES) ? int(w) : 0;
    return B;
}

// Builds the hierarchy for the current element scales and free mask (stream
// ordered; *P.flag tells whether the coarsest operator was singular).
static int mg_setup(const GridDev& d, MgPlan& P, const double* s_e64, const double* K064, const float* s_e32,
                    hipStream_t s) {
    const int nl = P.nlev;
    const MgLevel& L0 = P.L[0];
    if (mg_bottom_args(P).lds_words)
        HIPCHK(hipFuncSetAttribute((const void*)k_mg_bottom<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   MG_BOTTOM_LDS_BYTES));
    static_assert(sizeof(MgConst) % sizeof(unsigned) == 0, "MgConst is cleared as words");
    hipLaunchKernelGGL(k_fill<unsigned>, dim3(1), dim3(64), 0, s,
                       (long long)(sizeof(MgConst) / sizeof(unsigned) * MG_MAX_LEVELS),
                       reinterpret_cast<unsigned*>(P.cst), 0u);
    hipLaunchKernelGGL(k_mg_mc, dim3(18), dim3(256), 0, s, K064, L0.rx, L0.ry, L0.rz, P.M);
    const long long m = (long long)d.nx * d.ny * d.nz;
    hipLaunchKernelGGL(k_mg_elem_mask, dim3(mg_blocks(m)), dim3(256), 0, s, d, (const float*)L0.dinv, P.emask);
    hipLaunchKernelGGL(k_mg_gersh0, dim3(mg_blocks(d.nnodes)), dim3(256), 0, s, d, s_e32, K064,
                       (const float*)L0.dinv, P.cst);
    {
        // a thread per (coarse node, neighbour block) pair of level 1
        const MgLevel& L1 = P.L[1];
        const long long pairs = 27LL * L1.nnodes;
        const long long nblk = std::min<long long>((pairs + 255) / 256, 16384);
        hipLaunchKernelGGL(k_mg_level1, dim3(unsigned(nblk)), dim3(256), 0, s, d, L1, L0.rx, L0.ry, L0.rz, s_e64,
                           K064, (const double*)P.M, (const unsigned*)P.emask);
    }
    hipLaunchKernelGGL(k_mg_diag, dim3(mg_blocks(P.L[1].ndof)), dim3(256), 0, s, P.L[1], P.cst + 1);
    for (int l = 1; l + 1 < nl; ++l) {
        const MgLevel& F = P.L[l];
        // lanes per RAP entry: about 2^20 threads per launch, 1..32
        const long long entries = 243 * P.L[l + 1].nnodes;
        int lg = 0;
        while (lg < 5 && (entries << (lg + 1)) <= (1LL << 20)) ++lg;
        if (lg == 0) {
            // a large level: one entry per block over a chunk of coarse nodes, XCD-local chunks
            const long long nblk = 243 * ((P.L[l + 1].nnodes + MG_RAP_CHUNK - 1) / MG_RAP_CHUNK);
            hipLaunchKernelGGL(k_mg_rap_chunked, dim3(unsigned(nblk)), dim3(MG_RAP_CHUNK), 0, s, F, P.L[l + 1], F.rx,
                               F.ry, F.rz);
        } else {
            hipLaunchKernelGGL(k_mg_rap, dim3(mg_blocks(entries << lg)), dim3(256), 0, s, F, P.L[l + 1], F.rx, F.ry,
                               F.rz, lg);
        }
        hipLaunchKernelGGL(k_mg_diag, dim3(mg_blocks(P.L[l + 1].ndof)), dim3(256), 0, s, P.L[l + 1],
                           P.cst + l + 1);
    }
    const MgLevel& Lc = P.L[nl - 1];
    const long long nc = Lc.ndof;
    hipLaunchKernelGGL(k_fill<double>, dim3(mg_blocks(nc * nc)), dim3(256), 0, s, nc * nc, P.A, 0.0);
    hipLaunchKernelGGL(k_mg_dense, dim3(mg_blocks(27 * nc)), dim3(256), 0, s, Lc, P.A);
    if (nc <= MG_LDS_GJ) {
        // the register form: row/column buffers of both parities and the diagonal
        const unsigned lds = unsigned(5 * nc * sizeof(double));
        const long long per = (nc * nc + 1023) / 1024;      // register slots per thread
        if (per <= 4)
            hipLaunchKernelGGL((k_mg_gauss_jordan<true, 4>), dim3(1), dim3(1024), lds, s, P.A, int(nc), P.Ainv, P.flag);
        else if (per <= 8)
            hipLaunchKernelGGL((k_mg_gauss_jordan<true, 8>), dim3(1), dim3(1024), lds, s, P.A, int(nc), P.Ainv, P.flag);
        else
            hipLaunchKernelGGL((k_mg_gauss_jordan<true, 16>), dim3(1), dim3(1024), lds, s, P.A, int(nc), P.Ainv,
                               P.flag);
    } else if (nc <= MG_SWEEP_MAX) {
        // the symmetric sweep on the upper triangle in registers (a 3 x 3 x 3 coarsest grid)
        static_assert(MG_SWEEP_MAX * (MG_SWEEP_MAX + 1) / 2 <= 1024 * 19, "sweep slots");
        hipLaunchKernelGGL(k_mg_sweep_sym<19>, dim3(1), dim3(1024), unsigned(3 * nc * sizeof(double)), s,
                           (const double*)P.A, int(nc), P.Ainv, P.flag);
    } else {
        hipLaunchKernelGGL(k_mg_gauss_jordan<false>, dim3(1), dim3(1024), unsigned(2 * nc * sizeof(double)), s, P.A,
                           int(nc), P.Ainv, P.flag);
    }
    hipLaunchKernelGGL(k_mg_consts, dim3(1), dim3(64), 0, s, P.cst, nl);
    LAUNCHCHK();
    return H8_OK;
}

// The small levels' kernel: staged in LDS when they fit.
static int launch_bottom(const MgPlan& P, const PcgState* st, hipStream_t s) {
    const MgBottom B = mg_bottom_args(P);
    if (B.lds_words) {
        // (the kernel's dynamic-LDS limit is raised by mg_setup, outside any capture)
        const unsigned lds = unsigned(B.lds_words) * sizeof(float);
        hipLaunchKernelGGL(k_mg_bottom<true>, dim3(1), dim3(MG_BOTTOM_THREADS), lds, s, B, st);
    } else {
        hipLaunchKernelGGL(k_mg_bottom<false>, dim3(1), dim3(MG_BOTTOM_THREADS), 0, s, B, st);
    }
    LAUNCHCHK();
    return H8_OK;
}

// Stencil-level sweeps: the quad form on small levels, the row form on large ones.
static void launch_cheb_st(const MgLevel& L, const float* xin, float* xout, const MgConst* cst, int step,
                           const PcgState* st, hipStream_t s) {
    if (L.ndof <= MG_QUAD_DOF)
        hipLaunchKernelGGL(k_mg_cheb_st<true>, dim3(mg_blocks(4 * L.ndof)), dim3(256), 0, s, L, xin, xout, cst, step,
                           st);
    else
        hipLaunchKernelGGL(k_mg_cheb_st<false>, dim3(mg_blocks(L.ndof)), dim3(256), 0, s, L, xin, xout, cst, step,
                           st);
}
static void launch_resid_st(const MgLevel& L, const float* x, float* r, const PcgState* st, hipStream_t s) {
    if (L.ndof <= MG_QUAD_DOF)
        hipLaunchKernelGGL(k_mg_resid_st<true>, dim3(mg_blocks(4 * L.ndof)), dim3(256), 0, s, L, x, r, st);
    else
        hipLaunchKernelGGL(k_mg_resid_st<false>, dim3(mg_blocks(L.ndof)), dim3(256), 0, s, L, x, r, st);
}

// z = V(r): one V-cycle (degree-2 Chebyshev smoothing before and after the
// coarse correction).  P.L[0] holds b = r32, x1 = z, dinv = dinv32; prod0 the
// level-0 operator arguments (FP32 sector kernel; st and partials of the
// solve).  Level 0 smooths with the fused sector Chebyshev step; a stencil
// level gets its first step from the restriction into it; the small levels
// run in one k_mg_bottom launch.  fuse_rz: the last step also reduces r.z and
// sets beta (the multigrid iteration's k_mg_rz).
static int mg_vcycle(const GridDev& d32, const MgPlan& P, const NodeArgs<float>& prod0, const PcgState* st,
                     hipStream_t s, bool fuse_rz) {
    const int nl = P.nlev, lb = mg_bottom_level(P);
    const MgLevel* L = P.L;
    const MgConst* cst = P.cst;
    const MgLevel& L0 = L[0];
    auto resid0 = [&](const float* u) {          // L0.y = M (b - K u)
        NodeArgs<float> A = prod0;
        A.u = u;
        A.y = L0.y;
        A.cb_b = L0.b;
        A.dinv = L0.dinv;
        return launch_op<float, MODE_MGRES>(d32, A, s);
    };
    auto first0 = [&](float* out) {              // x0 = c0 dinv b, then one step: out = x0 + d
        NodeArgs<float> A = prod0;
        A.dinv = L0.dinv;
        A.cheb = &cst[0].c0;
        A.cb_b = L0.b;
        A.cb_d = L0.d;
        A.cb_out = out;
        return launch_op<float, MODE_CHEB0>(d32, A, s);
    };
    auto cheb0 = [&](const float* u, float* out, int step, int rz) {
        NodeArgs<float> A = prod0;
        A.u = u;
        A.dinv = L0.dinv;
        A.cheb = &cst[0].c0;
        A.cb_b = L0.b;
        A.cb_d = L0.d;
        A.cb_out = out;
        A.cheb_step = step;
        A.cheb_rz = rz;
        return launch_op<float, MODE_CHEB>(d32, A, s);
    };
    int rc;
    // Level 0 smooths deg0 steps per pass.  Pre: the first two in one pass
    // (into x1), steps 2.. alternating between x1 and x0; post: deg0 steps
    // from the corrected iterate.  2 deg0 - 2 swaps in all, so the result
    // lands in x1 whatever the degree.
    const int deg0 = P.deg0;
    float* cur0 = L0.x1;
    auto other0 = [&](float* v) { return v == L0.x1 ? L0.x0 : L0.x1; };
    for (int l = 0; l < lb; ++l) {
        const MgLevel& F = L[l];
        const MgLevel& C = L[l + 1];
        if (l == 0) {
            // x0 = d = c0 dinv b and x1 = x0 + d' in one pass; y = M (b - K x)
            if ((rc = first0(F.x1))) return rc;
            for (int k = 2; k < deg0; ++k) {
                if ((rc = cheb0(cur0, other0(cur0), k, 0))) return rc;
                cur0 = other0(cur0);
            }
            if ((rc = resid0(cur0))) return rc;
        } else {
            launch_cheb_st(F, F.x0, F.x1, cst + l, 1, st, s);
            launch_resid_st(F, F.x1, F.y, st, s);
        }
        hipLaunchKernelGGL(k_mg_restrict, dim3(mg_blocks(C.nnodes)), dim3(256), 0, s, F, C, F.rx, F.ry, F.rz,
                           (const float*)F.y, cst + l + 1, (l + 1 < lb) ? 1 : 0, st);
    }
    if ((rc = launch_bottom(P, st, s))) return rc;
    for (int l = lb - 1; l >= 0; --l) {
        const MgLevel& F = L[l];
        const MgLevel& C = L[l + 1];
        hipLaunchKernelGGL(k_mg_prolong, dim3(mg_blocks(F.nnodes)), dim3(256), 0, s, F, C, F.rx, F.ry, F.rz,
                           l == 0 ? cur0 : F.x1, (const float*)C.x1, (const float*)F.dinv, st);
        if (l == 0) {
            for (int k = 0; k < deg0; ++k) {
                if ((rc = cheb0(cur0, other0(cur0), k, (fuse_rz && k == deg0 - 1) ? 1 : 0))) return rc;
                cur0 = other0(cur0);
            }
        } else {
            launch_cheb_st(F, F.x1, F.x0, cst + l, 0, st, s);
            launch_cheb_st(F, F.x0, F.x1, cst + l, 1, st, s);
        }
    }
    LAUNCHCHK();
    return H8_OK;
}

// The refined solve of refine_solve with z = V(r) in the FP32 iterations:
// per iteration p = z + beta p (sector kernel, MODE_DIRZ), q = K p, alpha;
// x += alpha p, r -= alpha q and the replacement test; on a replacement the
// FP64 residual at u + x (a no-op otherwise; the next update kernel folds x
// into u, and one fold follows the loop); z = V(r); r.z -> beta.  All of it is
// captured in graphs of MG_CHUNK (or, near the end, 4 or 2) iterations.
constexpr int MG_CHUNK = 8;

static int refine_solve_mg(const h8_grid* g, const double* s_e64, const double* K064, const float* s_e32,
                           const float* K032, const double* dinv64, const double* f, double* u, double tol,
                           int64_t maxiter, double delta, int warm, void* ws, size_t ws_bytes, void* mg_ws,
                           size_t mg_ws_bytes, h8_pcg_stats* stats, void* stream) {
    GridDev d;
    int rc = make_grid(g, d);
    if (rc) return rc;
    if (!s_e64 || !K064 || !s_e32 || !K032 || !dinv64 || !f || !u || !ws || !mg_ws)
        return fail(H8_ERR_ARG, "NULL pointer");
    if (ws_bytes < ws_bytes_ref(d))
        return fail(H8_ERR_WORKSPACE, "workspace too small: %zu < %zu", ws_bytes, ws_bytes_ref(d));
    if (!(tol > 0.0)) return fail(H8_ERR_ARG, "tol must be > 0");
    if (!(delta > 0.0 && delta < 1.0)) return fail(H8_ERR_ARG, "delta must be in (0, 1)");
    if (maxiter < 0) return fail(H8_ERR_ARG, "maxiter must be >= 0");
    MgPlan P;
    if (!mg_plan(d, P))
        return fail(H8_ERR_ARG, "multigrid: a %dx%dx%d grid does not coarsen to at most %lld DOFs by halving its even "
                    "axes", d.nx, d.ny, d.nz, (long long)MG_DENSE_MAX);
    const size_t need = mg_layout(P, nullptr);
    if (mg_ws_bytes < need) return fail(H8_ERR_WORKSPACE, "multigrid workspace too small: %zu < %zu", mg_ws_bytes, need);
    mg_layout(P, (char*)mg_ws);
    hipStream_t s = (hipStream_t)stream;
    char* p = (char*)ws;
    WsRef w;
    w.r = (float*)carve(p, sizeof(float) * d.ndof);
    w.x = (float*)carve(p, sizeof(float) * d.ndof);
    w.p0 = (float*)carve(p, sizeof(float) * d.ndof);
    w.p1 = (float*)carve(p, sizeof(float) * d.ndof);
    w.q = (float*)carve(p, sizeof(float) * d.ndof);
    w.dinv32 = (float*)carve(p, sizeof(float) * d.ndof);
    w.partials = (double*)carve(p, sizeof(double) * partial_slots(d));
    w.st = (PcgState*)carve(p, sizeof(PcgState));
    float* z = (float*)carve(p, sizeof(float) * d.ndof);      // the first z-exchange vector of the resident solver
    (void)carve(p, sizeof(float) * d.ndof);                      // the second (unused here)
    // f and the FP64 solution live in the workspace during the solve, so that
    // the captured graphs hold no per-call pointers (a caller that allocates
    // them afresh every solve, as a SIMP loop does, would otherwise capture
    // every graph again); u is written once at the end
    double* f_ws = (double*)carve(p, sizeof(double) * d.ndof);
    double* u_ws = (double*)carve(p, sizeof(double) * d.ndof);
    double* const u_out = u;
    MgLevel& L0 = P.L[0];
    L0.b = w.r;
    L0.x1 = z;
    L0.dinv = w.dinv32;
    // FP32 products on the sector kernel (the direction mode needs it)
    GridDev d32 = d;
    if (d32.variant != H8_VARIANT_SECTOR && d32.variant != H8_VARIANT_SECTOR_NARROW) {
        d32.variant = H8_VARIANT_SECTOR;
        d32.layers = 0;
    }

    Ev e0, e1;
    HIPCHK(e0.create());
    HIPCHK(e1.create());
    HIPCHK(hipEventRecord(e0, s));
    const int sb = stream_blocks(d.ndof);
    hipLaunchKernelGGL(k_cast_f64_f32, dim3(sb), dim3(STREAM_THREADS), 0, s, d.ndof, dinv64, w.dinv32);
    HIPCHK(hipMemcpyAsync(f_ws, f, sizeof(double) * d.ndof, hipMemcpyDeviceToDevice, s));
    f = f_ws;
    if (warm)
        hipLaunchKernelGGL(k_masked_sub, dim3(sb), dim3(STREAM_THREADS), 0, s, d.ndof, (const double*)u_out,
                           (const double*)nullptr, dinv64, u_ws);
    else
        hipLaunchKernelGGL(k_fill<double>, dim3(sb), dim3(STREAM_THREADS), 0, s, d.ndof, u_ws, 0.0);
    u = u_ws;
    hipLaunchKernelGGL(k_fill<float>, dim3(sb), dim3(STREAM_THREADS), 0, s, d.ndof, w.x, 0.0f);
    LAUNCHCHK();
    rc = mg_setup(d, P, s_e64, K064, s_e32, s);
    if (rc) return rc;
    PcgState hst{};
    hst.tol = tol;
    hst.maxiter = maxiter;
    hst.delta = delta;
    hst.best_true = INFINITY;
    hst.status = ST_RUNNING;
    hst.mg = 1;
    HIPCHK(hipMemcpyAsync(w.st, &hst, sizeof(PcgState), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_pcg_init<double>, dim3(sb), dim3(STREAM_THREADS), 0, s, d.ndof, f,
                       (const double*)nullptr, dinv64, (double*)w.p0, w.st, w.partials);
    LAUNCHCHK();
    int flag = 0;
    HIPCHK(hipMemcpyAsync(&hst, w.st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&flag, P.flag, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (flag) {
        // (u holds the masked initial guess, as documented)
        HIPCHK(hipMemcpyAsync(u_out, u_ws, sizeof(double) * d.ndof, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        return fail(H8_MG_DECLINED, "multigrid: the coarsest operator is singular (the supports leave rigid-body "
                    "motions free); use the Jacobi preconditioner");
    }
    if (warm && hst.status == ST_DONE) {
        hst.status = ST_RUNNING;
        HIPCHK(hipMemcpyAsync(w.st, &hst, sizeof(PcgState), hipMemcpyHostToDevice, s));
    }
    if (hst.status == ST_RUNNING && maxiter == 0) {
        hst.status = ST_MAXITER;
        hst.iters = 0;
    }
    // The FP64 residual runs every iteration and returns at once unless a
    // replacement is due; the sector kernel's no-op launch is the cheapest
    // (the node kernel's 8x8x4 bricks launch about 1.7x as many workgroups),
    // so it is used whatever the caller's FP64 variant.
    GridDev d64 = d;
    if (d64.variant != H8_VARIANT_SECTOR && d64.variant != H8_VARIANT_SECTOR_NARROW) {
        d64.variant = H8_VARIANT_SECTOR;
        d64.layers = 0;
    }
    NodeArgs<double> R{};
    R.g = d64; R.s_e = s_e64; R.K0 = K064; R.u = u; R.dinv = dinv64; R.f = f; R.r32 = w.r;
    R.dinv32 = w.dinv32; R.st = w.st; R.partials = w.partials;
    R.x32 = w.x;     // the FP64 residual is formed at u + x (x is folded by the next update)
    NodeArgs<float> prod0{};
    prod0.g = d32; prod0.s_e = s_e32; prod0.K0 = K032; prod0.u = z; prod0.y = L0.y; prod0.st = w.st;
    prod0.partials = w.partials;
    NodeArgs<float> D{};
    D.g = d32; D.s_e = s_e32; D.K0 = K032; D.r = z; D.dinv = w.dinv32; D.p0 = w.p0; D.p1 = w.p1; D.q = w.q;
    D.x = w.x; D.st = w.st; D.partials = w.partials;
    long long launches = 0;
    hst.true_rel = hst.rel;
    if (hst.status == ST_RUNNING) {
        GraphKey key{};
        key.kind = 4;
        key.variant = d32.variant;
        key.layers = d32.layers;
        key.arith = d.arith;
        HIPCHK(hipGetDevice(&key.dev));
        key.nx = d.nx; key.ny = d.ny; key.nz = d.nz;
        const void* ptrs[] = {s_e64, K064, s_e32, K032, dinv64, ws, mg_ws};   // (f and u are in ws)
        std::memcpy(key.ptr, ptrs, sizeof ptrs);
        key.extra[0] = (unsigned long long)d64.variant;   // the FP64 residual's kernel
        key.extra[1] = (unsigned long long)d64.layers;
        key.extra[3] = (unsigned long long)P.deg0;         // the level-0 steps of every captured cycle
        NodeArgs<double> Rq = R;
        Rq.only_on_replace = 1;
        NodeArgs<double> R0 = R;
        R0.only_on_replace = 2;
        // one captured graph per chunk length (8, 4, 2 iterations): the solve
        // launches full chunks while far from the tolerance and shorter ones at
        // the end, where the iterations of a full chunk past the stop would be
        // launches that return at once (about 3 us each, 29 per iteration).  The
        // first chunk also holds the first residual and z = V(r), r.z (otherwise
        // some 30 launches made one by one from the host)
        auto graph_of = [&](int chunk, bool first, GraphRef* exec) -> int {
            key.chunk = chunk;
            key.extra[2] = first ? 1 : 0;
            return get_graph(key, [&](hipStream_t cs) -> int {
                if (first) {
                    int r0 = launch_op<double, MODE_RESID>(d64, R0, cs);
                    if (r0) return r0;
                    if ((r0 = mg_vcycle(d32, P, prod0, w.st, cs, true))) return r0;     // with r.z and beta
                }
                for (int it = 0; it < chunk; ++it) {
                    int r1 = launch_op<float, MODE_DIRZ>(d32, D, cs);
                    if (r1) return r1;
                    // (after a replacement the update folds x into u first)
                    hipLaunchKernelGGL((k_pcg_update<float, true, true, true>), dim3(update_blocks<float>(d.ndof)),
                                       dim3(STREAM_THREADS), 0, cs, d.ndof, w.x, w.r, (const float*)w.p0,
                                       (const float*)w.p1, (const float*)w.q, (const float*)w.dinv32, w.st,
                                       w.partials, u);
                    if ((r1 = launch_op<double, MODE_RESID>(d64, Rq, cs))) return r1;
                    if ((r1 = mg_vcycle(d32, P, prod0, w.st, cs, true))) return r1;
                }
                LAUNCHCHK();
                return H8_OK;
            }, exec);
        };
        int chunk = int(std::min<long long>(MG_CHUNK, maxiter));
        double rel_prev = hst.rel;
        long long k_prev = hst.k;
        for (bool first = true;; first = false) {
            GraphRef exec;
            rc = graph_of(chunk, first, &exec);
            if (rc) return rc;
            HIPCHK(hipGraphLaunch(exec->exec, s));
            ++launches;
            HIPCHK(hipMemcpyAsync(&hst, w.st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            if (hst.status != ST_RUNNING) break;
            // iterations still needed, from the decay of the recurrence residual
            // over the last chunk (a replacement inside it can make the estimate
            // short: the next chunk is then launched as well)
            const double its = double(hst.k - k_prev);
            int next = MG_CHUNK;
            if (its > 0 && hst.rel > 0.0 && hst.rel < rel_prev && rel_prev > 0.0) {
                const double rate = std::log(hst.rel / rel_prev) / its;       // < 0
                const double need = std::log(tol / hst.rel) / rate + 1.0;     // + the replacement's iteration
                next = need <= 2.0 ? 2 : (need <= 4.0 ? 4 : MG_CHUNK);
            }
            chunk = int(std::min<long long>(next, maxiter));
            rel_prev = hst.rel;
            k_prev = hst.k;
        }
    }
    // the FP32 iterate not yet taken into u (the last replacement's, or the
    // progress since it when the solve stopped otherwise)
    hipLaunchKernelGGL(k_fold_final, dim3(sb), dim3(STREAM_THREADS), 0, s, d.ndof, (const double*)u_ws, w.x,
                       u_out);
    LAUNCHCHK();
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    fill_stats(stats, hst, ms, launches);
    if (stats) {
        stats->outer = int(std::max(0LL, hst.outer - 1));
        stats->true_rel_residual = hst.true_rel;
        stats->inner_iterations = hst.iters;
        stats->bnorm = hst.bnorm;
    }
    if (hst.status == ST_NOT_SPD)
        return fail(H8_NOT_SPD, "preconditioned_cg: p^T A p = %.3e <= 0 at iteration %lld -- A is not SPD.",
                    hst.bad_pq, hst.iters);
    if (hst.status == ST_MAXITER && hst.stalled)
        return fail(H8_MAXITER,
                    "refined solve: the true relative residual stagnated at %.3e (tol=%.1e) after %lld FP32 "
                    "iterations -- FP32 iterations cannot reduce it further; use the FP64 solve or a larger tol.",
                    hst.true_rel, tol, (long long)hst.iters);
    if (hst.status == ST_MAXITER)
        return fail(H8_MAXITER,
                    "refined solve: did not converge in %lld iterations (true relative residual %.3e >= "
                    "tol=%.1e).",
                    (long long)maxiter, hst.true_rel, tol);
    return H8_OK;
}

// One V-cycle z = V(r) after building the hierarchy (diagnostics: the
// preconditioner of h8_refine_solve_mg as an operator).
extern "C" int h8_mg_vcycle_f32(const h8_grid* g, const double* s_e64, const double* K064, const float* s_e32,
                                const float* K032, const double* dinv64, const float* r, float* z, void* ws,
                                size_t ws_bytes, void* mg_ws, size_t mg_ws_bytes, void* stream) {
    GridDev d;
    int rc = make_grid(g, d);
    if (rc) return rc;
    if (!s_e64 || !K064 || !s_e32 || !K032 || !dinv64 || !r || !z || !ws || !mg_ws)
        return fail(H8_ERR_ARG, "NULL pointer");
    if (ws_bytes < ws_bytes_ref(d))
        return fail(H8_ERR_WORKSPACE, "workspace too small: %zu < %zu", ws_bytes, ws_bytes_ref(d));
    MgPlan P;
    if (!mg_plan(d, P)) return fail(H8_ERR_ARG, "multigrid: the grid does not coarsen far enough");
    const size_t need = mg_layout(P, nullptr);
    if (mg_ws_bytes < need) return fail(H8_ERR_WORKSPACE, "multigrid workspace too small: %zu < %zu", mg_ws_bytes, need);
    mg_layout(P, (char*)mg_ws);
    hipStream_t s = (hipStream_t)stream;
    char* p = (char*)ws;
    (void)carve(p, sizeof(float) * d.ndof * 5);
    float* dinv32 = (float*)carve(p, sizeof(float) * d.ndof);
    hipLaunchKernelGGL(k_cast_f64_f32, dim3(stream_blocks(d.ndof)), dim3(STREAM_THREADS), 0, s, d.ndof, dinv64, dinv32);
    LAUNCHCHK();
    MgLevel& L0 = P.L[0];
    L0.b = const_cast<float*>(r);
    L0.x1 = z;
    L0.dinv = dinv32;
    GridDev d32 = d;
    d32.variant = H8_VARIANT_SECTOR;
    d32.layers = 0;
    if ((rc = mg_setup(d, P, s_e64, K064, s_e32, s))) return rc;
    int flag = 0;
    HIPCHK(hipMemcpyAsync(&flag, P.flag, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (flag) return fail(H8_MG_DECLINED, "multigrid: the coarsest operator is singular");
    NodeArgs<float> prod0{};
    prod0.g = d32; prod0.s_e = s_e32; prod0.K0 = K032; prod0.u = z; prod0.y = L0.y;
    return mg_vcycle(d32, P, prod0, nullptr, s, false);
}

// Average device time (ms) of the multigrid kernels of one FP32 iteration,
// each graph-launched `reps` times alone between HIP events on `stream`:
// avg_ms[0] = the level-0 product y = K x (sector kernel, FP32), avg_ms[1] =
// one Chebyshev step on level 1 (stencil product fused with the update; level
// 1 is a stencil level when the hierarchy has more than two levels, else 0),
// avg_ms[2] = one whole V-cycle.  info[0..3] = levels, DOFs of level 1,
// nodes of level 1, Chebyshev steps per level-0 smoothing pass.  Measurement helper with no counterpart in the reference;
// scribbles on both workspaces.
extern "C" int h8_mg_time_kernels(const h8_grid* g, const double* s_e64, const double* K064, const float* s_e32,
                                  const float* K032, const double* dinv64, void* ws, size_t ws_bytes, void* mg_ws,
                                  size_t mg_ws_bytes, int reps, double* avg_ms, int64_t* info, void* stream) {
    GridDev d;
    int rc = make_grid(g, d);
    if (rc) return rc;
    if (!s_e64 || !K064 || !s_e32 || !K032 || !dinv64 || !ws || !mg_ws || !avg_ms || reps < 1)
        return fail(H8_ERR_ARG, "bad argument");
    if (ws_bytes < ws_bytes_ref(d)) return fail(H8_ERR_WORKSPACE, "workspace too small");
    MgPlan P;
    if (!mg_plan(d, P)) return fail(H8_ERR_ARG, "multigrid: the grid does not coarsen far enough");
    if (mg_ws_bytes < mg_layout(P, nullptr)) return fail(H8_ERR_WORKSPACE, "multigrid workspace too small");
    mg_layout(P, (char*)mg_ws);
    hipStream_t s = (hipStream_t)stream;
    char* p = (char*)ws;
    float* r = (float*)carve(p, sizeof(float) * d.ndof);
    float* z = (float*)carve(p, sizeof(float) * d.ndof);
    (void)carve(p, sizeof(float) * d.ndof * 3);
    float* dinv32 = (float*)carve(p, sizeof(float) * d.ndof);
    const int sb = stream_blocks(d.ndof);
    hipLaunchKernelGGL(k_cast_f64_f32, dim3(sb), dim3(STREAM_THREADS), 0, s, d.ndof, dinv64, dinv32);
    hipLaunchKernelGGL(k_fill<float>, dim3(sb), dim3(STREAM_THREADS), 0, s, d.ndof, r, 1.0f);
    hipLaunchKernelGGL(k_fill<float>, dim3(sb), dim3(STREAM_THREADS), 0, s, d.ndof, z, 0.0f);
    LAUNCHCHK();
    MgLevel& L0 = P.L[0];
    L0.b = r;
    L0.x1 = z;
    L0.dinv = dinv32;
    GridDev d32 = d;
    if (d32.variant != H8_VARIANT_SECTOR && d32.variant != H8_VARIANT_SECTOR_NARROW) {
        d32.variant = H8_VARIANT_SECTOR;
        d32.layers = 0;
    }
    if ((rc = mg_setup(d, P, s_e64, K064, s_e32, s))) return rc;
    const MgLevel& L1 = P.L[1];
    if (L1.nnodes > 0 && P.nlev > 2) {
        hipLaunchKernelGGL(k_fill<float>, dim3(mg_blocks(L1.ndof)), dim3(256), 0, s, L1.ndof, L1.x0, 1.0f);
        hipLaunchKernelGGL(k_fill<float>, dim3(mg_blocks(L1.ndof)), dim3(256), 0, s, L1.ndof, L1.b, 1.0f);
        hipLaunchKernelGGL(k_fill<float>, dim3(mg_blocks(L1.ndof)), dim3(256), 0, s, L1.ndof, L1.d, 0.0f);
    }
    NodeArgs<float> prod0{};
    prod0.g = d32; prod0.s_e = s_e32; prod0.K0 = K032; prod0.u = z; prod0.y = L0.y;
    const bool has1 = P.nlev > 2;           // level 1 is a smoothed stencil level
    const int lb = mg_bottom_level(P);
    // one kernel (or the whole cycle) per slot, `reps` launches per graph
    auto body = [&](int which, hipStream_t cs) -> int {
        NodeArgs<float> A = prod0;
        A.dinv = L0.dinv; A.cheb = &P.cst[0].c0; A.cb_b = L0.b; A.cb_d = L0.d;
        switch (which) {
        case 0: return launch_op<float, MODE_PLAIN>(d32, prod0, cs);
        case 1:
            launch_cheb_st(L1, L1.x0, L1.x1, P.cst + 1, 1, nullptr, cs);
            return H8_OK;
        case 2: return mg_vcycle(d32, P, prod0, nullptr, cs, false);
        case 3:
            A.u = L0.x0; A.cb_out = L0.x1; A.cheb_step = 1;
            return launch_op<float, MODE_CHEB>(d32, A, cs);
        case 4:
            A.cb_out = L0.x1;
            return launch_op<float, MODE_CHEB0>(d32, A, cs);
        case 5:
            A.u = L0.x1; A.y = L0.y;
            return launch_op<float, MODE_MGRES>(d32, A, cs);
        case 6:
            hipLaunchKernelGGL(k_mg_restrict, dim3(mg_blocks(L1.nnodes)), dim3(256), 0, cs, L0, L1, L0.rx, L0.ry,
                               L0.rz, (const float*)L0.y, (const MgConst*)(P.cst + 1), 1, (const PcgState*)nullptr);
            return H8_OK;
        case 7:
            hipLaunchKernelGGL(k_mg_prolong, dim3(mg_blocks(L0.nnodes)), dim3(256), 0, cs, L0, L1, L0.rx, L0.ry,
                               L0.rz, L0.x1, (const float*)L1.x1, (const float*)L0.dinv, (const PcgState*)nullptr);
            return H8_OK;
        case 8: return launch_bottom(P, nullptr, cs);
        default:
            launch_resid_st(L1, L1.x1, L1.y, nullptr, cs);
            return H8_OK;
        }
    };
    // slots that need level 1 as a smoothed level (1, 9) or a restriction from
    // level 0 into a smoothed level (6 with pre0)
    auto skip = [&](int which) { return (which == 1 || which == 9 || which == 6) && (!has1 || lb < 1); };
    Ev ev[2];
    for (auto& e : ev) HIPCHK(e.create());
    // the whole cycle first: one launch fills every level's vectors
    const int order[H8_MG_TIMES] = {2, 0, 1, 3, 4, 5, 6, 7, 8, 9};
    for (int o = 0; o < H8_MG_TIMES; ++o) {
        const int which = order[o];
        avg_ms[which] = 0.0;
        if (skip(which)) continue;
        GraphExec ex;
        rc = capture_graph([&](hipStream_t cs) -> int {
            for (int i = 0; i < reps; ++i)
                if (int r1 = body(which, cs)) return r1;
            LAUNCHCHK();
            return H8_OK;
        }, ex);
        if (rc) return rc;
        HIPCHK(hipGraphLaunch(ex.exec, s));    // warm
        HIPCHK(hipEventRecord(ev[0], s));
        HIPCHK(hipGraphLaunch(ex.exec, s));
        HIPCHK(hipEventRecord(ev[1], s));
        HIPCHK(hipEventSynchronize(ev[1]));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[1]));
        avg_ms[which] = ms / reps;
    }
    if (info) {
        info[0] = P.nlev;
        info[1] = has1 ? L1.ndof : 0;
        info[2] = has1 ? L1.nnodes : 0;
        info[3] = P.deg0;
    }
    return H8_OK;
}

extern "C" int h8_refine_solve_mg(const h8_grid* g, const double* s_e64, const double* K064, const float* s_e32,
                                  const float* K032, const double* dinv64, const double* f, double* u, double tol,
                                  int64_t maxiter, double delta, int use_x0, void* ws, size_t ws_bytes, void* mg_ws,
                                  size_t mg_ws_bytes, h8_pcg_stats* stats, void* stream) {
    return refine_solve_mg(g, s_e64, K064, s_e32, K032, dinv64, f, u, tol, maxiter, delta, use_x0 != 0, ws, ws_bytes,
                           mg_ws, mg_ws_bytes, stats, stream);
}

// ---------------------------------------------------------------------------
// Measurement helper: average device time of the two kernels of one PCG
// iteration (direction+product, update), each launched `reps` times alone.
// ---------------------------------------------------------------------------
template <typename T, bool REFINED>
static int time_kernels(const GridDev& d, const T* s_e, const T* K0, const T* dinv, T* x, T* r, T* p0, T* p1,
                        T* q, const ZRing<T>& zr, PcgState* st, double* partials, int reps, double* avg_ms,
                        hipStream_t s) {
    // r = 1 and a zero previous direction: every launch forms p = dinv*r (zero at
    // fixed DOFs), so p^T K p > 0 and no launch takes the early exit.
    const int sb = stream_blocks(d.ndof);
    hipLaunchKernelGGL(k_fill<T>, dim3(sb), dim3(STREAM_THREADS), 0, s, d.ndof, r, T(1));
    hipLaunchKernelGGL(k_fill<T>, dim3(sb), dim3(STREAM_THREADS), 0, s, d.ndof, p0, T(0));
    hipLaunchKernelGGL(k_fill<T>, dim3(sb), dim3(STREAM_THREADS), 0, s, d.ndof, p1, T(0));
    hipLaunchKernelGGL(k_fill<T>, dim3(sb), dim3(STREAM_THREADS), 0, s, d.ndof, x, T(0));
    LAUNCHCHK();
    PcgState h{};
    h.status = ST_RUNNING;
    h.k = 2;
    h.first = 0;
    h.beta = 0.0;
    h.alpha = 1e-9;
    h.rz_old = 1.0;
    h.bnorm = 1.0;
    h.tol = 0.0;
    h.maxiter = (1LL << 62);
    h.rmax = INFINITY;            // no replacement trigger fires while timing
    h.k_rep = (1LL << 62);
    h.pending = defer_x<T>(d) ? 1 : 0;   // as inside a solve of this size and kernel
    if (d.variant == H8_VARIANT_RESIDENT) {
        // the resident solver: `reps` iterations in one launch (stopped by
        // maxiter); avg_ms[0] = time per iteration, avg_ms[1] = 0
        h.maxiter = h.k + reps - 1;
        Ev ev[2];
        for (auto& e : ev) HIPCHK(e.create());
        PcgState hh{};
        long long launches = 0;
        float ms = 0.f;
        for (int rep = 0; rep < 2; ++rep) {      // warm-up, then timed
            HIPCHK(hipMemcpyAsync(st, &h, sizeof h, hipMemcpyHostToDevice, s));
            int rc = run_resident<T, REFINED>(d, s_e, K0, dinv, x, r, p0, p1, zr, st, partials, s, hh, launches,
                                              ev[0], ev[1]);
            if (rc) return rc;
            HIPCHK(hipEventSynchronize(ev[1]));
            HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[1]));
            if (hh.status == ST_NOT_SPD || hh.k - 2 + (hh.status == ST_REPLACE ? 0 : 1) < reps)
                return fail(H8_ERR_ARG, "resident timing run stopped early (status %d at iteration %lld)", hh.status,
                            hh.k);
        }
        avg_ms[0] = ms / reps;
        avg_ms[1] = 0.0;
        return H8_OK;
    }
    NodeArgs<T> A{};
    A.g = d; A.s_e = s_e; A.K0 = K0; A.r = r; A.dinv = dinv; A.p0 = p0; A.p1 = p1; A.q = q; A.x = x;
    A.st = st; A.partials = partials;
    // the launches are captured into graphs, as in the solves
    GraphExec ex[2];      // destroyed on every return
    for (int which = 0; which < 2; ++which) {
        const int rc = capture_graph([&](hipStream_t cs) -> int {
            for (int i = 0; i < reps; ++i) {
                if (which == 0) {
                    const int r1 = launch_dir<T>(d, A, cs);
                    if (r1) return r1;
                } else {
                    launch_update<T, REFINED>(d, x, r, p0, p1, q, dinv, st, partials, cs);
                }
            }
            LAUNCHCHK();
            return H8_OK;
        }, ex[which]);
        if (rc) return rc;
    }
    Ev ev[4];
    for (auto& e : ev) HIPCHK(e.create());
    double ms[2];
    for (int which = 0; which < 2; ++which) {
        HIPCHK(hipMemcpyAsync(st, &h, sizeof h, hipMemcpyHostToDevice, s));
        HIPCHK(hipGraphLaunch(ex[which].exec, s));   // warm
        HIPCHK(hipMemcpyAsync(st, &h, sizeof h, hipMemcpyHostToDevice, s));
        HIPCHK(hipEventRecord(ev[2 * which], s));
        HIPCHK(hipGraphLaunch(ex[which].exec, s));
        HIPCHK(hipEventRecord(ev[2 * which + 1], s));
        HIPCHK(hipEventSynchronize(ev[2 * which + 1]));
        float m = 0.f;
        HIPCHK(hipEventElapsedTime(&m, ev[2 * which], ev[2 * which + 1]));
        ms[which] = m;
        PcgState hh{};
        HIPCHK(hipMemcpy(&hh, st, sizeof hh, hipMemcpyDeviceToHost));
        if (hh.status != ST_RUNNING)
            return fail(H8_ERR_ARG, "timing loop %d left the running state (%d)", which, hh.status);
    }
    avg_ms[0] = ms[0] / reps;
    avg_ms[1] = ms[1] / reps;
    return H8_OK;
}


extern "C" int h8_refine_solve(const h8_grid* g, const double* s_e64, const double* K064,
                               const float* s_e32, const float* K032, const double* dinv64,
                               const double* f, double* u, double tol, int64_t maxiter, double delta,
                               void* ws, size_t ws_bytes, h8_pcg_stats* stats, void* stream) {
    return refine_solve(g, s_e64, K064, s_e32, K032, dinv64, f, u, tol, maxiter, delta, ws, ws_bytes, stats, stream,
                        false);
}

extern "C" int h8_refine_solve_x0(const h8_grid* g, const double* s_e64, const double* K064,
                                  const float* s_e32, const float* K032, const double* dinv64,
                                  const double* f, double* u, double tol, int64_t maxiter, double delta,
                                  void* ws, size_t ws_bytes, h8_pcg_stats* stats, void* stream) {
    return refine_solve(g, s_e64, K064, s_e32, K032, dinv64, f, u, tol, maxiter, delta, ws, ws_bytes, stats, stream,
                        true);
}

extern "C" int h8_time_kernels(const h8_grid* g, int precision, const void* s_e, const void* K0,
                               const double* dinv64, void* ws, size_t ws_bytes, int reps, double* avg_ms,
                               void* stream) {
    GridDev d;
    int rc = make_grid(g, d);
    if (rc) return rc;
    if (!s_e || !K0 || !dinv64 || !ws || !avg_ms || reps < 1) return fail(H8_ERR_ARG, "bad argument");
    hipStream_t s = (hipStream_t)stream;
    char* p = (char*)ws;
    if (precision == 64) {
        if (ws_bytes < ws_bytes_64(d)) return fail(H8_ERR_WORKSPACE, "workspace too small");
        double* r = (double*)carve(p, sizeof(double) * d.ndof);
        double* p0 = (double*)carve(p, sizeof(double) * d.ndof);
        double* p1 = (double*)carve(p, sizeof(double) * d.ndof);
        double* q = (double*)carve(p, sizeof(double) * d.ndof);
        double* partials = (double*)carve(p, sizeof(double) * partial_slots(d));
        PcgState* st = (PcgState*)carve(p, sizeof(PcgState));
        double* x = (double*)carve(p, sizeof(double) * d.ndof);
        const ZRing<double> zr = carve_ring<double>(p, d, q);
        return time_kernels<double, false>(d, (const double*)s_e, (const double*)K0, dinv64, x, r, p0, p1, q,
                                           zr, st, partials, reps, avg_ms, s);
    }
    if (precision == 32) {
        if (ws_bytes < ws_bytes_ref(d)) return fail(H8_ERR_WORKSPACE, "workspace too small");
        float* r = (float*)carve(p, sizeof(float) * d.ndof);
        float* x = (float*)carve(p, sizeof(float) * d.ndof);
        float* p0 = (float*)carve(p, sizeof(float) * d.ndof);
        float* p1 = (float*)carve(p, sizeof(float) * d.ndof);
        float* q = (float*)carve(p, sizeof(float) * d.ndof);
        float* dinv32 = (float*)carve(p, sizeof(float) * d.ndof);
        double* partials = (double*)carve(p, sizeof(double) * partial_slots(d));
        PcgState* st = (PcgState*)carve(p, sizeof(PcgState));
        const ZRing<float> zr = carve_ring<float>(p, d, q);
        hipLaunchKernelGGL(k_cast_f64_f32, dim3(stream_blocks(d.ndof)), dim3(STREAM_THREADS), 0, s, d.ndof,
                           dinv64, dinv32);
        LAUNCHCHK();
        return time_kernels<float, true>(d, (const float*)s_e, (const float*)K0, dinv32, x, r, p0, p1, q, zr, st,
                                         partials, reps, avg_ms, s);
    }
    return fail(H8_ERR_ARG, "precision must be 64 or 32");
}

// ---------------------------------------------------------------------------
// Batched solve
// ---------------------------------------------------------------------------
namespace {
constexpr int BCH = 2048;   // update/init chunk (elements); k_binit/k_bupdate assume it
// sector march of the batched direction kernel: 8 x 16 element columns (7 x 15
// owned nodes: small columns waste few threads on the small instances of a
// batch) and chunks of 16 node layers (a whole instance of up to 16 layers in
// one march, without the extra slab of a chunk boundary)
constexpr int BEJ = 8, BEI = 16, BKD = 16;

struct BLayout {
    int n_sys = 0, n_tiles = 0, n_chunks = 0, pstride = 0;
    size_t off_r, off_p0, off_p1, off_q, off_inst, off_tiles, off_chunks, off_sys, off_active, off_part, off_order,
        off_queue, total;
};

int batched_layout(const h8_instance* inst, int n_inst, long long total_vec, BLayout& L) {
    if (!inst || n_inst < 1) return fail(H8_ERR_ARG, "no instances");
    if (total_vec < 0) return fail(H8_ERR_ARG, "total_vec must be >= 0, got %lld", total_vec);
    int maxt = 1, maxc = 1;
    long long n_sys = 0, n_tiles = 0, n_chunks = 0;     // (the tables index them with int)
    for (int i = 0; i < n_inst; ++i) {
        const h8_instance& I = inst[i];
        if (I.nx < 1 || I.ny < 1 || I.nz < 1 || I.nrhs < 1)
            return fail(H8_ERR_ARG, "instance %d: nx, ny, nz, nrhs must be >= 1", i);
        GridDev d;
        h8_grid g{I.nx, I.ny, I.nz, 0};
        int rc = make_grid(&g, d);
        if (rc) return rc;
        sector_tiles(d, BEJ, BEI, BKD);
        const long long nch = (d.ndof + BCH - 1) / BCH;
        n_sys += I.nrhs;
        n_tiles += (long long)d.ntiles * I.nrhs;
        n_chunks += nch * I.nrhs;
        if (n_tiles > INT_MAX || n_chunks > INT_MAX)
            return fail(H8_ERR_ARG, "batch too large: more than 2^31 column tiles or vector chunks");
        maxt = std::max(maxt, d.ntiles);
        maxc = std::max(maxc, int(nch));
    }
    L.n_sys = int(n_sys);
    L.n_tiles = int(n_tiles);
    L.n_chunks = int(n_chunks);
    L.pstride = std::max(maxt, 3 * maxc);
    size_t p = 0;
    auto take = [&](size_t bytes) { size_t q = p; p += align_up(bytes); return q; };
    L.off_r = take(sizeof(double) * total_vec);
    L.off_p0 = take(sizeof(double) * total_vec);
    L.off_p1 = take(sizeof(double) * total_vec);
    L.off_q = take(sizeof(double) * total_vec);
    L.off_inst = take(sizeof(BInst) * n_inst);
    L.off_tiles = take(sizeof(BSTile) * L.n_tiles);
    L.off_chunks = take(sizeof(BChunk) * L.n_chunks);
    L.off_sys = take(sizeof(BSys) * L.n_sys);
    L.off_active = take(sizeof(int));
    L.off_part = take(sizeof(double) * (size_t)L.n_sys * L.pstride);
    L.off_order = take(sizeof(int) * (size_t)L.n_sys);
    L.off_queue = take(sizeof(int));
    L.total = p;
    return H8_OK;
}
}  // namespace

extern "C" size_t h8_batched_workspace_bytes(const h8_instance* inst, int32_t n_inst, int64_t total_vec) {
    BLayout L;
    if (batched_layout(inst, n_inst, total_vec, L)) return 0;
    return L.total;
}

// The persistent form: k_bsolve<double, EJ, EI> over the queue, one launch.
template <int EJ, int EI>
static int batched_persistent(const BArgs& A, const BVecs<double>& V, int* order, int n_sys, int* queue,
                              hipStream_t s) {
    const void* fn = (const void*)k_bsolve<double, EJ, EI>;
    int per = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, EJ * EI, 0));
    const int grid = std::max(1, std::min(n_sys, std::max(1, per) * device_cus()));
    hipLaunchKernelGGL((k_bsolve<double, EJ, EI>), dim3(grid), dim3(EJ * EI), 0, s, A, V, (const int*)order, n_sys,
                       queue, BKD);
    LAUNCHCHK();
    return H8_OK;
}

static int pcg_batched(const h8_instance* inst, int32_t n_inst, int64_t total_nodes, int64_t total_elems,
                       int64_t total_vec, const double* s_e, const double* K0, int32_t k0_per_instance,
                       const double* dinv, const double* b, double* x, double tol, int64_t maxiter, void* ws,
                       size_t ws_bytes, int32_t* status_out, int64_t* iters_out, double* rel_out, void* stream,
                       int form) {
    if (form < H8_BATCHED_LAUNCHED || form > H8_BATCHED_PERSISTENT_16X16)
        return fail(H8_ERR_ARG, "batched form %d: expected 0 (launched), 1 or 2 (persistent)", form);
    BLayout L;
    int rc = batched_layout(inst, n_inst, total_vec, L);
    if (rc) return rc;
    if (!s_e || !K0 || !dinv || !b || !x || !ws) return fail(H8_ERR_ARG, "NULL pointer");
    if (ws_bytes < L.total) return fail(H8_ERR_WORKSPACE, "workspace too small: %zu < %zu", ws_bytes, L.total);
    if (!(tol >= 0.0) || maxiter < 0) return fail(H8_ERR_ARG, "bad tol / maxiter");
    // host tables
    std::vector<BInst> hi(n_inst);
    std::vector<BSTile> ht;
    std::vector<BChunk> hc;
    std::vector<BSys> hs;
    ht.reserve(L.n_tiles);
    hc.reserve(L.n_chunks);
    hs.reserve(L.n_sys);
    for (int i = 0; i < n_inst; ++i) {
        const h8_instance& I = inst[i];
        GridDev d;
        h8_grid g{I.nx, I.ny, I.nz, 0};
        make_grid(&g, d);
        const long long m = (long long)I.nx * I.ny * I.nz;
        if (I.node_off < 0 || I.node_off + d.nnodes > total_nodes || I.elem_off < 0 || I.elem_off + m > total_elems ||
            I.vec_off < 0 || I.vec_off + d.ndof * I.nrhs > total_vec)
            return fail(H8_ERR_ARG, "instance %d: offsets outside the batch arrays", i);
        BInst& B = hi[i];
        B.nx = I.nx; B.ny = I.ny; B.nz = I.nz; B.nrhs = I.nrhs;
        B.node_off = I.node_off; B.elem_off = I.elem_off; B.vec_off = I.vec_off;
        B.sys_off = int(hs.size());
        B.k0_off = k0_per_instance ? i : 0;
        GridDev ds = d;
        sector_tiles(ds, BEJ, BEI, BKD);
        for (int c = 0; c < I.nrhs; ++c) {
            for (int t = 0; t < ds.ntiles; ++t) ht.push_back(BSTile{int(hs.size()), t});
            BSys S{};
            S.status = ST_RUNNING;
            S.inst = i;
            S.off = I.vec_off + (long long)c * d.ndof;
            S.ndof = int(d.ndof);
            S.case_idx = c;
            const int sysid = int(hs.size());
            hs.push_back(S);
            for (long long st = 0; st < d.ndof; st += BCH)
                hc.push_back(BChunk{sysid, int(st), int(std::min<long long>(BCH, d.ndof - st)), 0});
        }
    }
    hipStream_t s = (hipStream_t)stream;
    char* w = (char*)ws;
    BArgs A{};
    A.inst = (const BInst*)(w + L.off_inst);
    A.chunks = (const BChunk*)(w + L.off_chunks);
    A.sys = (BSys*)(w + L.off_sys);
    A.active = (int*)(w + L.off_active);
    A.partials = (double*)(w + L.off_part);
    A.pstride = L.pstride;
    A.tol = tol;
    A.maxiter = maxiter;
    BVecs<double> V{};
    V.s_e = s_e; V.K0 = K0; V.dinv = dinv; V.b = b; V.x = x;
    V.r = (double*)(w + L.off_r); V.p0 = (double*)(w + L.off_p0); V.p1 = (double*)(w + L.off_p1);
    V.q = (double*)(w + L.off_q);

    Ev e0, e1;
    HIPCHK(e0.create());
    HIPCHK(e1.create());
    HIPCHK(hipEventRecord(e0, s));
    HIPCHK(hipMemcpyAsync((void*)A.inst, hi.data(), sizeof(BInst) * hi.size(), hipMemcpyHostToDevice, s));
    const BSTile* stiles = (const BSTile*)(w + L.off_tiles);
    HIPCHK(hipMemcpyAsync((void*)stiles, ht.data(), sizeof(BSTile) * ht.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync((void*)A.chunks, hc.data(), sizeof(BChunk) * hc.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync((void*)A.sys, hs.data(), sizeof(BSys) * hs.size(), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(STREAM_THREADS), 0, s, 1LL, A.active, L.n_sys);
    LAUNCHCHK();
    int active = 0;
    if (form == H8_BATCHED_LAUNCHED) {
        hipLaunchKernelGGL(k_binit<double>, dim3(L.n_chunks), dim3(256), 0, s, A, V);
        LAUNCHCHK();
        HIPCHK(hipMemcpyAsync(&active, A.active, sizeof(int), hipMemcpyDeviceToHost, s));
    }
    HIPCHK(hipStreamSynchronize(s));   // also keeps the host tables alive until the copies completed
    long long launches = 0;
    if (form != H8_BATCHED_LAUNCHED) {
        // the persistent form re-initialises every system itself: systems in
        // decreasing size (the longest solves start first), one launch
        std::vector<int> ord(L.n_sys);
        for (int i = 0; i < L.n_sys; ++i) ord[i] = i;
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int c) { return hs[a].ndof > hs[c].ndof; });
        int* dord = (int*)(w + L.off_order);
        int* dq = (int*)(w + L.off_queue);
        HIPCHK(hipMemcpyAsync(dord, ord.data(), sizeof(int) * ord.size(), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(STREAM_THREADS), 0, s, 1LL, dq, 0);
        LAUNCHCHK();
        rc = form == H8_BATCHED_PERSISTENT_8X16 ? batched_persistent<8, 16>(A, V, dord, L.n_sys, dq, s)
                                                : batched_persistent<16, 16>(A, V, dord, L.n_sys, dq, s);
        if (rc) return rc;
        launches = 1;
        HIPCHK(hipStreamSynchronize(s));   // the order table stays alive until its copy completed
    }
    if (active > 0) {
        GraphKey key{};
        key.kind = 3;
        HIPCHK(hipGetDevice(&key.dev));
        key.nx = n_inst; key.ny = L.n_tiles; key.nz = L.n_chunks;
        const void* ptrs[] = {s_e, K0, dinv, b, x, ws};
        std::memcpy(key.ptr, ptrs, sizeof ptrs);
        key.chunk = chunk_for(maxiter);
        // k_bupdate reads tol and maxiter from the captured arguments, and the
        // workspace offsets of the vectors depend on total_vec
        const unsigned long long extra[] = {dbits(tol), (unsigned long long)maxiter, (unsigned long long)total_vec,
                                            (unsigned long long)L.n_sys, (unsigned long long)L.pstride};
        std::memcpy(key.extra, extra, sizeof extra);
        GraphRef exec;
        rc = get_graph(key, [&](hipStream_t cs) -> int {
            for (int it = 0; it < key.chunk; ++it) {
                hipLaunchKernelGGL((k_bsector<double, BEJ, BEI>), dim3(L.n_tiles), dim3(BEJ * BEI), 0, cs, A, V,
                                   stiles, BKD);
                hipLaunchKernelGGL(k_bupdate<double>, dim3(L.n_chunks), dim3(256), 0, cs, A, V);
            }
            LAUNCHCHK();
            return H8_OK;
        }, &exec);
        if (rc) return rc;
        while (active > 0) {
            HIPCHK(hipGraphLaunch(exec->exec, s));
            ++launches;
            HIPCHK(hipMemcpyAsync(&active, A.active, sizeof(int), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
        }
        // the last iteration's x += alpha p of every system (deferred by the update kernel)
        hipLaunchKernelGGL(k_bfinal<double>, dim3(L.n_chunks), dim3(256), 0, s, A, V);
        LAUNCHCHK();
    }
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    HIPCHK(hipMemcpy(hs.data(), A.sys, sizeof(BSys) * hs.size(), hipMemcpyDeviceToHost));
    int first_bad = H8_OK;
    for (int i = 0; i < L.n_sys; ++i) {
        const int st = hs[i].status;
        const int code = st == ST_DONE ? H8_OK : (st == ST_NOT_SPD ? H8_NOT_SPD : H8_MAXITER);
        if (status_out) status_out[i] = code;
        if (iters_out) iters_out[i] = hs[i].iters;
        if (rel_out) rel_out[i] = hs[i].rel;
        if (code != H8_OK && first_bad == H8_OK) first_bad = code;
    }
    g_batched_ms = ms;
    g_batched_launches = launches;
    if (first_bad == H8_NOT_SPD) return fail(H8_NOT_SPD, "batched PCG: p^T A p <= 0 in at least one system");
    if (first_bad == H8_MAXITER)
        return fail(H8_MAXITER, "batched PCG: at least one system did not converge in %lld iterations",
                    (long long)maxiter);
    return H8_OK;
}

extern "C" int h8_pcg_batched_f64(const h8_instance* inst, int32_t n_inst, int64_t total_nodes, int64_t total_elems,
                                  int64_t total_vec, const double* s_e, const double* K0, int32_t k0_per_instance,
                                  const double* dinv, const double* b, double* x, double tol, int64_t maxiter,
                                  void* ws, size_t ws_bytes, int32_t* status_out, int64_t* iters_out,
                                  double* rel_out, void* stream) {
    return pcg_batched(inst, n_inst, total_nodes, total_elems, total_vec, s_e, K0, k0_per_instance, dinv, b, x, tol,
                       maxiter, ws, ws_bytes, status_out, iters_out, rel_out, stream, H8_BATCHED_LAUNCHED);
}

extern "C" int h8_pcg_batched_form_f64(const h8_instance* inst, int32_t n_inst, int64_t total_nodes,
                                       int64_t total_elems, int64_t total_vec, const double* s_e, const double* K0,
                                       int32_t k0_per_instance, const double* dinv, const double* b, double* x,
                                       double tol, int64_t maxiter, void* ws, size_t ws_bytes, int32_t* status_out,
                                       int64_t* iters_out, double* rel_out, int32_t form, void* stream) {
    return pcg_batched(inst, n_inst, total_nodes, total_elems, total_vec, s_e, K0, k0_per_instance, dinv, b, x, tol,
                       maxiter, ws, ws_bytes, status_out, iters_out, rel_out, stream, form);
}

extern "C" double h8_batched_last_ms(void) { return g_batched_ms; }
extern "C" int64_t h8_batched_last_launches(void) { return g_batched_launches; }

// ---------------------------------------------------------------------------
// Hex20: element stiffness, and the structured-box operator, diagonal and PCG
// (h8_hex20.h)
// ---------------------------------------------------------------------------
extern "C" int h8_ke20_batched_f64(const double* X, const double* D, int64_t d_stride, int64_t m, double* ke,
                                   void* stream) {
    if (m < 0) return fail(H8_ERR_ARG, "m < 0");
    if (m == 0) return H8_OK;
    if (!X || !D || !ke) return fail(H8_ERR_ARG, "NULL pointer");
    if (d_stride != 0 && d_stride != 36) return fail(H8_ERR_ARG, "d_stride must be 0 or 36");
    if (m >= (1LL << 31)) return fail(H8_ERR_ARG, "too many elements");
    hipLaunchKernelGGL(k_ke20, dim3(unsigned(m)), dim3(256), 0, (hipStream_t)stream, (long long)m, X, D,
                       (long long)d_stride, ke);
    LAUNCHCHK();
    return H8_OK;
}

static unsigned h20_blocks(const H20Grid& g) {
    return unsigned(std::min<long long>((g.nnodes + H20_THREADS - 1) / H20_THREADS, H20_MAX_BLOCKS));
}

static int h20_args(int32_t nx, int32_t ny, int32_t nz, H20Grid& g) {
    if (nx < 1 || ny < 1 || nz < 1) return fail(H8_ERR_ARG, "hex20 box: nx, ny, nz must be >= 1");
    g = h20_grid(nx, ny, nz);
    if (g.ndof >= (1LL << 31)) return fail(H8_ERR_ARG, "hex20 box: %lld DOFs, the limit is 2^31", g.ndof);
    return H8_OK;
}

// y = K u (diag: y = the diagonal of K) over every DOF of the box, one launch
static void h20_node_launch(const H20Grid& g, bool diag, const double* s_e, const double* K0, const double* u,
                            double* y, hipStream_t s) {
    const unsigned blocks = h20_blocks(g);
    if (diag)
        hipLaunchKernelGGL(k_h20_node<true>, dim3(blocks), dim3(H20_THREADS), 0, s, g, s_e, K0, u, y);
    else
        hipLaunchKernelGGL(k_h20_node<false>, dim3(blocks), dim3(H20_THREADS), 0, s, g, s_e, K0, u, y);
}

extern "C" int h20_spmv_f64(int32_t nx, int32_t ny, int32_t nz, const double* s_e, const double* K0,
                            const double* u, double* y, void* stream) {
    H20Grid g;
    int rc = h20_args(nx, ny, nz, g);
    if (rc) return rc;
    if (!s_e || !K0 || !u || !y) return fail(H8_ERR_ARG, "NULL pointer");
    hipStream_t s = (hipStream_t)stream;
    h20_node_launch(g, false, s_e, K0, u, y, s);
    LAUNCHCHK();
    return H8_OK;
}

extern "C" int h20_diag_f64(int32_t nx, int32_t ny, int32_t nz, const double* s_e, const double* K0,
                            const uint8_t* free_mask, double* dinv, int32_t* d_flag, void* stream) {
    H20Grid g;
    int rc = h20_args(nx, ny, nz, g);
    if (rc) return rc;
    if (!s_e || !K0 || !free_mask || !dinv || !d_flag) return fail(H8_ERR_ARG, "NULL pointer");
    hipStream_t s = (hipStream_t)stream;
    const int sb = stream_blocks(g.ndof);
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(STREAM_THREADS), 0, s, 1LL, (int*)d_flag, 0);
    h20_node_launch(g, true, s_e, K0, nullptr, dinv, s);
    hipLaunchKernelGGL(k_h20_dinv, dim3(sb), dim3(256), 0, s, g.ndof, (const double*)dinv, free_mask, dinv,
                       (int*)d_flag);
    LAUNCHCHK();
    int32_t flag = 0;
    HIPCHK(hipMemcpyAsync(&flag, d_flag, sizeof flag, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (flag) return fail(H8_ZERO_DIAG, "jacobi_preconditioner: A has a zero diagonal entry.");
    return H8_OK;
}

static size_t h20_ws_bytes(const H20Grid& g) {
    return 4 * align_up(sizeof(double) * g.ndof) + align_up(sizeof(double) * 4 * 2048) + align_up(sizeof(PcgState));
}

extern "C" size_t h20_pcg_workspace_bytes(int32_t nx, int32_t ny, int32_t nz) {
    H20Grid g;
    if (h20_args(nx, ny, nz, g)) return 0;
    return h20_ws_bytes(g);
}

extern "C" int h20_pcg_f64(int32_t nx, int32_t ny, int32_t nz, const double* s_e, const double* K0,
                           const double* dinv, const double* b, double* x, double tol, int64_t maxiter, void* ws,
                           size_t ws_bytes, h8_pcg_stats* stats, void* stream) {
    H20Grid g;
    int rc = h20_args(nx, ny, nz, g);
    if (rc) return rc;
    if (!s_e || !K0 || !dinv || !b || !x || !ws) return fail(H8_ERR_ARG, "NULL pointer");
    if (!(tol >= 0.0)) return fail(H8_ERR_ARG, "tol must be >= 0");
    if (maxiter < 0) return fail(H8_ERR_ARG, "maxiter must be >= 0");
    if (ws_bytes < h20_ws_bytes(g))
        return fail(H8_ERR_WORKSPACE, "workspace too small: %zu < %zu", ws_bytes, h20_ws_bytes(g));
    hipStream_t s = (hipStream_t)stream;
    char* p = (char*)ws;
    double* r = (double*)carve(p, sizeof(double) * g.ndof);
    double* p0 = (double*)carve(p, sizeof(double) * g.ndof);
    double* p1 = (double*)carve(p, sizeof(double) * g.ndof);
    double* q = (double*)carve(p, sizeof(double) * g.ndof);
    double* partials = (double*)carve(p, sizeof(double) * 4 * 2048);
    PcgState* st = (PcgState*)carve(p, sizeof(PcgState));
    const long long n = g.ndof;
    const int sb = stream_blocks(n);
    Ev e0, e1;
    HIPCHK(e0.create());
    HIPCHK(e1.create());
    HIPCHK(hipEventRecord(e0, s));
    PcgState hst{};
    hst.tol = tol;
    hst.maxiter = maxiter;
    hst.status = ST_RUNNING;
    HIPCHK(hipMemcpyAsync(st, &hst, sizeof(PcgState), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_fill<double>, dim3(sb), dim3(STREAM_THREADS), 0, s, n, x, 0.0);
    hipLaunchKernelGGL(k_pcg_init<double>, dim3(sb), dim3(STREAM_THREADS), 0, s, n, b, (const double*)nullptr, dinv,
                       r, st, partials);
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(&hst, st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    long long launches = 0;
    if (hst.status == ST_RUNNING && maxiter > 0) {
        GraphKey key{};
        key.kind = 6;
        HIPCHK(hipGetDevice(&key.dev));
        key.nx = nx; key.ny = ny; key.nz = nz;
        const void* ptrs[] = {s_e, K0, dinv, x, r, p0, p1, q, st, partials};
        std::memcpy(key.ptr, ptrs, sizeof ptrs);
        key.chunk = chunk_for(maxiter);
        GraphRef exec;
        rc = get_graph(key, [&](hipStream_t cs) -> int {
            for (int it = 0; it < key.chunk; ++it) {
                hipLaunchKernelGGL(k_h20_dir, dim3(sb), dim3(256), 0, cs, n, (const double*)r, dinv, p0, p1,
                                   (const PcgState*)st);
                hipLaunchKernelGGL(k_h20_node_pcg, dim3(h20_blocks(g)), dim3(H20_THREADS), 0, cs, g, s_e, K0,
                                   (const double*)p0, (const double*)p1, q, st, partials);
                hipLaunchKernelGGL((k_pcg_update<double, false, true>), dim3(sb), dim3(STREAM_THREADS), 0, cs, n, x,
                                   r, (const double*)p0, (const double*)p1, (const double*)q, dinv, st, partials,
                                   (double*)nullptr);
            }
            LAUNCHCHK();
            return H8_OK;
        }, &exec);
        if (rc) return rc;
        for (;;) {
            HIPCHK(hipGraphLaunch(exec->exec, s));
            ++launches;
            HIPCHK(hipMemcpyAsync(&hst, st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            if (hst.status != ST_RUNNING) break;
        }
    } else if (hst.status == ST_RUNNING) {
        hst.status = ST_MAXITER;
        hst.iters = 0;
    }
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    fill_stats(stats, hst, ms, launches);
    if (hst.status == ST_NOT_SPD)
        return fail(H8_NOT_SPD, "preconditioned_cg: p^T A p = %.3e <= 0 at iteration %lld -- A is not SPD.",
                    hst.bad_pq, hst.iters);
    if (hst.status == ST_MAXITER)
        return fail(H8_MAXITER,
                    "preconditioned_cg: did not converge in %lld iterations (final relative residual "
                    "%.3e >= tol=%.1e).",
                    (long long)maxiter, hst.rel, tol);
    return H8_OK;
}
