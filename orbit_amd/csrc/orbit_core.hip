// liborbit_hip: MI355X periodic-orbit transport solve.
//
// Device layout: one vector per tracer slot, idx = col*16 + k (col = j*nx + i,
// k = level padded to 16).  A_tau = I + dtau*L_tau is held as
//   V   : per column, a 16x16 block of identity + dtau*(in-column off-diagonals),
//   d   : per tracer, dtau * diag(L) (loss, deposition and the transport diagonal),
//   H   : per cell, dtau * the 4 same-level horizontal off-diagonals (W,E,S,N),
//   P   : per tracer, the inverse of each column block V + diag(d)
// so that P_tau = A_tau^{-1} is applied by red-black column Gauss-Seidel sweeps in
// correction form  x_c += P_c (b - A x)_c  (colour c = (i+j) mod 2).  A fixed
// sweep count makes the approximate P a fixed linear map whose transpose is the
// same sweep run on A^T, which the adjoint uses.
//
// Reference call surface replaced (SumilThakr/orbit v1.0.4):
//   _factor_one_bin / _UmfpackLU       orbit/core/orbit.py:79-96, solve.py:210-283
//   apply_monodromy / compute_s         orbit/core/orbit.py:99-120
//   solve_orbit_one_species             orbit/core/orbit.py:123-210
//   apply_monodromy_adjoint ... solve   orbit/core/orbit_adjoint.py:36-144
//   mass_balance_diagnostics            orbit/core/operator.py:598-652
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <type_traits>

#include <hip/hip_ext.h>

#include "orbit_internal.h"

namespace orbit {

static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }
int fail(int code, const std::string& m) {
    g_err = m;
    return code;
}

int Ctx::ensure(int slot, size_t bytes) {
    if (scratch[slot].bytes >= bytes) return ORBIT_OK;
    if (scratch[slot].p) {
        hipStreamSynchronize(stream);
        hipFree(scratch[slot].p);
        scratch[slot].p = nullptr;
        scratch[slot].bytes = 0;
    }
    size_t b = std::max<size_t>(bytes, 256);
    ORBIT_HIP_CHECK(hipMalloc(&scratch[slot].p, b));
    scratch[slot].bytes = b;
    return ORBIT_OK;
}

// Host-to-device copy of a small host array that completes before returning:
// a pageable source may be read by the runtime when the stream reaches the copy,
// so the source must not be freed or rewritten before then.
int Ctx::ensure_pinned(size_t n) {
    if (pinned_n >= n) return ORBIT_OK;
    if (pinned) hipHostFree(pinned);
    pinned = nullptr;
    pinned_n = 0;
    const size_t want = std::max<size_t>(n, 256);
    if (hipHostMalloc((void**)&pinned, want * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        pinned = nullptr;
        return fail(ORBIT_ERR_NOMEM, "solve: pinned host allocation failed");
    }
    pinned_n = want;
    return ORBIT_OK;
}

int h2d_sync(hipStream_t st, void* dst, const void* src, size_t bytes) {
    ORBIT_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
    ORBIT_HIP_CHECK(hipStreamSynchronize(st));
    return ORBIT_OK;
}

// Fixed device scratch of orbit_solve_forward_dev_noalloc: one array per GPU,
// held from the library's load until the process ends.  In host code the
// array's name is only a stand-in for the symbol; its device address on each
// GPU comes from fixed_scratch_address and is the only form passed to kernels
// or copies.
static __device__ __align__(256) unsigned char scratch[ORBIT_FIXED_SCRATCH_BYTES];

// Device address of `scratch` on `device` (which must be current), looked up
// once per device and kept: each GPU holds its own copy at its own address.
int fixed_scratch_address(int device, void** out) {
    static std::mutex mu;
    static std::map<int, void*> addr;
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(device);
    if (it == addr.end()) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(ORBIT_ERR_HIP, std::string("fixed scratch array: hipGetSymbolAddress on device ") +
                                           std::to_string(device) + ": " + hipGetErrorString(e));
        it = addr.emplace(device, p).first;
    }
    *out = it->second;
    return ORBIT_OK;
}

// Per-slot flags and counts (up to 64 slots) travel as kernel arguments: the
// write is ordered on the stream like any launch and the host array can be
// dropped at once, so the host does not wait for the device.
struct SmallInts {
    int v[64];
};
__global__ void k_set_ints(int* dst, SmallInts a, int n) {
    const int i = threadIdx.x;
    if (i < n) dst[i] = a.v[i];
}
int set_ints(hipStream_t st, int* dst, const int* src, int n) {
    if (n > 64) return h2d_sync(st, dst, src, (size_t)n * sizeof(int));
    SmallInts a;
    for (int i = 0; i < n; ++i) a.v[i] = src[i];
    hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(64), 0, st, dst, a, n);
    ORBIT_HIP_CHECK(hipGetLastError());
    return ORBIT_OK;
}

template <class T>
T* Opset::alloc(size_t n) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
    owned.push_back(p);
    bytes += (double)n * sizeof(T);
    return reinterpret_cast<T*>(p);
}

// ------------------------------------------------------------------------- //
// element loads
// ------------------------------------------------------------------------- //
__device__ __forceinline__ double ldv(double v) { return v; }
__device__ __forceinline__ float ldv(float v) { return v; }
__device__ __forceinline__ float ldv(uint16_t v) { return bf16_to_f32(v); }

// ------------------------------------------------------------------------- //
// column-block factorization: P = (V + diag(d))^{-1} by unpivoted Gauss-Jordan.
// A_tau is a nonsingular M-matrix (every column block is one), so pivots stay
// positive.  64 threads = 4 blocks, lane = row.
// ------------------------------------------------------------------------- //
__global__ __launch_bounds__(64) void k_factor(const double* __restrict__ V, long long vstride,
                                               const double* __restrict__ d, long long dstride,
                                               double* __restrict__ P, long long pstride, int ncol) {
    __shared__ double piv[4][32];
    const int t = blockIdx.y;
    const int grp = threadIdx.x >> 4, lane = threadIdx.x & 15;
    const int col = blockIdx.x * 4 + grp;
    const bool valid = col < ncol;
    const int cc = valid ? col : 0;
    const double* Vr = V + (size_t)t * vstride + (size_t)cc * 256 + lane * 16;
    const double dd = d[(size_t)t * dstride + (size_t)cc * 16 + lane];
    double row[32];
#pragma unroll
    for (int m = 0; m < 16; ++m) row[m] = Vr[m] + (m == lane ? dd : 0.0);
#pragma unroll
    for (int m = 0; m < 16; ++m) row[16 + m] = (m == lane) ? 1.0 : 0.0;
#pragma unroll
    for (int p = 0; p < 16; ++p) {
        if (lane == p) {
#pragma unroll
            for (int m = 0; m < 32; ++m) piv[grp][m] = row[m];
        }
        __syncthreads();
        const double pv = piv[grp][p];
        if (lane == p) {
            const double inv = 1.0 / pv;
#pragma unroll
            for (int m = 0; m < 32; ++m) row[m] *= inv;
        } else {
            const double f = row[p] / pv;
#pragma unroll
            for (int m = 0; m < 32; ++m) row[m] -= f * piv[grp][m];
        }
        __syncthreads();
    }
    if (valid) {
        double* Pr = P + (size_t)t * pstride + (size_t)col * 256 + lane * 16;
#pragma unroll
        for (int m = 0; m < 16; ++m) Pr[m] = row[16 + m];
    }
}

// dst block b = transpose of src block b (16x16), optional bf16 copy of dst.
template <class T>
__global__ void k_transpose_blocks(const T* __restrict__ src, T* __restrict__ dst, size_t nblk) {
    size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nblk * 256) return;
    size_t b = g >> 8;
    int e = (int)(g & 255), i = e >> 4, j = e & 15;
    dst[b * 256 + j * 16 + i] = src[g];
}

__global__ void k_to_bf16(const double* __restrict__ src, uint16_t* __restrict__ dst, size_t n) {
    size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) dst[g] = f32_to_bf16((float)src[g]);
}

template <class S, class D>
__global__ void k_convert(const S* __restrict__ src, D* __restrict__ dst, size_t n) {
    size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) dst[g] = (D)src[g];
}

// Transposed 4-point stencil: A^T[c, nbr] = A[nbr, c], i.e. the W coefficient of
// cell c in A^T is the E coefficient of its west neighbour in A.
template <class T>
__global__ void k_transpose_stencil(const T* __restrict__ H, T* __restrict__ HT, int nx, int ny, int px) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    int ncol = nx * ny;
    if (g >= ncol * 16) return;
    int col = g >> 4, k = g & 15;
    int j = col / nx, i = col - j * nx;
    const int cw = west_col(col, i, nx, px), ce = east_col(col, i, nx, px);
    T w = cw >= 0 ? H[(size_t)cw * 64 + 16 + k] : T(0);
    T e = ce >= 0 ? H[(size_t)ce * 64 + 0 + k] : T(0);
    T s = j > 0 ? H[(size_t)(col - nx) * 64 + 48 + k] : T(0);
    T n = j < ny - 1 ? H[(size_t)(col + nx) * 64 + 32 + k] : T(0);
    HT[(size_t)col * 64 + 0 + k] = w;
    HT[(size_t)col * 64 + 16 + k] = e;
    HT[(size_t)col * 64 + 32 + k] = s;
    HT[(size_t)col * 64 + 48 + k] = n;
}

// ------------------------------------------------------------------------- //
// red-black column half-sweep, correction form
// ------------------------------------------------------------------------- //
struct SweepArgs {
    int ncols;
    int nx, ny;
    int nxc;            // width of the red-black part of a row (nx, or nx - 1 with four colours)
    const void* V;
    long long vstride;  // per tracer (0 when the block is shared)
    const void* Vt;     // transpose of V (read by the CS mapping, which needs vstride 0)
    const void* H;
    long long hstride;
    const void* d;
    long long dstride;
    const void* P;
    long long pstride;
    const void* b;
    long long bstride;  // per slot
    void* x;
    long long xstride;  // per slot
    const int* tmap;    // slot -> tracer
    int tmap_identity;  // slot s is tracer s (no tmap load before the per-tracer rows)
    const int* active;  // nullable
    int own_zero;       // x of this colour is known to be zero (not read)
    int nbr_zero;       // x of the other colour is known to be zero (not read)
    int nslots;
    int tiles_per_xcd;  // column tiles given to each of the 8 XCDs (xcd_map)
    int xcd_map;        // 1: XCD-contiguous tile ranges; 0: linear block order
    const int* counts;  // nullable: per-slot sweep count; slots past theirs skip
    int sweep_idx;
    int colour;         // 0 red ((i + j) even), 1 black
    int px;             // periodic x (wrap couplings)
    float omega;        // relaxation of the correction (1: Gauss-Seidel)
    // group launches (slots of several operator sets, four-slot mapping): the
    // slot groups' operators and geometry, and per block its (group, tile)
    const struct GroupOp* gops;
    const int* btab;    // group << 16 | tile, -1 for a block with no work
    // epilogue of a slot's last sweep (epi_out null: none): at the colour's
    // columns epi_out = x + epi_alpha * epi_add (FP64 sweeps) or x + (float)
    // epi_add (FP32 sweeps; epi_add FP32 when epi_add_f32), the next bin's
    // right-hand side or the outer
    // residual, so no separate pass forms it
    void* epi_out;
    long long epi_os;
    const void* epi_add;
    int epi_add_f32;
    long long epi_as;
    double epi_alpha;
    int nsweeps;        // sweeps of the inner solve (a slot's last is min(count, nsweeps) - 1)
};

// One slot group of a group launch: up to four consecutive slots of one
// operator set, with that set's operator arrays for the launch's bin and sweep
// precision (tracer t's diagonal and inverse blocks at + t * stride) and grid.
struct GroupOp {
    const void* V;
    const void* Vt;
    const void* H;
    const void* d;
    const void* P;
    long long dstride, pstride;
    int ncols[4];       // columns of each colour
    int nx, ny, nxc, px;
    int first, count;   // slots first .. first + count - 1
    float omega;
};

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2):
// block L runs on XCD L%8.  Give each XCD a contiguous range of column tiles and
// run all slots of a tile there back to back, so the shared column blocks and
// stencil of a tile, and the neighbouring tiles' x, are fetched into one L2.
__device__ __forceinline__ void xcd_tile_slot(int L, int nslots, int tiles_per_xcd, int& tile, int& slot) {
    const int xcd = L & 7, r = L >> 3;
    const int lt = r / nslots;
    slot = r - lt * nslots;
    tile = xcd * tiles_per_xcd + lt;
}

// Linear order: consecutive blocks take the slot groups of one tile.
__device__ __forceinline__ void linear_tile_slot(int L, int nslots, int& tile, int& slot) {
    tile = L / nslots;
    slot = L - tile * nslots;
}

// Column id of the idx-th column of colour c in row-major order.  Colours 0 / 1
// (red / black) cover the first w columns of every row: rows 2p and 2p+1
// together hold exactly w columns of each, row 2p's start at i = c.  Colours
// 2 / 3 are the last column of the even / odd rows (a periodic x with an odd
// nx, where w = nx - 1).
__device__ __forceinline__ void colour_col(int idx, int c, int w, int nx, int& col, int& i, int& j) {
    if (c >= 2) {
        j = 2 * idx + (c - 2);
        i = nx - 1;
    } else {
        const int p = idx / w, r = idx - p * w;
        const int n0 = (w + 1 - c) >> 1;
        if (r < n0) {
            j = 2 * p;
            i = c + 2 * r;
        } else {
            j = 2 * p + 1;
            i = (1 - c) + 2 * (r - n0);
        }
    }
    col = j * nx + i;
}

// Lanes exchanging through LDS here always belong to one wavefront (a column's
// 16 levels are 16 consecutive lanes), so wave-level ordering is enough.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One colour's half-sweep: x_c += P_c (b - A x)_c for every column c of the colour
// and every tracer slot.  Lane k of a 16-lane group owns level k of one column of
// one slot.  CS = 0: a workgroup of W threads is W/16 columns of one slot; CS = 1:
// W/64 columns x 4 slots, so each wave reads the shared column block and stencil
// of one column once for four slots.  W (64, 128 or 256) and the block order
// (XCD-contiguous or linear) are launch choices; LDS is sized by the launch
// (2*W values of CT, plus W/64 staged blocks of AT for CS = 1).  OZ / NZ: x of
// this / the other colour is known to be zero and is not read.
// Every load is in bounds and issued up front (no early exits, no divergent
// branches around loads): missing edge neighbours read this column's x and are
// replaced by zero with a select, so stale values there never enter; inactive
// slots and padding columns compute but do not store.  The per-lane arithmetic
// does not depend on CS, so a slot's result is the same in any batch.
// Matrix-core product of the shared column block with the four slots' x
// (CS = 1): D (16x16) = A (16x4) B (4x16) accumulated over four k-slices,
// lane l holding A[l % 16][l / 16], B[l / 16][l % 16] and D[i][l % 16] with
// i = 4 (l / 16) + r (FP32) or l / 16 + 4 r (FP64), r = 0..3 (measured).
typedef float mf_f4 __attribute__((ext_vector_type(4)));
typedef double mf_d4 __attribute__((ext_vector_type(4)));
template <class T> struct MfAcc;
template <> struct MfAcc<float> { using V = mf_f4; };
template <> struct MfAcc<double> { using V = mf_d4; };
__device__ __forceinline__ void mfma_16x16x4(float a, float b, mf_f4& d) {
    d = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, d, 0, 0, 0);
}
__device__ __forceinline__ void mfma_16x16x4(double a, double b, mf_d4& d) {
    d = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, d, 0, 0, 0);
}

template <class XT, class AT, class PT, class CT, int CS, bool OZ, bool NZ, bool MF = false, bool GRP = false>
__global__ __launch_bounds__(256) void k_halfsweep(SweepArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    AT* sV = (AT*)smem;  // CS = 1 only: blockDim.x / 64 blocks of 256
    CT* sx = (CT*)(smem + (CS ? (blockDim.x >> 6) * 256 * sizeof(AT) : 0));
    CT* sr = sx + blockDim.x;
    const int lane = threadIdx.x & 15, grp = threadIdx.x >> 4;
    const int tc = CS ? (int)(blockDim.x >> 6) : (int)(blockDim.x >> 4);  // columns per tile
    // this block's operator and grid: the launch's, or (group launch) its slot group's
    int ncols, nx, ny, nxc, px, tile, slot, idx, ls;
    bool slot_ok;
    float omega;
    const void *Vb, *Vtb, *Hb, *db, *Pb;
    long long vstride, hstride, dstride, pstride;
    if constexpr (GRP) {  // four-slot mapping only
        const int e = a.btab[blockIdx.x];
        if (e < 0) return;  // uniform over the block
        const GroupOp& go = a.gops[e >> 16];
        tile = e & 0xffff;
        ncols = go.ncols[a.colour];
        nx = go.nx; ny = go.ny; nxc = go.nxc; px = go.px; omega = go.omega;
        Vb = go.V; Vtb = go.Vt; Hb = go.H; db = go.d; Pb = go.P;
        vstride = 0; hstride = 0; dstride = go.dstride; pstride = go.pstride;
        const int q = grp & 3;
        slot = go.first + q;
        slot_ok = q < go.count;
        ls = go.first + (q < go.count ? q : go.count - 1);
        idx = tile * tc + (grp >> 2);
    } else {
        const int nsg = CS ? (a.nslots + 3) >> 2 : a.nslots;
        int sg;
        if (a.xcd_map)
            xcd_tile_slot(blockIdx.x, nsg, a.tiles_per_xcd, tile, sg);
        else
            linear_tile_slot(blockIdx.x, nsg, tile, sg);
        ncols = a.ncols;
        if (tile * tc >= ncols) return;  // uniform over the block
        nx = a.nx; ny = a.ny; nxc = a.nxc; px = a.px; omega = a.omega;
        Vb = a.V; Vtb = a.Vt; Hb = a.H; db = a.d; Pb = a.P;
        vstride = a.vstride; hstride = a.hstride; dstride = a.dstride; pstride = a.pstride;
        if (CS) {
            slot = sg * 4 + (grp & 3);
            idx = tile * tc + (grp >> 2);
        } else {
            slot = sg;
            idx = tile * tc + grp;
        }
        slot_ok = slot < a.nslots;
        ls = slot < a.nslots ? slot : a.nslots - 1;
    }
    const bool valid = idx < ncols && slot_ok;
    int col, i, j;
    colour_col(idx < ncols ? idx : 0, a.colour, nxc, nx, col, i, j);
    const size_t base = (size_t)col * 16 + lane;
    const XT* __restrict__ b = (const XT*)a.b + (size_t)ls * a.bstride;
    XT* __restrict__ x = (XT*)a.x + (size_t)ls * a.xstride;
    const int tr = a.tmap_identity ? ls : a.tmap[ls];
    int act = 1, cnt = 0x7fffffff;
    if (a.active) act = a.active[ls];
    if (a.counts) cnt = a.counts[ls];

    // tracer-independent loads: rhs, own x, neighbours' x
    const CT bv = (CT)b[base];
    const CT xo = OZ ? CT(0) : (CT)x[base];
    CT xw = 0, xe = 0, xs = 0, xn = 0;
    const int cw = west_col(col, i, nx, px), ce = east_col(col, i, nx, px);
    if (!NZ) {
        xw = (CT)x[(size_t)(cw >= 0 ? cw : col) * 16 + lane];
        xe = (CT)x[(size_t)(ce >= 0 ? ce : col) * 16 + lane];
        xs = (CT)x[(size_t)(j > 0 ? col - nx : col) * 16 + lane];
        xn = (CT)x[(size_t)(j < ny - 1 ? col + nx : col) * 16 + lane];
    }
    // per-tracer operator rows
    const AT* __restrict__ V = (const AT*)Vb + (size_t)tr * vstride;
    const AT* __restrict__ H = (const AT*)Hb + (size_t)tr * hstride;
    const AT* __restrict__ d = (const AT*)db + (size_t)tr * dstride;
    const PT* __restrict__ P = (const PT*)Pb + (size_t)tr * pstride;
    const PT* Pr = P + (size_t)col * 256 + lane * 16;
    PT prow[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) prow[m] = Pr[m];
    // CS: the wave's column block (shared by its four slots) is staged in LDS,
    // transposed, 4 elements a lane: sV[m*16 + k] = V[k][m].  Else a row a lane.
    AT vrow[CS ? 4 : 16];
    CT dv = 0;
    if (!OZ) {
        if (CS && MF) {
            // the matrix cores' A operand of k-slice q: lane l takes V[l % 16][4q + l / 16],
            // i.e. element 64q + l of the transposed block (four coalesced 64-lane rows)
            const AT* Vt = (const AT*)Vtb + (size_t)col * 256 + (threadIdx.x & 63);
#pragma unroll
            for (int q = 0; q < 4; ++q) vrow[q] = Vt[64 * q];
        } else if (CS) {
            const AT* Vt = (const AT*)Vtb + (size_t)col * 256 + (threadIdx.x & 63) * 4;
#pragma unroll
            for (int q = 0; q < 4; ++q) vrow[q] = Vt[q];
        } else {
            // row `lane` from the block's transpose (element m at m * 16 + lane):
            // the 16 lanes of a column read contiguous bytes per element
            const AT* Vr = (const AT*)Vtb + (size_t)tr * vstride + (size_t)col * 256 + lane;
#pragma unroll
            for (int m = 0; m < 16; ++m) vrow[m] = Vr[m * 16];
        }
        dv = (CT)d[base];
    }
    CT hw = 0, he = 0, hs = 0, hn = 0;
    if (!NZ) {
        const AT* Hc = H + (size_t)col * 64 + lane;
        hw = (CT)Hc[0];
        he = (CT)Hc[16];
        hs = (CT)Hc[32];
        hn = (CT)Hc[48];
    }
    const bool on = valid & (act != 0) & (a.sweep_idx < cnt);

    CT r = bv;
    if (!NZ) {
        CT hx = 0;
        hx += hw * (cw >= 0 ? xw : CT(0));
        hx += he * (ce >= 0 ? xe : CT(0));
        hx += hs * (j > 0 ? xs : CT(0));
        hx += hn * (j < ny - 1 ? xn : CT(0));
        r -= hx;
    }
    if (!OZ) {
        sx[threadIdx.x] = xo;
        CT acc = dv * xo;
        if (CS && MF) {
            // V [x_0 x_1 x_2 x_3] on the matrix cores: B's column j < 4 is slot j's
            // x (the wave's four 16-lane groups), the other columns are zero; the
            // product's column j goes back to slot j's lanes through LDS
            const int wl = threadIdx.x & 63, jc = wl & 15, kk = wl >> 4;
            const CT* swx = sx + (threadIdx.x - wl);
            wave_sync();
            typename MfAcc<CT>::V dacc = {CT(0), CT(0), CT(0), CT(0)};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const CT bq = jc < 4 ? swx[jc * 16 + 4 * q + kk] : CT(0);
                mfma_16x16x4((CT)vrow[q], bq, dacc);
            }
            // D rows of lane l: 4 (l / 16) + r for the FP32 product, l / 16 + 4 r for FP64
            constexpr bool f64 = sizeof(CT) == 8;
            CT* so = (CT*)(sV + (threadIdx.x >> 6) * 256);
            if (jc < 4) {
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) so[jc * 16 + (f64 ? kk + 4 * rr : 4 * kk + rr)] = dacc[rr];
            }
            wave_sync();
            acc += so[wl];
        } else if (CS) {
            AT* sv = sV + (threadIdx.x >> 6) * 256;
#pragma unroll
            for (int q = 0; q < 4; ++q) sv[(threadIdx.x & 63) * 4 + q] = vrow[q];
            wave_sync();
#pragma unroll
            for (int m = 0; m < 16; ++m) acc += (CT)sv[m * 16 + lane] * sx[(grp << 4) + m];
        } else {
            wave_sync();
#pragma unroll
            for (int m = 0; m < 16; ++m) acc += (CT)vrow[m] * sx[(grp << 4) + m];
        }
        r -= acc;
    }
    sr[threadIdx.x] = r;
    wave_sync();
    CT dx = 0;
#pragma unroll
    for (int m = 0; m < 16; ++m) dx += (CT)ldv(prow[m]) * sr[(grp << 4) + m];
    if (on) {
        const XT xn = (XT)(xo + (CT)omega * dx);
        x[base] = xn;
        if (a.epi_out && a.sweep_idx == (cnt < a.nsweeps ? cnt : a.nsweeps) - 1) {
            const size_t ia = (size_t)ls * a.epi_as + base;
            const double ad = a.epi_add_f32 ? (double)((const float*)a.epi_add)[ia] : ((const double*)a.epi_add)[ia];
            if constexpr (sizeof(XT) == 4)
                ((float*)a.epi_out)[(size_t)ls * a.epi_os + base] = (float)xn + (float)ad;
            else
                ((double*)a.epi_out)[(size_t)ls * a.epi_os + base] = (double)xn + a.epi_alpha * ad;
        }
    }
}

// The first sweep of an FP32 inner solve (x = 0; BF16 inverse blocks, one
// shared transport, two colours) in one launch over column patches (the
// geometry is PatchCol's).  From x = 0 a red column's new value is
// x_r = omega P_r b_r (no neighbour, no column block), so a workgroup computes
// the reds of its patch and the ring reds its black columns touch, passes them
// to the blacks through LDS, and the blacks finish with
// x_b = omega P_b (b_b - H x_r): the arithmetic of the two half-sweeps, hence
// their bits, with no kernel boundary between them.  A ring red is computed
// by every patch it borders.  A workgroup of W waves takes one patch and four
// slots; a wave is one column x four slots (lane = level + 16 * slot) and
// takes the red items w, w + W, ... of the patch, then its black items the
// same way.  Every load is in bounds and issued before the first product;
// missing neighbours and columns past the grid's edge read a column of the
// patch and are masked or not stored; inactive slots compute but do not store.
// With one sweep per inner solve this is also the last sweep, so it carries
// the half-sweeps' epilogue.
struct FirstSweepArgs {
    int nx, ny, px;
    int npx, npatches, tiles_per_xcd;
    const float* H;       // shared stencil (transposed for the adjoint)
    const uint16_t* P;    // BF16 inverse blocks
    long long pstride;    // per tracer
    const float* b;
    long long bstride;    // per slot
    float* x;
    long long xstride;
    const int* tmap;
    int tmap_identity;
    const int* active;    // nullable
    int nslots;
    float omega;
    // epilogue (epi_out null: none): epi_out = x + epi_add at every column of
    // a slot whose last sweep this is (min(counts[slot], nsweeps) == 1)
    float* epi_out;
    long long epi_os;
    const void* epi_add;  // FP32 when epi_add_f32, else FP64
    int epi_add_f32;
    long long epi_as;
    const int* counts;    // nullable: per-slot sweep counts
    int nsweeps;
    // group launches (slots of several operator sets): per block its
    // (slot group, patch); geometry and operators are the slot group's
    const GroupOp* gops;
    const int* btab;      // group << 16 | patch, -1 for a block with no work
    // hand-off (HO): the previous bin's last black half-sweep in the same
    // launch.  Each black item first finishes the previous bin's sweep at its
    // column, x_b += omega P_b (b - A x)_b with the previous bin's operator
    // (every red of the previous bin is final: the launch before wrote them),
    // and that new value (plus the previous bin's epilogue term, when it has
    // one) is this bin's right-hand side there, which the unfused sequence
    // would read back from memory.
    const GroupOp* hgops; // group launches: the slot groups' operators of the previous bin
    const float* hVt;     // previous bin: transposed column blocks (shared)
    const float* hH;      // stencil (shared)
    const float* hd;      // diagonals, per tracer (stride hdstride)
    const uint16_t* hP;   // BF16 inverse blocks, per tracer (stride pstride)
    long long hdstride;
    const float* hb;      // the previous bin's right-hand side
    long long hbstride;
    float* hx;            // the previous bin's x (its reds final)
    long long hxstride;
    float hepi_omega;     // the previous bin's relaxation
    float* hepi_out;      // previous bin's epilogue (null: none): out = x + add
    long long hepi_os;
    const void* hepi_add;
    int hepi_add_f32;
    long long hepi_as;
};

template <int PA, int PB, int W, bool GRP = false, bool HO = false>
__global__ __launch_bounds__(W * 64) void k_first_sweep(FirstSweepArgs a) {
    constexpr int NRO = PA * PB / 2, NR = NRO + PA + PB, NB = PA * PB / 2, NPOS = (PA + 2) * (PB + 2);
    constexpr int RW = (NR + W - 1) / W, BW = (NB + W - 1) / W, IW = RW > BW ? RW : BW;
    __shared__ float sgrid[NPOS][64];  // new red values by ext position
    __shared__ float sr[W][IW][64];    // a column's right-hand side / residual, per item of the wave
    __shared__ float shv[HO ? W : 1][HO ? BW * 256 : 1];  // hand-off: the black items' column blocks
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, k = l & 15, q = l >> 4, g16 = q << 4;
    // this block's patch and slots, and its operator and grid: the launch's,
    // or (group launch) its slot group's
    int patch, npx, nx, ny, px, ls;
    bool slot_ok;
    const uint16_t* Pb;
    const float* Hg;
    long long pstride;
    float omega;
    // the previous bin's operator (hand-off): the launch's, or the slot group's
    const float *hVtb = a.hVt, *hHb = a.hH, *hdb = a.hd;
    const uint16_t* hPb = a.hP;
    long long hdstride = a.hdstride, hpstride = a.pstride;
    float homega = a.hepi_omega;
    if constexpr (GRP) {
        const int e = a.btab[blockIdx.x];
        if (e < 0) return;  // uniform over the block
        const GroupOp& go = a.gops[e >> 16];
        if constexpr (HO) {
            const GroupOp& hg = a.hgops[e >> 16];
            hVtb = (const float*)hg.Vt; hHb = (const float*)hg.H; hdb = (const float*)hg.d;
            hPb = (const uint16_t*)hg.P; hdstride = hg.dstride; hpstride = hg.pstride; homega = hg.omega;
        }
        patch = e & 0xffff;
        nx = go.nx; ny = go.ny; px = go.px;
        npx = (nx + PA - 1) / PA;
        Pb = (const uint16_t*)go.P; Hg = (const float*)go.H; pstride = go.pstride; omega = go.omega;
        slot_ok = q < go.count;
        ls = go.first + (slot_ok ? q : go.count - 1);
    } else {
        const int nsg = (a.nslots + 3) >> 2;
        int sg;
        xcd_tile_slot(blockIdx.x, nsg, a.tiles_per_xcd, patch, sg);
        if (patch >= a.npatches) return;  // uniform over the block
        npx = a.npx; nx = a.nx; ny = a.ny; px = a.px;
        Pb = a.P; Hg = a.H; pstride = a.pstride; omega = a.omega;
        const int slot = sg * 4 + q;
        slot_ok = slot < a.nslots;
        ls = slot_ok ? slot : a.nslots - 1;
    }
    const int pj = patch / npx, i0 = (patch - pj * npx) * PA, j0 = pj * PB;
    const int tr = a.tmap_identity ? ls : a.tmap[ls];
    const bool on = slot_ok && (!a.active || a.active[ls] != 0);
    // the epilogue at this slot's last sweep only
    const int cnt = a.counts ? a.counts[ls] : 0x7fffffff;
    const bool epi = a.epi_out && (cnt < a.nsweeps ? cnt : a.nsweeps) == 1;
    const float* __restrict__ b = a.b + (size_t)ls * a.bstride;
    float* __restrict__ x = a.x + (size_t)ls * a.xstride;
    const uint16_t* __restrict__ P = Pb + (size_t)tr * pstride;

    // ---- every load of the wave's items, red and black
    int rcol[RW], rpos[RW];
    bool rown[RW], rgo[RW];
    float rb[RW];
    uint16_t rp[RW][16];
#pragma unroll
    for (int u = 0; u < RW; ++u) {
        const int t = wv + W * u;
        const bool has = t < NR;
        const PatchCol c = patch_red<PA, PB>(has ? t : 0, i0, j0, nx, ny, px);
        rcol[u] = c.col;
        rpos[u] = c.pos;
        rown[u] = has && t < NRO && c.valid;
        rgo[u] = has;
        rb[u] = b[(size_t)c.col * 16 + k];
        const uint16_t* Pr = P + (size_t)c.col * 256 + k * 16;
#pragma unroll
        for (int m = 0; m < 16; ++m) rp[u][m] = Pr[m];
    }
    int bcol[BW], bq[BW][4];
    bool bgo[BW], bst[BW], bw_[BW], be_[BW], bs_[BW], bn_[BW];
    float bb[BW], bh[BW][4];
    uint16_t bp[BW][16];
#pragma unroll
    for (int u = 0; u < BW; ++u) {
        const int t = wv + W * u;
        const bool has = t < NB;
        const PatchCol c = patch_black<PA, PB>(has ? t : 0, i0, j0, nx, ny);
        bcol[u] = c.col;
        bgo[u] = has;
        bst[u] = has && c.valid;
        patch_black_nbrs<PA>(c.pos, c.i, i0, nx, bq[u]);
        bw_[u] = west_col(c.col, c.i, nx, px) >= 0;
        be_[u] = east_col(c.col, c.i, nx, px) >= 0;
        bs_[u] = c.j > 0;
        bn_[u] = c.j < ny - 1;
        bb[u] = b[(size_t)c.col * 16 + k];
        const uint16_t* Pr = P + (size_t)c.col * 256 + k * 16;
#pragma unroll
        for (int m = 0; m < 16; ++m) bp[u][m] = Pr[m];
        const float* Hc = Hg + (size_t)c.col * 64 + k;
#pragma unroll
        for (int e = 0; e < 4; ++e) bh[u][e] = Hc[16 * e];
    }
    // hand-off loads: the previous bin's rows at the black items' columns, in
    // the order and form of the four-slot half-sweep (k_halfsweep, CS = 1)
    float hbv[HO ? BW : 1], hxo[HO ? BW : 1], hnx[HO ? BW : 1][4], hdv[HO ? BW : 1], hhv[HO ? BW : 1][4];
    float hvr[HO ? BW : 1][4];
    uint16_t hpr[HO ? BW : 1][16];
    if constexpr (HO) {
        const float* __restrict__ hb = a.hb + (size_t)ls * a.hbstride;
        const float* __restrict__ hx = a.hx + (size_t)ls * a.hxstride;
        const uint16_t* __restrict__ hP = hPb + (size_t)tr * hpstride;
        const float* __restrict__ hd = hdb + (size_t)tr * hdstride;
#pragma unroll
        for (int u = 0; u < BW; ++u) {
            const int col = bcol[u];
            const size_t base = (size_t)col * 16 + k;
            const int ci = col % nx, cj = col / nx;
            const int cw = west_col(col, ci, nx, px), ce = east_col(col, ci, nx, px);
            hbv[u] = hb[base];
            hxo[u] = hx[base];
            hnx[u][0] = hx[(size_t)(cw >= 0 ? cw : col) * 16 + k];
            hnx[u][1] = hx[(size_t)(ce >= 0 ? ce : col) * 16 + k];
            hnx[u][2] = hx[(size_t)(cj > 0 ? col - nx : col) * 16 + k];
            hnx[u][3] = hx[(size_t)(cj < ny - 1 ? col + nx : col) * 16 + k];
            const uint16_t* Pr = hP + (size_t)col * 256 + k * 16;
#pragma unroll
            for (int m = 0; m < 16; ++m) hpr[u][m] = Pr[m];
            const float* Vt = hVtb + (size_t)col * 256 + l * 4;
#pragma unroll
            for (int qq = 0; qq < 4; ++qq) hvr[u][qq] = Vt[qq];
            hdv[u] = hd[base];
            const float* Hc = hHb + (size_t)col * 64 + k;
#pragma unroll
            for (int e = 0; e < 4; ++e) hhv[u][e] = Hc[16 * e];
        }
    }
    // a slot whose sweep count (group solves) ended the previous bin's solve
    // before its last sweep has no black half-sweep left there: its right-hand
    // side is the one in memory
    const bool hdo = !a.counts || a.nsweeps - 1 < cnt;
    // omega P v with v the wave's item u in sr: the half-sweep's product from x = 0
    auto apply_P = [&](const uint16_t* row, int u) {
        float dx = 0;
#pragma unroll
        for (int m = 0; m < 16; ++m) dx += (float)ldv(row[m]) * sr[wv][u][g16 + m];
        return (float)(0.0f + omega * dx);
    };
    auto store = [&](int col, float x1) {
        const size_t base = (size_t)col * 16 + k;
        x[base] = x1;
        if (epi) {
            const size_t ia = (size_t)ls * a.epi_as + base;
            const double ad = a.epi_add_f32 ? (double)((const float*)a.epi_add)[ia] : ((const double*)a.epi_add)[ia];
            a.epi_out[(size_t)ls * a.epi_os + base] = x1 + (float)ad;
        }
    };

    // ---- hand-off: the previous bin's last black half-sweep at the black
    // items' columns (k_halfsweep's arithmetic, so its bits); the new value is
    // stored as the previous bin's x and becomes this bin's right-hand side
    if constexpr (HO) {
#pragma unroll
        for (int u = 0; u < BW; ++u) {
            const int col = bcol[u];
            const int ci = col % nx, cj = col / nx;
            const int cw = west_col(col, ci, nx, px), ce = east_col(col, ci, nx, px);
            float r = hbv[u];
            float hx = 0;
            hx += hhv[u][0] * (cw >= 0 ? hnx[u][0] : 0.0f);
            hx += hhv[u][1] * (ce >= 0 ? hnx[u][1] : 0.0f);
            hx += hhv[u][2] * (cj > 0 ? hnx[u][2] : 0.0f);
            hx += hhv[u][3] * (cj < ny - 1 ? hnx[u][3] : 0.0f);
            r -= hx;
            const float xo = hxo[u];
            float* sv = shv[wv] + u * 256;
            sr[wv][u][l] = xo;
            float acc = hdv[u] * xo;
#pragma unroll
            for (int qq = 0; qq < 4; ++qq) sv[l * 4 + qq] = hvr[u][qq];
            wave_sync();
#pragma unroll
            for (int m = 0; m < 16; ++m) acc += sv[m * 16 + k] * sr[wv][u][g16 + m];
            r -= acc;
            wave_sync();
            sr[wv][u][l] = r;
            wave_sync();
            float dx = 0;
#pragma unroll
            for (int m = 0; m < 16; ++m) dx += (float)ldv(hpr[u][m]) * sr[wv][u][g16 + m];
            const float xn = (float)(xo + homega * dx);
            float bn = xn;
            const size_t base = (size_t)col * 16 + k;
            if (a.hepi_out) {
                const size_t ia = (size_t)ls * a.hepi_as + base;
                const double ad = a.hepi_add_f32 ? (double)((const float*)a.hepi_add)[ia]
                                                 : ((const double*)a.hepi_add)[ia];
                bn = xn + (float)ad;
            }
            if (bst[u] && on && hdo) {
                a.hx[(size_t)ls * a.hxstride + base] = xn;
                if (a.hepi_out) a.hepi_out[(size_t)ls * a.hepi_os + base] = bn;
            }
            if (hdo) bb[u] = bn;
        }
        wave_sync();
    }

    // ---- red: x_r = omega P_r b_r for the patch's reds and its ring reds
#pragma unroll
    for (int u = 0; u < RW; ++u) sr[wv][u][l] = rb[u];
    wave_sync();
#pragma unroll
    for (int u = 0; u < RW; ++u) {
        if (!rgo[u]) continue;  // uniform over the wave
        const float x1 = apply_P(rp[u], u);
        sgrid[rpos[u]][l] = x1;
        if (rown[u] && on) store(rcol[u], x1);
    }
    __syncthreads();
    // ---- black: x_b = omega P_b (b_b - H x_r) from the new reds
#pragma unroll
    for (int u = 0; u < BW; ++u) {
        float r = bb[u];
        float hx = 0;
        hx += bh[u][0] * (bw_[u] ? sgrid[bq[u][0]][l] : 0.0f);
        hx += bh[u][1] * (be_[u] ? sgrid[bq[u][1]][l] : 0.0f);
        hx += bh[u][2] * (bs_[u] ? sgrid[bq[u][2]][l] : 0.0f);
        hx += bh[u][3] * (bn_[u] ? sgrid[bq[u][3]][l] : 0.0f);
        r -= hx;
        sr[wv][u][l] = r;
    }
    wave_sync();
#pragma unroll
    for (int u = 0; u < BW; ++u) {
        if (!bgo[u]) continue;  // uniform over the wave
        const float x1 = apply_P(bp[u], u);
        if (bst[u] && on) store(bcol[u], x1);
    }
}

// y = A x (one operator apply, FP64), used for residual checks of the inner solves.
// VT: the transposes of the applied column blocks (row `lane` of a block read
// as element m at m * 16 + lane: contiguous bytes over a column's 16 lanes).
__global__ __launch_bounds__(256) void k_apply_A(const double* __restrict__ VT, long long vstride,
                                                 const double* __restrict__ H, long long hstride,
                                                 const double* __restrict__ d, long long dstride,
                                                 const double* __restrict__ x, double* __restrict__ y,
                                                 long long xstride, const int* tmap, int ncol, int nx,
                                                 int ny, int px) {
    __shared__ double sx[256];
    const int slot = blockIdx.y;
    const int lane = threadIdx.x & 15, grp = threadIdx.x >> 4;
    const int col0 = blockIdx.x * 16 + grp;
    const bool valid = col0 < ncol;
    const int col = valid ? col0 : 0;
    const int tr = tmap[slot];
    const double* xs = x + (size_t)slot * xstride;
    const double* Vt = VT + (size_t)tr * vstride;
    const double* Ht = H + (size_t)tr * hstride;
    const double* dt = d + (size_t)tr * dstride;
    const size_t base = (size_t)col * 16 + lane;
    sx[threadIdx.x] = xs[base];
    __syncthreads();
    double acc = dt[base] * sx[threadIdx.x];
    for (int m = 0; m < 16; ++m) acc += Vt[(size_t)col * 256 + m * 16 + lane] * sx[(grp << 4) + m];
    const int j = col / nx, i = col - j * nx;
    const double* Hc = Ht + (size_t)col * 64 + lane;
    const int cw = west_col(col, i, nx, px), ce = east_col(col, i, nx, px);
    if (cw >= 0) acc += Hc[0] * xs[(size_t)cw * 16 + lane];
    if (ce >= 0) acc += Hc[16] * xs[(size_t)ce * 16 + lane];
    if (j > 0) acc += Hc[32] * xs[(size_t)(col - nx) * 16 + lane];
    if (j < ny - 1) acc += Hc[48] * xs[(size_t)(col + nx) * 16 + lane];
    if (valid) y[(size_t)slot * xstride + base] = acc;
}

// ------------------------------------------------------------------------- //
// layout permutes and batched vector kernels
// ------------------------------------------------------------------------- //
// k-major (N,) per (slot, state) -> padded device layout, times a scale.
// The two layout permutes go through an LDS tile of 64 columns x 16 levels so
// both the k-major side (64 consecutive columns of one level) and the padded
// side (64 consecutive 16-level columns) are read and written in contiguous runs.
constexpr int PT_COLS = 64;
constexpr int PT_LD = 17;  // padded row of the tile (no LDS bank conflicts on the transpose)

// k-major (N,) per (slot, state) -> padded device layout, times a scale.
// grid: (ceil(ncol / 64), nslots * nstates), 256 threads.
__global__ __launch_bounds__(256) void k_ext_to_int(const double* __restrict__ src, long long sstride,
                                                    double* __restrict__ dst, long long dstride, int nstates,
                                                    int nz, int ncol, double scale) {
    __shared__ double tile[PT_COLS * PT_LD];
    const int s = blockIdx.y;  // slot*nstates + state
    const int slot = s / nstates, st = s - slot * nstates;
    const int col0 = blockIdx.x * PT_COLS;
    const double* in = src + (size_t)slot * sstride + (size_t)st * (size_t)nz * ncol;
    for (int e = threadIdx.x; e < PT_COLS * 16; e += blockDim.x) {
        const int k = e / PT_COLS, c = e - k * PT_COLS;
        double v = 0.0;
        if (k < nz && col0 + c < ncol) v = scale * in[(size_t)k * ncol + col0 + c];
        tile[c * PT_LD + k] = v;
    }
    __syncthreads();
    double* out = dst + (size_t)slot * dstride + (size_t)st * (size_t)ncol * 16 + (size_t)col0 * 16;
    for (int e = threadIdx.x; e < PT_COLS * 16; e += blockDim.x) {
        const int c = e >> 4, k = e & 15;
        if (col0 + c < ncol) out[e] = tile[c * PT_LD + k];
    }
}

// padded device layout -> k-major (N,) per (slot, state); same grid.
__global__ __launch_bounds__(256) void k_int_to_ext(const double* __restrict__ src, long long sstride,
                                                    double* __restrict__ dst, long long dstride, int nstates,
                                                    int nz, int ncol) {
    __shared__ double tile[PT_COLS * PT_LD];
    const int s = blockIdx.y;
    const int slot = s / nstates, st = s - slot * nstates;
    const int col0 = blockIdx.x * PT_COLS;
    const double* in = src + (size_t)slot * sstride + (size_t)st * (size_t)ncol * 16 + (size_t)col0 * 16;
    for (int e = threadIdx.x; e < PT_COLS * 16; e += blockDim.x) {
        const int c = e >> 4, k = e & 15;
        if (col0 + c < ncol) tile[c * PT_LD + k] = in[e];
    }
    __syncthreads();
    double* out = dst + (size_t)slot * dstride + (size_t)st * (size_t)nz * ncol;
    for (int e = threadIdx.x; e < PT_COLS * 16; e += blockDim.x) {
        const int k = e / PT_COLS, c = e - k * PT_COLS;
        if (k < nz && col0 + c < ncol) out[(size_t)k * ncol + col0 + c] = tile[c * PT_LD + k];
    }
}

// Group solves: slot s's k-major (nstates, N_s) input at src[s] -> its padded
// rows [s][state][n16max] (its ncol_s columns only), times scale; grid
// (ceil(max ncol / 64), nslots * nstates).
__global__ __launch_bounds__(256) void k_ext_to_int_g(const double* const* __restrict__ src, const int* __restrict__ sncol,
                                                      const int* __restrict__ snz, double* __restrict__ dst,
                                                      long long dstride, long long n16max, int nstates, double scale) {
    __shared__ double tile[PT_COLS * PT_LD];
    const int s = blockIdx.y;
    const int slot = s / nstates, st = s - slot * nstates;
    const int ncol = sncol[slot], nz = snz[slot];
    const int col0 = blockIdx.x * PT_COLS;
    if (col0 >= ncol) return;  // uniform over the block
    const double* in = src[slot] + (size_t)st * (size_t)nz * ncol;
    for (int e = threadIdx.x; e < PT_COLS * 16; e += blockDim.x) {
        const int k = e / PT_COLS, c = e - k * PT_COLS;
        double v = 0.0;
        if (k < nz && col0 + c < ncol) v = scale * in[(size_t)k * ncol + col0 + c];
        tile[c * PT_LD + k] = v;
    }
    __syncthreads();
    double* out = dst + (size_t)slot * dstride + (size_t)st * n16max + (size_t)col0 * 16;
    for (int e = threadIdx.x; e < PT_COLS * 16; e += blockDim.x) {
        const int c = e >> 4, k = e & 15;
        if (col0 + c < ncol) out[e] = tile[c * PT_LD + k];
    }
}

// Group solves: padded rows [s][state][n16max] -> slot s's k-major
// (nstates, N_s) output at dst[s]; same grid.
__global__ __launch_bounds__(256) void k_int_to_ext_g(const double* __restrict__ src, long long sstride,
                                                      long long n16max, double* const* __restrict__ dst,
                                                      const int* __restrict__ sncol, const int* __restrict__ snz,
                                                      int nstates) {
    __shared__ double tile[PT_COLS * PT_LD];
    const int s = blockIdx.y;
    const int slot = s / nstates, st = s - slot * nstates;
    const int ncol = sncol[slot], nz = snz[slot];
    const int col0 = blockIdx.x * PT_COLS;
    if (col0 >= ncol) return;  // uniform over the block
    const double* in = src + (size_t)slot * sstride + (size_t)st * n16max + (size_t)col0 * 16;
    for (int e = threadIdx.x; e < PT_COLS * 16; e += blockDim.x) {
        const int c = e >> 4, k = e & 15;
        if (col0 + c < ncol) tile[c * PT_LD + k] = in[e];
    }
    __syncthreads();
    double* out = dst[slot] + (size_t)st * (size_t)nz * ncol;
    for (int e = threadIdx.x; e < PT_COLS * 16; e += blockDim.x) {
        const int k = e / PT_COLS, c = e - k * PT_COLS;
        if (k < nz && col0 + c < ncol) out[(size_t)k * ncol + col0 + c] = tile[c * PT_LD + k];
    }
}

inline dim3 grid_permute(int ncol, int nstreams) { return dim3((unsigned)((ncol + PT_COLS - 1) / PT_COLS), nstreams); }

// Per-slot vector lengths of a group solve (slots of operator sets of
// different sizes in buffers of the largest): element g of slot s counts only
// where g % seg < len[s]; the rest is padding, never read or written.  A null
// len (single-set solves) makes every element count.
struct SlotLen {
    const int* len = nullptr;
    int seg = 1;
};
__device__ __forceinline__ bool in_slot(const SlotLen& L, int slot, long long g) {
    return !L.len || (int)(g % L.seg) < L.len[slot];
}

// out[slot] = a[slot] + alpha * b[slot]   (strides in elements)
__global__ void k_axpby(const double* __restrict__ a, long long as, const double* __restrict__ b,
                        long long bs, double alpha, double* __restrict__ out, long long os, int n,
                        const int* active, SlotLen sl) {
    const int slot = blockIdx.y;
    if (active && !active[slot]) return;
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n || !in_slot(sl, slot, g)) return;
    out[(size_t)slot * os + g] = a[(size_t)slot * as + g] + alpha * b[(size_t)slot * bs + g];
}

template <class S, class D>
__global__ void k_convert_slots(const S* __restrict__ src, long long ss, D* __restrict__ dst,
                                long long ds, int n, const int* active, SlotLen sl) {
    const int slot = blockIdx.y;
    if (active && !active[slot]) return;
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n || !in_slot(sl, slot, g)) return;
    dst[(size_t)slot * ds + g] = (D)src[(size_t)slot * ss + g];
}

// w = v - (double)u  (u the FP32 or FP64 result of the monodromy apply)
template <class U>
__global__ void k_sub_from(const double* __restrict__ v, long long vs, const U* __restrict__ u,
                           long long us, double* __restrict__ w, long long ws, int n,
                           const int* active, SlotLen sl) {
    const int slot = blockIdx.y;
    if (active && !active[slot]) return;
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n || !in_slot(sl, slot, g)) return;
    w[(size_t)slot * ws + g] = v[(size_t)slot * vs + g] - (double)u[(size_t)slot * us + g];
}

// ------------------------------------------------------------------------- //
// deterministic reductions: per-block partials, then a fixed-order sum
// ------------------------------------------------------------------------- //
constexpr int RED_THREADS = 256;
constexpr int RED_ELEMS = 4;
constexpr int RED_CHUNK = RED_THREADS * RED_ELEMS;

__device__ __forceinline__ double block_sum(double v, double* sh) {
    // wave64 shuffle reduction then across the 4 waves
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    return r;
}

// partials[(slot*nvec + i)*nb + blk] = sum_n V_i[n] * w[n]   for i < nvec
__global__ __launch_bounds__(RED_THREADS) void k_mdot(const double* __restrict__ Vb, long long vslot,
                                                      int nvec, const double* __restrict__ w,
                                                      long long wslot, int n, int nb,
                                                      double* __restrict__ partials,
                                                      const int* active, SlotLen sl) {
    __shared__ double sh[4];
    const int slot = blockIdx.y, blk = blockIdx.x;
    if (active && !active[slot]) return;
    const int nn = sl.len ? min(n, sl.len[slot]) : n;  // (single vectors: seg = n)
    const double* ws = w + (size_t)slot * wslot;
    const double* Vs = Vb + (size_t)slot * vslot;
    double wv[RED_ELEMS];
    int idx[RED_ELEMS];
#pragma unroll
    for (int e = 0; e < RED_ELEMS; ++e) {
        idx[e] = blk * RED_CHUNK + e * RED_THREADS + threadIdx.x;
        wv[e] = idx[e] < nn ? ws[idx[e]] : 0.0;
    }
    for (int i = 0; i < nvec; ++i) {
        const double* vi = Vs + (size_t)i * n;
        double acc = 0.0;
#pragma unroll
        for (int e = 0; e < RED_ELEMS; ++e)
            if (idx[e] < nn) acc += vi[idx[e]] * wv[e];
        double s = block_sum(acc, sh);
        if (threadIdx.x == 0) partials[((size_t)slot * nvec + i) * nb + blk] = s;
    }
}

// coef[slot*MAXM1 + i] = sum over blocks of partials (fixed order)
constexpr int MAXM1 = MAXM + 1;
// Sum of nb partials by one wavefront: lane l adds partials l, l+64, ... in
// order, then a fixed butterfly over the lanes (the result is in lane 0).  The
// order depends only on nb, so every slot's sum is deterministic.
__device__ __forceinline__ double wave_sum_partials(const double* __restrict__ p, int nb) {
    double s = 0.0;
    for (int b = threadIdx.x & 63; b < nb; b += 64) s += p[b];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    return s;
}
// blocks of up to 4 waves, one wave per output value
inline dim3 grid_waves(size_t n) { return dim3((unsigned)((n + 3) / 4)); }

__global__ __launch_bounds__(256) void k_finalize(const double* __restrict__ partials, int nvec, int nb,
                                                  double* __restrict__ out, int nslots) {
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= nslots * nvec) return;  // uniform over the wave
    const int slot = w / nvec, i = w - slot * nvec;
    const double s = wave_sum_partials(partials + ((size_t)slot * nvec + i) * nb, nb);
    if ((threadIdx.x & 63) == 0) out[(size_t)slot * MAXM1 + i] = s;
}

// w -= sum_i coef_i V_i ; optionally partials of ||w||^2
__global__ __launch_bounds__(RED_THREADS) void k_mupdate(const double* __restrict__ Vb, long long vslot,
                                                         int nvec, const double* __restrict__ coef,
                                                         double sign, double* __restrict__ w,
                                                         long long wslot, int n, int nb,
                                                         double* __restrict__ npartials,
                                                         const int* active, const int* counts, SlotLen sl) {
    __shared__ double sh[4];
    const int slot = blockIdx.y, blk = blockIdx.x;
    if (active && !active[slot]) return;
    const int nn = sl.len ? min(n, sl.len[slot]) : n;
    if (counts) nvec = counts[slot];  // only the slot's own valid basis vectors
    double* ws = w + (size_t)slot * wslot;
    const double* Vs = Vb + (size_t)slot * vslot;
    const double* c = coef + (size_t)slot * MAXM1;
    double acc2 = 0.0;
#pragma unroll
    for (int e = 0; e < RED_ELEMS; ++e) {
        int ix = blk * RED_CHUNK + e * RED_THREADS + threadIdx.x;
        if (ix < nn) {
            double v = ws[ix];
            for (int i = 0; i < nvec; ++i) v -= sign * c[i] * Vs[(size_t)i * n + ix];
            ws[ix] = v;
            acc2 += v * v;
        }
    }
    if (npartials) {
        double s = block_sum(acc2, sh);
        if (threadIdx.x == 0) npartials[(size_t)slot * nb + blk] = s;
    }
}

// partials of ||x||^2 per slot
__global__ __launch_bounds__(RED_THREADS) void k_norm2(const double* __restrict__ x, long long xslot,
                                                       int n, int nb, double* __restrict__ partials, SlotLen sl) {
    __shared__ double sh[4];
    const int slot = blockIdx.y, blk = blockIdx.x;
    const int nn = sl.len ? min(n, sl.len[slot]) : n;
    const double* xs = x + (size_t)slot * xslot;
    double acc = 0.0;
#pragma unroll
    for (int e = 0; e < RED_ELEMS; ++e) {
        int ix = blk * RED_CHUNK + e * RED_THREADS + threadIdx.x;
        if (ix < nn) acc += xs[ix] * xs[ix];
    }
    double s = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[(size_t)slot * nb + blk] = s;
}

// partial counts of negative entries per slot over nstates states
__global__ __launch_bounds__(RED_THREADS) void k_negcount(const double* __restrict__ x, long long xslot,
                                                          int n, int nb, double* __restrict__ partials, SlotLen sl) {
    __shared__ double sh[4];
    const int slot = blockIdx.y, blk = blockIdx.x;
    const double* xs = x + (size_t)slot * xslot;
    double acc = 0.0;
#pragma unroll
    for (int e = 0; e < RED_ELEMS; ++e) {
        int ix = blk * RED_CHUNK + e * RED_THREADS + threadIdx.x;
        if (ix < n && in_slot(sl, slot, ix) && xs[ix] < 0.0) acc += 1.0;
    }
    double s = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[(size_t)slot * nb + blk] = s;
}

__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ partials, int nb,
                                                      double* __restrict__ out, int nslots) {
    const int slot = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (slot >= nslots) return;  // uniform over the wave
    const double s = wave_sum_partials(partials + (size_t)slot * nb, nb);
    if ((threadIdx.x & 63) == 0) out[slot] = s;
}

// ------------------------------------------------------------------------- //
// GMRES state per slot (Hessenberg, Givens, residual vector g)
// ------------------------------------------------------------------------- //
struct GState {
    double H[MAXM1 * MAXM];  // column-major: H[j*MAXM1 + i]
    double cs[MAXM], sn[MAXM];
    double g[MAXM1];
    double y[MAXM];
    double beta;
    double resid;
    double hnext;
    double ynorm;  // ||y|| = ||delta|| of the last cycle (orthonormal basis)
};

__global__ void k_gather_ynorm(const GState* st, int nslots, double* out) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot < nslots) out[slot] = st[slot].ynorm;
}

// beta = ||r||, g = beta e1; the caller then scales V_0 = r / beta.
__global__ void k_gmres_start(GState* st, const double* norm2, int nslots, const int* active) {
    int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= nslots) return;
    if (active && !active[slot]) return;
    GState& s = st[slot];
    double beta = sqrt(norm2[slot]);
    s.beta = beta;
    for (int i = 0; i < MAXM1; ++i) s.g[i] = 0.0;
    s.g[0] = beta;
    s.resid = beta;
}

// out[slot] = in[slot] * (1/scale[slot]) where scale is read from the state
__global__ void k_scale_by(const double* __restrict__ in, long long is, double* __restrict__ out,
                           long long os, const GState* st, int which, int n, const int* active, SlotLen sl) {
    const int slot = blockIdx.y;
    if (active && !active[slot]) return;
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n || !in_slot(sl, slot, g)) return;
    double s = which == 0 ? st[slot].beta : st[slot].hnext;
    double inv = s > 0.0 ? 1.0 / s : 0.0;
    out[(size_t)slot * os + g] = in[(size_t)slot * is + g] * inv;
}

// Column j of H = h1 + h2 (CGS2 coefficients); h_{j+1,j} = ||w||; Givens update.
__global__ void k_gmres_column(GState* st, const double* __restrict__ h1, const double* __restrict__ h2,
                               const double* __restrict__ wnorm2, int j, int nslots, const int* active,
                               double* resid_out) {
    int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= nslots) return;
    GState& s = st[slot];
    if (active && !active[slot]) {
        resid_out[slot] = s.resid;
        return;
    }
    double* col = s.H + (size_t)j * MAXM1;
    for (int i = 0; i <= j; ++i) col[i] = h1[(size_t)slot * MAXM1 + i] + h2[(size_t)slot * MAXM1 + i];
    double hn = sqrt(wnorm2[slot]);
    col[j + 1] = hn;
    s.hnext = hn;
    for (int i = 0; i < j; ++i) {
        double a = col[i], b = col[i + 1];
        col[i] = s.cs[i] * a + s.sn[i] * b;
        col[i + 1] = -s.sn[i] * a + s.cs[i] * b;
    }
    double a = col[j], b = col[j + 1];
    double r = hypot(a, b);
    double c = r > 0.0 ? a / r : 1.0, sn = r > 0.0 ? b / r : 0.0;
    s.cs[j] = c;
    s.sn[j] = sn;
    col[j] = r;
    col[j + 1] = 0.0;
    s.g[j + 1] = -sn * s.g[j];
    s.g[j] = c * s.g[j];
    s.resid = fabs(s.g[j + 1]);
    resid_out[slot] = s.resid;
}

// y = R^{-1} g over the first k columns (k per slot), written to y and to coef
__global__ void k_gmres_solve(GState* st, const int* kcols, int nslots, double* coef) {
    int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= nslots) return;
    GState& s = st[slot];
    int k = kcols[slot];
    for (int i = 0; i < MAXM1; ++i) coef[(size_t)slot * MAXM1 + i] = 0.0;
    for (int i = k - 1; i >= 0; --i) {
        double v = s.g[i];
        for (int c = i + 1; c < k; ++c) v -= s.H[(size_t)c * MAXM1 + i] * s.y[c];
        double dgl = s.H[(size_t)i * MAXM1 + i];
        s.y[i] = dgl != 0.0 ? v / dgl : 0.0;
    }
    double yn = 0.0;
    for (int i = 0; i < k; ++i) {
        coef[(size_t)slot * MAXM1 + i] = s.y[i];
        yn += s.y[i] * s.y[i];
    }
    s.ynorm = sqrt(yn);
}

// ------------------------------------------------------------------------- //
// fused Arnoldi-step kernels: the same arithmetic and summation orders as the
// separate kernels above, fewer launches per step
// ------------------------------------------------------------------------- //
// w = v - u (the new Krylov direction, stored) and the partials of V_i . w
// (k_sub_from then k_mdot in one pass over w).
template <class U>
__global__ __launch_bounds__(RED_THREADS) void k_mdot_sub(const double* __restrict__ Vb, long long vslot, int nvec,
                                                          const double* __restrict__ v, long long vs,
                                                          const U* __restrict__ u, long long us,
                                                          double* __restrict__ w, long long wslot, int n, int nb,
                                                          double* __restrict__ partials, const int* active,
                                                          SlotLen sl) {
    __shared__ double sh[4];
    const int slot = blockIdx.y, blk = blockIdx.x;
    if (active && !active[slot]) return;
    const int nn = sl.len ? min(n, sl.len[slot]) : n;
    const double* Vs = Vb + (size_t)slot * vslot;
    double wv[RED_ELEMS];
    int idx[RED_ELEMS];
#pragma unroll
    for (int e = 0; e < RED_ELEMS; ++e) {
        idx[e] = blk * RED_CHUNK + e * RED_THREADS + threadIdx.x;
        wv[e] = 0.0;
        if (idx[e] < nn) {
            wv[e] = v[(size_t)slot * vs + idx[e]] - (double)u[(size_t)slot * us + idx[e]];
            w[(size_t)slot * wslot + idx[e]] = wv[e];
        }
    }
    for (int i = 0; i < nvec; ++i) {
        const double* vi = Vs + (size_t)i * n;
        double acc = 0.0;
#pragma unroll
        for (int e = 0; e < RED_ELEMS; ++e)
            if (idx[e] < nn) acc += vi[idx[e]] * wv[e];
        double s = block_sum(acc, sh);
        if (threadIdx.x == 0) partials[((size_t)slot * nvec + i) * nb + blk] = s;
    }
}

// w -= sum_i c_i V_i with c_i summed here from k_mdot's partials in k_finalize's
// order (every block the same sums; block 0 also stores them in coef_out), and
// optionally the partials of ||w||^2 (k_finalize then k_mupdate in one launch).
__global__ __launch_bounds__(RED_THREADS) void k_mupdate_fin(const double* __restrict__ Vb, long long vslot,
                                                             int nvec, const double* __restrict__ partials,
                                                             double* __restrict__ coef_out, double* __restrict__ w,
                                                             long long wslot, int n, int nb,
                                                             double* __restrict__ npartials, const int* active,
                                                             SlotLen sl) {
    __shared__ double sh[4];
    __shared__ double c[MAXM1];
    const int slot = blockIdx.y, blk = blockIdx.x;
    if (active && !active[slot]) return;
    for (int i = threadIdx.x >> 6; i < nvec; i += RED_THREADS / 64) {
        const double s = wave_sum_partials(partials + ((size_t)slot * nvec + i) * nb, nb);
        if ((threadIdx.x & 63) == 0) {
            c[i] = s;
            if (blk == 0) coef_out[(size_t)slot * MAXM1 + i] = s;
        }
    }
    __syncthreads();
    const int nn = sl.len ? min(n, sl.len[slot]) : n;
    double* ws = w + (size_t)slot * wslot;
    const double* Vs = Vb + (size_t)slot * vslot;
    const double sign = 1.0;
    double acc2 = 0.0;
#pragma unroll
    for (int e = 0; e < RED_ELEMS; ++e) {
        int ix = blk * RED_CHUNK + e * RED_THREADS + threadIdx.x;
        if (ix < nn) {
            double v = ws[ix];
            for (int i = 0; i < nvec; ++i) v -= sign * c[i] * Vs[(size_t)i * n + ix];
            ws[ix] = v;
            acc2 += v * v;
        }
    }
    if (npartials) {
        double s = block_sum(acc2, sh);
        if (threadIdx.x == 0) npartials[(size_t)slot * nb + blk] = s;
    }
}

// The end of Arnoldi step j: ||w||^2 summed from its partials (k_sum_partials'
// order) by every block; block 0 of a slot updates its Hessenberg column and
// Givens rotation (k_gmres_column, active slots) or reports its residual
// (inactive ones); every block of an active slot writes its part of
// V_{j+1} = w / ||w|| (k_scale_by).
__global__ __launch_bounds__(256) void k_arnoldi_tail(GState* st, const double* __restrict__ h1,
                                                      const double* __restrict__ h2,
                                                      const double* __restrict__ npartials, int nb, int j,
                                                      const int* active, double* resid_out,
                                                      const double* __restrict__ w, long long ws,
                                                      double* __restrict__ vnext, long long vs, int n, SlotLen sl,
                                                      float* __restrict__ fnext, long long fs) {
    __shared__ double hn_s;
    const int slot = blockIdx.y;
    const bool on = !active || active[slot];
    if (!on) {
        if (blockIdx.x == 0 && threadIdx.x == 0) resid_out[slot] = st[slot].resid;
        return;
    }
    if (threadIdx.x < 64) {
        const double s = wave_sum_partials(npartials + (size_t)slot * nb, nb);
        if (threadIdx.x == 0) {
            hn_s = sqrt(s);
            if (blockIdx.x == 0) {
                GState& S = st[slot];
                double* col = S.H + (size_t)j * MAXM1;
                for (int i = 0; i <= j; ++i)
                    col[i] = h1[(size_t)slot * MAXM1 + i] + (h2 ? h2[(size_t)slot * MAXM1 + i] : 0.0);
                const double hn = hn_s;
                col[j + 1] = hn;
                S.hnext = hn;
                for (int i = 0; i < j; ++i) {
                    double a = col[i], b = col[i + 1];
                    col[i] = S.cs[i] * a + S.sn[i] * b;
                    col[i + 1] = -S.sn[i] * a + S.cs[i] * b;
                }
                double a = col[j], b = col[j + 1];
                double r = hypot(a, b);
                double cc = r > 0.0 ? a / r : 1.0, sn = r > 0.0 ? b / r : 0.0;
                S.cs[j] = cc;
                S.sn[j] = sn;
                col[j] = r;
                col[j + 1] = 0.0;
                S.g[j + 1] = -sn * S.g[j];
                S.g[j] = cc * S.g[j];
                S.resid = fabs(S.g[j + 1]);
                resid_out[slot] = S.resid;
            }
        }
    }
    __syncthreads();
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n || !in_slot(sl, slot, g)) return;
    const double s = hn_s;
    const double inv = s > 0.0 ? 1.0 / s : 0.0;
    const double val = w[(size_t)slot * ws + g] * inv;
    vnext[(size_t)slot * vs + g] = val;
    if (fnext) fnext[(size_t)slot * fs + g] = (float)val;  // the inexact monodromy's FP32 input
}

// The start of a cycle: beta = ||r|| summed from k_norm2's partials by every
// block; block 0 of a slot sets g = beta e1 (k_gmres_start); every block of an
// active slot writes its part of V_0 = r / beta (k_scale_by).
__global__ __launch_bounds__(256) void k_gmres_start_scale(GState* st, const double* __restrict__ npartials, int nb,
                                                           const int* active, const double* __restrict__ r,
                                                           long long rs, double* __restrict__ v0, long long vs, int n,
                                                           SlotLen sl, float* __restrict__ f0, long long fs) {
    __shared__ double beta_s;
    const int slot = blockIdx.y;
    if (active && !active[slot]) return;
    if (threadIdx.x < 64) {
        const double s = wave_sum_partials(npartials + (size_t)slot * nb, nb);
        if (threadIdx.x == 0) {
            beta_s = sqrt(s);
            if (blockIdx.x == 0) {
                GState& S = st[slot];
                S.beta = beta_s;
                for (int i = 0; i < MAXM1; ++i) S.g[i] = 0.0;
                S.g[0] = beta_s;
                S.resid = beta_s;
            }
        }
    }
    __syncthreads();
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n || !in_slot(sl, slot, g)) return;
    const double s = beta_s;
    const double inv = s > 0.0 ? 1.0 / s : 0.0;
    const double val = r[(size_t)slot * rs + g] * inv;
    v0[(size_t)slot * vs + g] = val;
    if (f0) f0[(size_t)slot * fs + g] = (float)val;
}

static inline dim3 grid1(size_t n, int threads = 256) { return dim3((unsigned)((n + threads - 1) / threads)); }

}  // namespace orbit

// ========================================================================= //
// host side
// ========================================================================= //
using namespace orbit;


extern "C" const char* orbit_last_error(void) { return g_err.c_str(); }

extern "C" void orbit_opts_default(orbit_opts* o) {
    o->tol = 1e-8;
    o->maxiter = 200;
    o->restart = 30;
    o->precision = ORBIT_PREC_FP64;
    o->inner_sweeps = 0;
    o->acc_sweeps = 0;
    o->use_graph = 0;
    o->verbose = 0;
    o->max_rounds = 0;
    o->rhs_is_s = 0;
    o->warm_orbit = 0;
    o->check_sweeps = 0;
}

extern "C" int orbit_ctx_create(int device, orbit_ctx** out) {
    if (!out) return fail(ORBIT_ERR_INVALID, "orbit_ctx_create: out is NULL");
    int n = 0;
    ORBIT_HIP_CHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return fail(ORBIT_ERR_INVALID, "orbit_ctx_create: device " + std::to_string(device) +
                                           " out of range (" + std::to_string(n) + " devices)");
    ORBIT_HIP_CHECK(hipSetDevice(device));
    auto* c = new orbit_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(ORBIT_ERR_HIP, "orbit_ctx_create: hipStreamCreate failed");
    }
    // set-up of the allocation-free solve: this device's scratch address (a
    // failure here is reported by that entry point, which looks again)
    if (fixed_scratch_address(device, &c->fixed_scratch) != ORBIT_OK) c->fixed_scratch = nullptr;
    *out = c;
    return ORBIT_OK;
}

extern "C" int orbit_ctx_destroy(orbit_ctx* c) {
    if (!c) return ORBIT_OK;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    for (auto& s : c->scratch)
        if (s.p) hipFree(s.p);
    for (void* p : c->owned) hipFree(p);
    if (c->pinned) hipHostFree(c->pinned);
    hipStreamDestroy(c->stream);
    delete c;
    return ORBIT_OK;
}

extern "C" int orbit_ctx_sync(orbit_ctx* c) {
    if (!c) return fail(ORBIT_ERR_INVALID, "ctx is NULL");
    ORBIT_HIP_CHECK(hipStreamSynchronize(c->stream));
    return ORBIT_OK;
}

extern "C" void* orbit_ctx_stream(orbit_ctx* c) { return c ? (void*)c->stream : nullptr; }

// ------------------------------------------------------------------------- //
// opset construction
// ------------------------------------------------------------------------- //
int calibrate_opset(orbit_ctx* ctx, orbit_opset* op, double omega);

namespace orbit {
// Default sweep launch shapes: reduced-precision sweeps of four or more slots
// share each column's loads over four slots of a wave (the shared block needs a
// single transport group); 256 threads, XCD-contiguous tiles.
void default_shapes(orbit_opset* op) {
    for (int p = 0; p < N_SWEEP_PRECS; ++p)
        for (int b = 0; b < N_SLOT_BUCKETS; ++b) {
            SweepShape& sh = op->shapes[p][b];
            sh.cs = (p != 0 /* FP64 */ && b >= 2 && op->ngroups == 1) ? 1 : 0;
            sh.threads = 256;
            sh.xcd = 1;
        }
}

}  // namespace orbit

namespace orbit {

int alloc_bins(orbit_opset* op) {
    const int ncol = op->ncol, n16 = op->n16, G = op->ngroups, T = op->ntr;
    const size_t nV = (size_t)G * ncol * 256, nH = (size_t)G * ncol * 64, nd = (size_t)T * n16,
                 nP = (size_t)T * ncol * 256;
    op->bins.resize(op->nbins);
    for (int b = 0; b < op->nbins; ++b) {
        BinOps& B = op->bins[b];
        B.V64 = op->alloc<double>(nV); B.V32 = op->alloc<float>(nV);
        B.VT64 = op->alloc<double>(nV); B.VT32 = op->alloc<float>(nV);
        B.H64 = op->alloc<double>(nH); B.H32 = op->alloc<float>(nH);
        B.HT64 = op->alloc<double>(nH); B.HT32 = op->alloc<float>(nH);
        B.d64 = op->alloc<double>(nd); B.d32 = op->alloc<float>(nd);
        B.P64 = op->alloc<double>(nP); B.P16 = op->alloc<uint16_t>(nP);
        B.PT64 = op->alloc<double>(nP); B.PT16 = op->alloc<uint16_t>(nP);
        if (!B.V64 || !B.V32 || !B.VT64 || !B.VT32 || !B.H64 || !B.H32 || !B.HT64 || !B.HT32 ||
            !B.d64 || !B.d32 || !B.P64 || !B.P16 || !B.PT64 || !B.PT16)
            return fail(ORBIT_ERR_NOMEM, "opset: device allocation failed");
        if (op->has_tdiag) {
            B.tdiag = op->alloc<double>((size_t)n16);
            if (!B.tdiag) return fail(ORBIT_ERR_NOMEM, "opset: device allocation failed");
        }
    }
    return ORBIT_OK;
}

// FP32 copies, transposes, factored column blocks (FP64 and BF16) of one bin
// whose V64, H64 and d64 are in place.
int derive_bin(orbit_ctx* ctx, orbit_opset* op, int b) {
    hipStream_t st = ctx->stream;
    const int ncol = op->ncol, n16 = op->n16, G = op->ngroups, T = op->ntr;
    const size_t nV = (size_t)G * ncol * 256, nH = (size_t)G * ncol * 64, nd = (size_t)T * n16,
                 nP = (size_t)T * ncol * 256;
    BinOps& B = op->bins[b];
    hipLaunchKernelGGL((k_convert<double, float>), grid1(nV), dim3(256), 0, st, B.V64, B.V32, nV);
    hipLaunchKernelGGL((k_convert<double, float>), grid1(nH), dim3(256), 0, st, B.H64, B.H32, nH);
    hipLaunchKernelGGL((k_convert<double, float>), grid1(nd), dim3(256), 0, st, B.d64, B.d32, nd);
    hipLaunchKernelGGL(k_transpose_blocks<double>, grid1(nV), dim3(256), 0, st, B.V64, B.VT64, (size_t)G * ncol);
    hipLaunchKernelGGL((k_convert<double, float>), grid1(nV), dim3(256), 0, st, B.VT64, B.VT32, nV);
    for (int g = 0; g < G; ++g) {
        hipLaunchKernelGGL(k_transpose_stencil<double>, grid1((size_t)ncol * 16), dim3(256), 0, st,
                           B.H64 + (size_t)g * ncol * 64, B.HT64 + (size_t)g * ncol * 64, op->nx, op->ny, op->px);
    }
    hipLaunchKernelGGL((k_convert<double, float>), grid1(nH), dim3(256), 0, st, B.HT64, B.HT32, nH);
    const long long vstride = G == 1 ? 0 : (long long)ncol * 256;
    hipLaunchKernelGGL(k_factor, dim3((ncol + 3) / 4, T), dim3(64), 0, st, B.V64, vstride, B.d64, (long long)n16,
                       B.P64, (long long)ncol * 256, ncol);
    hipLaunchKernelGGL(k_to_bf16, grid1(nP), dim3(256), 0, st, B.P64, B.P16, nP);
    hipLaunchKernelGGL(k_transpose_blocks<double>, grid1(nP), dim3(256), 0, st, B.P64, B.PT64, (size_t)T * ncol);
    hipLaunchKernelGGL(k_to_bf16, grid1(nP), dim3(256), 0, st, B.PT64, B.PT16, nP);
    ORBIT_HIP_CHECK(hipGetLastError());
    return ORBIT_OK;
}

void default_shapes(orbit_opset* op);

// Colour counts, default launch shapes, then the sweep calibration.
int finalize_opset(orbit_ctx* ctx, orbit_opset* op) {
    // columns of each colour; the sweeps map a colour index to its column
    // arithmetically.  A periodic x with an odd nx (>= 3) keeps the last column
    // of every row out of the red-black pattern (colours 2 and 3).
    const bool four = op->px && op->nx >= 3 && (op->nx % 2);
    op->ncolours = four ? 4 : 2;
    op->nxc = four ? op->nx - 1 : op->nx;
    for (int c = 0; c < 4; ++c) op->ncols_c[c] = 0;
    for (int j = 0; j < op->ny; ++j)
        for (int i = 0; i < op->nx; ++i)
            op->ncols_c[i < op->nxc ? (i + j) % 2 : 2 + j % 2] += 1;
    default_shapes(op);
    ORBIT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return calibrate_opset(ctx, op, 0.0);
}

}  // namespace orbit

namespace {

struct HostBin {
    std::vector<double> V;  // ngroups * ncol * 256
    std::vector<double> H;  // ngroups * ncol * 64
    std::vector<double> d;  // ntr * ncol * 16
    std::vector<double> tdiag;
};

// Parse one CSR matrix of L into the structured form.  Returns 0 or an error.
// Couplings across the x wrap face (i = 0 <-> i = nx - 1, the reference's
// periodic longitude, orbit/core/convdiff.py:150-200) go to the W / E slots and
// set *wrap; with nx <= 2 they coincide with the ordinary neighbours.
int parse_csr(int nz, int ny, int nx, double dtau, const int32_t* rp, const int32_t* ci,
              const double* v, double* V, double* H, double* diag, int* wrap) {
    const int ncol = nx * ny, N = nz * ncol;
    for (int c = 0; c < ncol; ++c)
        for (int k = 0; k < NZP; ++k) V[(size_t)c * 256 + k * 16 + k] = 1.0;
    for (int row = 0; row < N; ++row) {
        const int k = row / ncol, col = row - k * ncol;
        const int j = col / nx, i = col - j * nx;
        if (rp[row + 1] < rp[row]) return fail(ORBIT_ERR_STRUCTURE, "CSR row pointers decrease");
        for (int p = rp[row]; p < rp[row + 1]; ++p) {
            const int c2 = ci[p];
            if (c2 < 0 || c2 >= N) return fail(ORBIT_ERR_STRUCTURE, "CSR column index out of range");
            const int k2 = c2 / ncol, col2 = c2 - k2 * ncol;
            const double a = dtau * v[p];
            if (col2 == col) {
                if (k2 == k)
                    diag[(size_t)col * 16 + k] += a;
                else
                    V[(size_t)col * 256 + k * 16 + k2] += a;
                continue;
            }
            if (k2 != k) goto bad;
            {
                const int j2 = col2 / nx, i2 = col2 - j2 * nx;
                int dir = -1;
                if (j2 == j && i2 == i - 1) dir = 0;
                else if (j2 == j && i2 == i + 1) dir = 1;
                else if (i2 == i && j2 == j - 1) dir = 2;
                else if (i2 == i && j2 == j + 1) dir = 3;
                else if (j2 == j && nx > 2 && i == 0 && i2 == nx - 1) dir = 0, *wrap = 1;
                else if (j2 == j && nx > 2 && i == nx - 1 && i2 == 0) dir = 1, *wrap = 1;
                if (dir < 0) goto bad;
                H[(size_t)col * 64 + dir * 16 + k] += a;
                continue;
            }
        bad:
            return fail(ORBIT_ERR_STRUCTURE,
                        "CSR entry (" + std::to_string(row) + "," + std::to_string(c2) +
                            ") couples cells that are neither in one column nor horizontal "
                            "neighbours on one level");
        }
    }
    return ORBIT_OK;
}

int upload(hipStream_t st, void* dst, const void* src, size_t bytes) {
    ORBIT_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
    return ORBIT_OK;
}

int finish_opset(orbit_ctx* ctx, orbit_opset* op, std::vector<HostBin>& hb) {
    hipStream_t st = ctx->stream;
    const int ncol = op->ncol, n16 = op->n16, G = op->ngroups, T = op->ntr;
    const size_t nV = (size_t)G * ncol * 256, nH = (size_t)G * ncol * 64, nd = (size_t)T * n16;
    int rc = alloc_bins(op);
    if (rc) return rc;
    for (int b = 0; b < op->nbins; ++b) {
        BinOps& B = op->bins[b];
        if (op->has_tdiag) {
            rc = upload(st, B.tdiag, hb[b].tdiag.data(), (size_t)n16 * 8);
            if (rc) return rc;
        }
        rc = upload(st, B.V64, hb[b].V.data(), nV * 8);
        if (!rc) rc = upload(st, B.H64, hb[b].H.data(), nH * 8);
        if (!rc) rc = upload(st, B.d64, hb[b].d.data(), nd * 8);
        if (!rc) rc = derive_bin(ctx, op, b);
        if (rc) return rc;
    }
    return finalize_opset(ctx, op);
}

int check_dims(int nz, int ny, int nx, int nbins, int ntr) {
    if (nz < 1 || nz > NZP) return fail(ORBIT_ERR_INVALID, "nz must be in [1, 16]");
    if (ny < 1 || nx < 1) return fail(ORBIT_ERR_INVALID, "ny and nx must be positive");
    if (nbins < 1 || nbins > 64) return fail(ORBIT_ERR_INVALID, "n_bins must be in [1, 64]");
    if (ntr < 1) return fail(ORBIT_ERR_INVALID, "n_tracers must be positive");
    return ORBIT_OK;
}

}  // namespace

extern "C" int orbit_opset_create_csr(orbit_ctx* ctx, int nz, int ny, int nx, int n_bins,
                                      int n_tracers, double dtau, const int32_t* const* rowptr,
                                      const int32_t* const* colidx, const double* const* vals,
                                      orbit_opset** out) {
    if (!ctx || !out || !rowptr || !colidx || !vals) return fail(ORBIT_ERR_INVALID, "NULL argument");
    int rc = check_dims(nz, ny, nx, n_bins, n_tracers);
    if (rc) return rc;
    hipSetDevice(ctx->device);
    const int ncol = nx * ny;
    // parse every (bin, tracer); decide whether tracers share off-diagonals
    std::vector<HostBin> hb(n_bins);
    bool shared = true;
    std::vector<std::vector<double>> Vt(n_tracers), Ht(n_tracers);
    int wrap = 0;
    for (int b = 0; b < n_bins && shared; ++b) {
        for (int t = 0; t < n_tracers; ++t) {
            Vt[t].assign((size_t)ncol * 256, 0.0);
            Ht[t].assign((size_t)ncol * 64, 0.0);
            std::vector<double> dg((size_t)ncol * 16, 0.0);
            const int idx = b * n_tracers + t;
            if (!rowptr[idx] || !colidx[idx] || !vals[idx]) return fail(ORBIT_ERR_INVALID, "NULL CSR");
            rc = parse_csr(nz, ny, nx, dtau, rowptr[idx], colidx[idx], vals[idx], Vt[t].data(),
                           Ht[t].data(), dg.data(), &wrap);
            if (rc) return rc;
            if (t > 0 && (Vt[t] != Vt[0] || Ht[t] != Ht[0])) shared = false;
            if (t == 0) hb[b].d.assign((size_t)n_tracers * ncol * 16, 0.0);
            std::copy(dg.begin(), dg.end(), hb[b].d.begin() + (size_t)t * ncol * 16);
        }
        if (shared) {
            hb[b].V = Vt[0];
            hb[b].H = Ht[0];
        }
    }
    const int G = shared ? 1 : n_tracers;
    if (!shared) {
        for (int b = 0; b < n_bins; ++b) {
            hb[b].V.assign((size_t)G * ncol * 256, 0.0);
            hb[b].H.assign((size_t)G * ncol * 64, 0.0);
            hb[b].d.assign((size_t)n_tracers * ncol * 16, 0.0);
            for (int t = 0; t < n_tracers; ++t) {
                const int idx = b * n_tracers + t;
                rc = parse_csr(nz, ny, nx, dtau, rowptr[idx], colidx[idx], vals[idx],
                               hb[b].V.data() + (size_t)t * ncol * 256,
                               hb[b].H.data() + (size_t)t * ncol * 64,
                               hb[b].d.data() + (size_t)t * ncol * 16, &wrap);
                if (rc) return rc;
            }
        }
    }
    auto* op = new orbit_opset();
    op->ctx = ctx;
    op->nz = nz; op->ny = ny; op->nx = nx; op->nbins = n_bins; op->ntr = n_tracers;
    op->ngroups = G; op->ncol = ncol; op->n16 = ncol * 16; op->N = nz * ncol; op->dtau = dtau;
    op->px = wrap;
    rc = finish_opset(ctx, op, hb);
    if (rc) {
        orbit_opset_destroy(op);
        return rc;
    }
    *out = op;
    return ORBIT_OK;
}

extern "C" int orbit_opset_create_shared(orbit_ctx* ctx, int nz, int ny, int nx, int n_bins,
                                         int n_tracers, double dtau, const int32_t* const* rowptr,
                                         const int32_t* const* colidx, const double* const* T_vals,
                                         const double* D, orbit_opset** out) {
    if (!ctx || !out || !rowptr || !colidx || !T_vals || !D)
        return fail(ORBIT_ERR_INVALID, "NULL argument");
    int rc = check_dims(nz, ny, nx, n_bins, n_tracers);
    if (rc) return rc;
    hipSetDevice(ctx->device);
    const int ncol = nx * ny, N = nz * ncol;
    std::vector<HostBin> hb(n_bins);
    int wrap = 0;
    for (int b = 0; b < n_bins; ++b) {
        hb[b].V.assign((size_t)ncol * 256, 0.0);
        hb[b].H.assign((size_t)ncol * 64, 0.0);
        hb[b].tdiag.assign((size_t)ncol * 16, 0.0);
        if (!rowptr[b] || !colidx[b] || !T_vals[b]) return fail(ORBIT_ERR_INVALID, "NULL CSR");
        rc = parse_csr(nz, ny, nx, dtau, rowptr[b], colidx[b], T_vals[b], hb[b].V.data(),
                       hb[b].H.data(), hb[b].tdiag.data(), &wrap);
        if (rc) return rc;
        hb[b].d.assign((size_t)n_tracers * ncol * 16, 0.0);
        for (int t = 0; t < n_tracers; ++t) {
            const double* Dt = D + ((size_t)b * n_tracers + t) * N;
            double* dd = hb[b].d.data() + (size_t)t * ncol * 16;
            for (int n = 0; n < N; ++n) {
                const int k = n / ncol, c = n - k * ncol;
                dd[(size_t)c * 16 + k] = hb[b].tdiag[(size_t)c * 16 + k] + dtau * Dt[n];
            }
        }
    }
    auto* op = new orbit_opset();
    op->ctx = ctx;
    op->nz = nz; op->ny = ny; op->nx = nx; op->nbins = n_bins; op->ntr = n_tracers;
    op->ngroups = 1; op->ncol = ncol; op->n16 = ncol * 16; op->N = N; op->dtau = dtau;
    op->has_tdiag = true;
    op->px = wrap;
    rc = finish_opset(ctx, op, hb);
    if (rc) {
        orbit_opset_destroy(op);
        return rc;
    }
    *out = op;
    return ORBIT_OK;
}

extern "C" int orbit_opset_destroy(orbit_opset* op) {
    if (!op) return ORBIT_OK;
    if (op->ctx) {
        hipSetDevice(op->ctx->device);
        hipStreamSynchronize(op->ctx->stream);
    }
    for (auto& kv : op->graphs) hipGraphExecDestroy(kv.second);
    for (void* p : op->owned) hipFree(p);
    delete op;
    return ORBIT_OK;
}

extern "C" int orbit_opset_stats_get(orbit_opset* op, orbit_opset_stats* st) {
    if (!op || !st) return fail(ORBIT_ERR_INVALID, "NULL argument");
    st->nz = op->nz; st->ny = op->ny; st->nx = op->nx; st->n_bins = op->nbins;
    st->n_tracers = op->ntr; st->n_groups = op->ngroups;
    st->default_inner_sweeps = op->def_inner; st->default_acc_sweeps = op->def_acc;
    st->dtau = op->dtau; st->device_bytes = op->bytes; st->sweep_contraction = op->rho;
    st->periodic_x = op->px;
    st->relaxation = op->omega;
    st->mfma = op->mfma ? 1 : 0;
    st->default_check_sweeps = op->def_check;
    st->patch_shape = op->patch;
    return ORBIT_OK;
}

namespace {
void drop_graphs(orbit_opset* op);
}  // namespace

extern "C" int orbit_opset_set_relaxation(orbit_ctx* ctx, orbit_opset* op, double omega) {
    if (!ctx || !op) return fail(ORBIT_ERR_INVALID, "set_relaxation: NULL argument");
    if (!(omega == 0.0 || (omega >= 0.5 && omega <= 1.5)))
        return fail(ORBIT_ERR_INVALID, "set_relaxation: omega must be 0 (choose) or in [0.5, 1.5]");
    hipSetDevice(ctx->device);
    const double old_w = op->omega;
    const double old_rho = op->rho;
    const int old_acc = op->def_acc, old_inner = op->def_inner, old_check = op->def_check;
    drop_graphs(op);  // captured sweeps carry the old relaxation
    int rc = calibrate_opset(ctx, op, omega);
    if (rc) {  // keep the opset usable as it was
        op->omega = old_w;
        op->rho = old_rho;
        op->def_acc = old_acc;
        op->def_inner = old_inner;
        op->def_check = old_check;
    }
    return rc;
}

extern "C" int orbit_opset_update_diag(orbit_ctx* ctx, orbit_opset* op, const double* D) {
    if (!ctx || !op || !D) return fail(ORBIT_ERR_INVALID, "NULL argument");
    if (!op->has_tdiag) return fail(ORBIT_ERR_INVALID, "update_diag needs an opset made with the shared form");
    hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    const int ncol = op->ncol, n16 = op->n16, T = op->ntr, N = op->N;
    std::vector<double> tdiag((size_t)n16), dd((size_t)T * n16, 0.0);
    for (int b = 0; b < op->nbins; ++b) {
        BinOps& B = op->bins[b];
        ORBIT_HIP_CHECK(hipMemcpyAsync(tdiag.data(), B.tdiag, (size_t)n16 * 8, hipMemcpyDeviceToHost, st));
        ORBIT_HIP_CHECK(hipStreamSynchronize(st));
        for (int t = 0; t < T; ++t) {
            const double* Dt = D + ((size_t)b * T + t) * N;
            for (int n = 0; n < N; ++n) {
                const int k = n / ncol, c = n - k * ncol;
                dd[(size_t)t * n16 + (size_t)c * 16 + k] = tdiag[(size_t)c * 16 + k] + op->dtau * Dt[n];
            }
        }
        ORBIT_HIP_CHECK(hipMemcpyAsync(B.d64, dd.data(), dd.size() * 8, hipMemcpyHostToDevice, st));
        const size_t nd = (size_t)T * n16, nP = (size_t)T * ncol * 256;
        hipLaunchKernelGGL((k_convert<double, float>), grid1(nd), dim3(256), 0, st, B.d64, B.d32, nd);
        const long long vstride = op->ngroups == 1 ? 0 : (long long)ncol * 256;
        hipLaunchKernelGGL(k_factor, dim3((ncol + 3) / 4, T), dim3(64), 0, st, B.V64, vstride, B.d64,
                           (long long)n16, B.P64, (long long)ncol * 256, ncol);
        hipLaunchKernelGGL(k_to_bf16, grid1(nP), dim3(256), 0, st, B.P64, B.P16, nP);
        hipLaunchKernelGGL(k_transpose_blocks<double>, grid1(nP), dim3(256), 0, st, B.P64, B.PT64,
                           (size_t)T * ncol);
        hipLaunchKernelGGL(k_to_bf16, grid1(nP), dim3(256), 0, st, B.PT64, B.PT16, nP);
        ORBIT_HIP_CHECK(hipGetLastError());
        ORBIT_HIP_CHECK(hipStreamSynchronize(st));
    }
    for (auto& kv : op->graphs) hipGraphExecDestroy(kv.second);
    op->graphs.clear();
    // new diagonals: re-measure at the current relaxation (choose again only if
    // the sweeps stopped contracting with it)
    int rc = calibrate_opset(ctx, op, op->omega);
    if (rc == ORBIT_ERR_INVALID) rc = calibrate_opset(ctx, op, 0.0);
    return rc;
}

// ------------------------------------------------------------------------- //
// solver
// ------------------------------------------------------------------------- //
namespace orbit {
// The start of a forward forced sweep: orbit[0] = c (or 0, every slot) and the
// first bin's right-hand side ta = orbit[0] + E[0] (k_zero's and k_axpby's
// expressions; c and ta for the slots the mask keeps).
__global__ void k_forced_start(const double* __restrict__ c, long long cs, int zero_start,
                               const double* __restrict__ E, long long es, double* __restrict__ o0, long long os,
                               double* __restrict__ ta, long long ts, int n, const int* active, SlotLen sl) {
    const int slot = blockIdx.y;
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n || !in_slot(sl, slot, g)) return;
    const bool keep = !active || active[slot];
    double v = 0.0;
    if (zero_start) {
        o0[(size_t)slot * os + g] = 0.0;
    } else if (keep) {
        const double cv = c[(size_t)slot * cs + g];
        v = cv + 0.0 * cv;
        o0[(size_t)slot * os + g] = v;
    }
    if (keep) ta[(size_t)slot * ts + g] = v + 1.0 * E[(size_t)slot * es + g];
}

__global__ void k_zero(double* __restrict__ x, long long xslot, int n, SlotLen sl) {
    const int slot = blockIdx.y;
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n && in_slot(sl, slot, g)) x[(size_t)slot * xslot + g] = 0.0;
}

// x[col*16 + k] = 0 for the last column of every row (i = nx - 1) of every
// slot the mask keeps (all when null); grid (ceil(ny*16/256), nslots).
template <class T>
__global__ void k_zero_last_col(T* __restrict__ x, long long xslot, int nx, int ny, const int* active) {
    const int slot = blockIdx.y;
    if (active && !active[slot]) return;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ny * 16) return;
    const int j = g >> 4, k = g & 15;
    x[(size_t)slot * xslot + ((size_t)j * nx + nx - 1) * 16 + k] = T(0);
}


// ------------------------------------------------------------------------- //
// orbit-state refinement (mixed precision)
// ------------------------------------------------------------------------- //
// The states x_0 .. x_nb of an orbit satisfy one backward-Euler equation per
// bin.  Forward: A_t x_{t+1} = x_t + dtau e_t.  Adjoint (lam_t = P_t^T lam_{t+1}
// + R_t): A_t^T (x_t - R_t) = x_{t+1}.  The periodic closure ties the first and
// last state.  Its exact residual is one FP64 operator apply per bin:
//   forward  r_t = x_t + E_t - A_t x_{t+1}
//   adjoint  r_t = x_{t+1} - A_t^T (x_t - E_t)
// (E = dtau e, resp. R).  Every bin is independent: one launch covers all bins
// and slots.  Propagating r through the bins from zero gives the periodic
// residual s - (I - M) c exactly; no accurate inner solve is needed.
constexpr int CYC_MAXB = 64;
struct CycArgs {
    const double* V[CYC_MAXB];  // per bin: transposes of the column blocks the residual applies
    const double* H[CYC_MAXB];  // per bin: stencil (transposed for the adjoint)
    const double* d[CYC_MAXB];  // per bin: diagonals of every tracer
    long long vstride, hstride, dstride;  // per tracer
    int ncol, nx, ny, px, nslots;
    int tiles_per_xcd;    // single-set solves: 16-column tiles given to each of the 8 XCDs
    const GroupOp* gops;  // group solves: GroupOps [bin][precision][group], else null
    int ngroups, gprec;   // gprec: a precision row whose V/H/d are the FP64 operator
    const int* tmap;
    const double* X;      // states [slot][nbins + 1][n16]
    long long xs;
    const double* E;      // forcing [slot][nbins][n16]
    long long es;
    float* R;             // out: residuals [slot][nbins][n16] (they feed the FP32 residual sweep)
    long long rs;
    int n16, adj;
    int xzero;            // every state is zero (not read)
    const int* active;    // nullable
};

// 256 threads: 16 columns x 16 levels; grid (., ., nbins).  A chunk is up to
// four slots that share one operator: the block reads its columns' blocks and
// stencil once (when the transport is shared) and loops over the chunk's
// slots.  Single-set solves: grid.x = 8 XCDs x tiles_per_xcd x chunks, the
// chunks of a tile consecutive on one XCD (block L runs on XCD L % 8), so a
// tile's operator comes from that XCD's L2 after the first chunk; group
// solves: grid (ceil(max ncol / 16), groups), a chunk per slot group.  The 16 lanes of a column are one wavefront's, so
// the block product exchanges through LDS with wave-level ordering only.
__global__ __launch_bounds__(256) void k_cyc_residual(CycArgs a) {
    __shared__ double sy[256];
    const int t = blockIdx.z;
    const double *V, *H, *d;
    long long vst, hst, dst;
    int nx, ny, px, first, count, tile;
    if (a.gops) {
        const GroupOp& go = a.gops[((size_t)t * N_SWEEP_PRECS + a.gprec) * a.ngroups + blockIdx.y];
        V = (const double*)go.Vt; H = (const double*)go.H; d = (const double*)go.d;
        vst = 0; hst = 0; dst = go.dstride;
        nx = go.nx; ny = go.ny; px = go.px;
        first = go.first; count = go.count;
        tile = blockIdx.x;
    } else {
        V = a.V[t]; H = a.H[t]; d = a.d[t];
        vst = a.vstride; hst = a.hstride; dst = a.dstride;
        nx = a.nx; ny = a.ny; px = a.px;
        const int nch = (a.nslots + 3) >> 2, L = blockIdx.x, r = L >> 3;
        const int chunk = r % nch;
        tile = (L & 7) * a.tiles_per_xcd + r / nch;
        first = 4 * chunk; count = min(4, a.nslots - first);
    }
    const int ncol = nx * ny;
    if (tile * 16 >= ncol) return;  // uniform over the block
    const int lane = threadIdx.x & 15, grp = threadIdx.x >> 4;
    const int col0 = tile * 16 + grp;
    const bool valid = col0 < ncol;
    const int col = valid ? col0 : 0;
    const int j = col / nx, i = col - j * nx;
    const int cw = west_col(col, i, nx, px), ce = east_col(col, i, nx, px);
    const size_t base = (size_t)col * 16 + lane;
    const size_t n16 = (size_t)a.n16;
    const bool shared_op = vst == 0 && hst == 0;
    // the shared operator row of this lane (reloaded per slot for per-tracer
    // transports), read from the block's transpose: element m of row `lane`
    // is at m * 16 + lane there, so the 16 lanes of a column read 128
    // contiguous bytes per element
    double vrow[16], hw = 0.0, he = 0.0, hs = 0.0, hn = 0.0;
    auto load_op = [&](int tr) {
        const double* Vr = V + (size_t)tr * vst + (size_t)col * 256 + lane;
#pragma unroll
        for (int m = 0; m < 16; ++m) vrow[m] = Vr[m * 16];
        const double* Hc = H + (size_t)tr * hst + (size_t)col * 64 + lane;
        hw = Hc[0]; he = Hc[16]; hs = Hc[32]; hn = Hc[48];
    };
    if (shared_op) load_op(0);
    // Every slot's loads are issued before the first product (up to four slots
    // in flight; a per-tracer transport loads its rows and vectors per slot).
    double y0[4], yw[4], ye[4], ys[4], yn[4], uu[4], dd[4];
    bool on[4];
    auto load_slot = [&](int q) {
        const int slot = first + q;
        const int tr = a.tmap[slot];
        const double* __restrict__ Xs = a.X + (size_t)slot * a.xs;
        const double* __restrict__ Et = a.E + (size_t)slot * a.es + (size_t)t * n16;
        // y: the state the operator multiplies; u: the one it is compared with
        const double* __restrict__ Y = Xs + (size_t)(a.adj ? t : t + 1) * n16;
        const double* __restrict__ U = Xs + (size_t)(a.adj ? t + 1 : t) * n16;
        auto yv = [&](size_t ix) {
            const double x = a.xzero ? 0.0 : Y[ix];
            return a.adj ? x - Et[ix] : x;
        };
        y0[q] = yv(base);
        if (cw >= 0) yw[q] = yv((size_t)cw * 16 + lane);
        if (ce >= 0) ye[q] = yv((size_t)ce * 16 + lane);
        if (j > 0) ys[q] = yv((size_t)(col - nx) * 16 + lane);
        if (j < ny - 1) yn[q] = yv((size_t)(col + nx) * 16 + lane);
        uu[q] = (a.xzero ? 0.0 : U[base]) + (a.adj ? 0.0 : Et[base]);
        dd[q] = d[(size_t)tr * dst + base];
    };
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        on[q] = q < count && (!a.active || a.active[first + q]);
        y0[q] = yw[q] = ye[q] = ys[q] = yn[q] = uu[q] = dd[q] = 0.0;
        if (on[q] && shared_op) load_slot(q);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (!on[q]) continue;  // uniform over the block
        if (!shared_op) {
            load_op(a.tmap[first + q]);
            load_slot(q);
        }
        wave_sync();  // the previous slot's reads of sy are done
        sy[threadIdx.x] = y0[q];
        wave_sync();
        double acc = dd[q] * y0[q];
#pragma unroll
        for (int m = 0; m < 16; ++m) acc += vrow[m] * sy[(grp << 4) + m];
        acc += hw * yw[q];
        acc += he * ye[q];
        acc += hs * ys[q];
        acc += hn * yn[q];
        if (valid) a.R[(size_t)(first + q) * a.rs + (size_t)t * n16 + base] = (float)(uu[q] - acc);
    }
}

// Correction of every state after a GMRES cycle on (I - M~) delta = z_end:
// the intermediate states (1 .. nbins-1) get z_t + sum_i y_i W_i[t] (the
// FP32 states of the residual sweep and of each Krylov vector's monodromy
// apply, in Krylov order), the free state becomes c (already corrected by
// the cycle) and the other end c - scl (scl: the s of a solve given s, else
// zero) plus zend (nullable: the residual sweep's last state, when a
// converged orbit takes its states from that sweep).  Forward: free state 0,
// end state nbins; adjoint the reverse.
// Four consecutive elements a thread (n16 and every slot's length are
// multiples of 16): grid (ceil(n16 / 1024), nslots, nbins + 1).
__global__ void k_cyc_combine(double* __restrict__ X, long long xs, const float* __restrict__ Z, long long zs,
                              const float* __restrict__ W, long long wj, long long ws, const double* __restrict__ coef,
                              const int* __restrict__ kcols, const double* __restrict__ c, long long cs,
                              const double* __restrict__ scl, long long ss, const double* __restrict__ zend, int n16,
                              int nbins, int adj, const int* active, SlotLen sl) {
    const int slot = blockIdx.y, t = blockIdx.z;
    if (active && !active[slot]) return;
    const int g = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
    if (g >= n16 || !in_slot(sl, slot, g)) return;
    double* x = X + (size_t)slot * xs + (size_t)t * n16 + g;
    const int tfree = adj ? nbins : 0;
    double v[4];
    if (t == tfree || t == nbins - tfree) {
        const double2* cp = (const double2*)(c + (size_t)slot * cs + g);
        const double2 c0 = cp[0], c1 = cp[1];
        v[0] = c0.x; v[1] = c0.y; v[2] = c1.x; v[3] = c1.y;
        if (t != tfree) {
            if (scl) {
                const double2* sp = (const double2*)(scl + (size_t)slot * ss + g);
                const double2 s0 = sp[0], s1 = sp[1];
                v[0] -= s0.x; v[1] -= s0.y; v[2] -= s1.x; v[3] -= s1.y;
            }
            if (zend) {
                const double2* zp = (const double2*)(zend + (size_t)slot * cs + g);
                const double2 z0 = zp[0], z1 = zp[1];
                v[0] += z0.x; v[1] += z0.y; v[2] += z1.x; v[3] += z1.y;
            }
        }
    } else {
        const size_t off = (size_t)slot * ws + (size_t)(t - 1) * n16 + g;
        const float4 z = *(const float4*)(Z + (size_t)slot * zs + (size_t)(t - 1) * n16 + g);
        double u[4] = {(double)z.x, (double)z.y, (double)z.z, (double)z.w};
        const int k = kcols[slot];
        const double* cf = coef + (size_t)slot * MAXM1;
        for (int i = 0; i < k; ++i) {
            const float4 w = *(const float4*)(W + (size_t)i * wj + off);
            const double ci = cf[i];
            u[0] += ci * (double)w.x;
            u[1] += ci * (double)w.y;
            u[2] += ci * (double)w.z;
            u[3] += ci * (double)w.w;
        }
        const double2 x0 = ((const double2*)x)[0], x1 = ((const double2*)x)[1];
        v[0] = x0.x + u[0]; v[1] = x0.y + u[1]; v[2] = x1.x + u[2]; v[3] = x1.y + u[3];
    }
    ((double2*)x)[0] = make_double2(v[0], v[1]);
    ((double2*)x)[1] = make_double2(v[2], v[3]);
}
}  // namespace orbit

namespace {

enum Dir { FWD = 0, ADJ = 1 };
enum SweepPrec { SP_F64 = 0, SP_ACC_MIXED = 1, SP_F32 = 2 };

// A group solve: slots of several operator sets (any grid sizes) in one batch.
// Slots are ordered by set; up to four consecutive slots of one set form a slot
// group, the unit of the four-slot sweep mapping.  Device tables: GroupOps per
// (bin, sweep precision, group), per colour the (group, tile) of every block
// (XCD-contiguous), per slot the vector length, grid and in/out pointers.
struct GroupSpec {
    int ngroups = 0;
    size_t n16max = 0;
    int ncolmax = 0;
    bool mf = false;
    std::vector<orbit_opset*> slot_op;  // per slot (internal order)
    const GroupOp* gops = nullptr;      // [bin][prec][group]
    const int* btab[2] = {nullptr, nullptr};
    int nblocks[2] = {0, 0};
    const int* ptab = nullptr;          // the first sweep's (group, 4 x 2 patch) per block; null: half-sweeps
    bool handoff = false;               // every set has the bin hand-off on
    int npblocks = 0;
    double first_bytes = 0.0;           // algorithmic bytes of a first-sweep launch
    double bytes[3][2] = {{0, 0}, {0, 0}, {0, 0}};  // algorithmic bytes of a full launch, per precision and colour
    const int* slen = nullptr;          // n16 of each slot's set
    const int* sncol = nullptr;
    const int* snz = nullptr;
    const double* const* in_ptrs = nullptr;  // per slot: (n_bins, N) input
    double* const* out_ptrs = nullptr;       // per slot: (n_bins + 1, N) output
    int* inner_cnt = nullptr;           // per slot sweep counts of its set (filled by the solve)
    int* acc_cnt = nullptr;
    int* chk_cnt = nullptr;
};

struct Batch {
    orbit_ctx* ctx;
    orbit_opset* op;
    int nslots;
    int* tmap_dev;
    int* active_dev;   // slots still iterating inside a GMRES cycle
    int* part_dev;     // slots taking part in the current round
    const GroupSpec* grp = nullptr;  // a group solve's tables, else null
    bool tmap_identity = false;      // slot s solves tracer s (single-set solves only)
};

// Algorithmic bytes of one half-sweep launch: per column of the colour the
// shared block V (unless x of this colour is known zero) and stencil H (unless
// the other colour's x is), per slot the diagonal d (unless own x is zero),
// the inverse block P, the rhs b, own x read (unless zero) and written, and
// the other colour's x once (unless zero).
double halfsweep_bytes(int ncols, int nslots, size_t ae, size_t pe, size_t xe, bool own_zero, bool nbr_zero) {
    const double shared = (own_zero ? 0.0 : 256.0 * ae) + (nbr_zero ? 0.0 : 64.0 * ae);
    const double per_slot = (own_zero ? 0.0 : 16.0 * ae) + 256.0 * pe + 16.0 * xe + (own_zero ? 0.0 : 16.0 * xe) +
                            16.0 * xe + (nbr_zero ? 0.0 : 16.0 * xe);
    return (double)ncols * (shared + nslots * per_slot);
}

template <class XT, class AT, class PT, class CT>
int launch_sweep_kernel(orbit_ctx* ctx, int prec, SweepArgs a, const SweepShape& sh, bool own_zero, bool nbr_zero,
                        hipStream_t st, bool mf = false) {
    // the matrix-core block product exists for the BF16-inverse sweeps (four-slot mapping)
    constexpr bool kMfOk = std::is_same<PT, uint16_t>::value;
    mf = mf && kMfOk;
    const int cs = mf ? 1 : sh.cs, W = sh.threads;
    const int tc = cs ? W / 64 : W / 16;  // columns per tile
    const int ntiles = (a.ncols + tc - 1) / tc;
    const int nsg = cs ? (a.nslots + 3) / 4 : a.nslots;
    a.xcd_map = sh.xcd;
    a.tiles_per_xcd = (ntiles + 7) / 8;
    const dim3 grid(sh.xcd ? 8 * a.tiles_per_xcd * nsg : ntiles * nsg), block(W);
    const size_t lds = (cs ? (size_t)(W / 64) * 256 * sizeof(AT) : 0) + 2 * (size_t)W * sizeof(CT);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (SweepTiming* tm = ctx->timing) {
        if (tm->next + 2 <= tm->ev.size()) {
            e0 = tm->ev[tm->next];
            e1 = tm->ev[tm->next + 1];
            tm->next += 2;
            tm->recs.push_back({prec * 16 + (mf ? 8 : 0) + cs * 4 + (own_zero ? 2 : 0) + (nbr_zero ? 1 : 0),
                                halfsweep_bytes(a.ncols, a.nslots, sizeof(AT), sizeof(PT), sizeof(XT), own_zero,
                                                nbr_zero)});
        } else {
            ++tm->dropped;
        }
    }
#define ORBIT_SWEEP_LAUNCH(CSV, OZV, NZV)                                                                    \
    do {                                                                                                     \
        if (e0)                                                                                              \
            hipExtLaunchKernelGGL((k_halfsweep<XT, AT, PT, CT, CSV, OZV, NZV>), grid, block, lds, st, e0, e1, \
                                  0, a);                                                                     \
        else                                                                                                 \
            hipLaunchKernelGGL((k_halfsweep<XT, AT, PT, CT, CSV, OZV, NZV>), grid, block, lds, st, a);        \
    } while (0)
#define ORBIT_SWEEP_LAUNCH_MF(OZV, NZV)                                                                         \
    do {                                                                                                         \
        if (e0)                                                                                                  \
            hipExtLaunchKernelGGL((k_halfsweep<XT, AT, PT, CT, 1, OZV, NZV, true>), grid, block, lds, st, e0, e1, \
                                  0, a);                                                                         \
        else                                                                                                     \
            hipLaunchKernelGGL((k_halfsweep<XT, AT, PT, CT, 1, OZV, NZV, true>), grid, block, lds, st, a);        \
    } while (0)
    // (own_zero, nbr_zero) = (true, false) is the second colour of a cold first sweep
    if (mf) {
        if constexpr (kMfOk) {
            if (own_zero && nbr_zero) ORBIT_SWEEP_LAUNCH_MF(true, true);
            else if (own_zero) ORBIT_SWEEP_LAUNCH_MF(true, false);
            else if (nbr_zero) ORBIT_SWEEP_LAUNCH_MF(false, true);
            else ORBIT_SWEEP_LAUNCH_MF(false, false);
        }
    } else if (cs) {
        if (own_zero && nbr_zero) ORBIT_SWEEP_LAUNCH(1, true, true);
        else if (own_zero) ORBIT_SWEEP_LAUNCH(1, true, false);
        else if (nbr_zero) ORBIT_SWEEP_LAUNCH(1, false, true);
        else ORBIT_SWEEP_LAUNCH(1, false, false);
    } else {
        if (own_zero && nbr_zero) ORBIT_SWEEP_LAUNCH(0, true, true);
        else if (own_zero) ORBIT_SWEEP_LAUNCH(0, true, false);
        else if (nbr_zero) ORBIT_SWEEP_LAUNCH(0, false, true);
        else ORBIT_SWEEP_LAUNCH(0, false, false);
    }
#undef ORBIT_SWEEP_LAUNCH
#undef ORBIT_SWEEP_LAUNCH_MF
    return ORBIT_OK;
}

// A half-sweep of a group solve: every slot group's tiles of this colour in
// one launch (four-slot mapping, 256 threads, the block table's order).
template <class XT, class AT, class PT, class CT>
int launch_group_kernel(const Batch& B, int bin, int prec, SweepArgs a, bool own_zero, bool nbr_zero) {
    const GroupSpec& G = *B.grp;
    const int colour = a.colour;
    if (G.nblocks[colour] == 0) return ORBIT_OK;
    a.gops = G.gops + ((size_t)bin * N_SWEEP_PRECS + prec) * G.ngroups;
    a.btab = G.btab[colour];
    const dim3 grid(G.nblocks[colour]), block(256);
    const size_t lds = (size_t)4 * 256 * sizeof(AT) + 2 * 256 * sizeof(CT);
    constexpr bool kMfOk = std::is_same<PT, uint16_t>::value;
    const bool mf = G.mf && kMfOk;
    hipStream_t st = B.ctx->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (SweepTiming* tm = B.ctx->timing) {
        if (tm->next + 2 <= tm->ev.size()) {
            e0 = tm->ev[tm->next];
            e1 = tm->ev[tm->next + 1];
            tm->next += 2;
            tm->recs.push_back({prec * 16 + (mf ? 8 : 0) + 4 + (own_zero ? 2 : 0) + (nbr_zero ? 1 : 0),
                                G.bytes[prec][colour]});
        } else {
            ++tm->dropped;
        }
    }
#define ORBIT_GROUP_LAUNCH(MFV, OZV, NZV)                                                                        \
    do {                                                                                                         \
        if (e0)                                                                                                  \
            hipExtLaunchKernelGGL((k_halfsweep<XT, AT, PT, CT, 1, OZV, NZV, MFV, true>), grid, block, lds, st, e0, \
                                  e1, 0, a);                                                                     \
        else                                                                                                     \
            hipLaunchKernelGGL((k_halfsweep<XT, AT, PT, CT, 1, OZV, NZV, MFV, true>), grid, block, lds, st, a);   \
    } while (0)
    if (mf) {
        if constexpr (kMfOk) {
            if (own_zero && nbr_zero) ORBIT_GROUP_LAUNCH(true, true, true);
            else if (own_zero) ORBIT_GROUP_LAUNCH(true, true, false);
            else if (nbr_zero) ORBIT_GROUP_LAUNCH(true, false, true);
            else ORBIT_GROUP_LAUNCH(true, false, false);
        }
    } else {
        if (own_zero && nbr_zero) ORBIT_GROUP_LAUNCH(false, true, true);
        else if (own_zero) ORBIT_GROUP_LAUNCH(false, true, false);
        else if (nbr_zero) ORBIT_GROUP_LAUNCH(false, false, true);
        else ORBIT_GROUP_LAUNCH(false, false, false);
    }
#undef ORBIT_GROUP_LAUNCH
    return ORBIT_OK;
}

// Epilogue of an inner solve's last sweep (see SweepArgs::epi_out).
struct Epi {
    void* out = nullptr;
    long long os = 0;
    const void* add = nullptr;
    bool add_f32 = false;  // add is FP32 (else FP64)
    long long as = 0;
    double alpha = 0.0;
};

void set_epi(SweepArgs& a, const Epi* epi, int nsweeps) {
    a.epi_out = epi ? epi->out : nullptr;
    a.epi_os = epi ? epi->os : 0;
    a.epi_add = epi ? epi->add : nullptr;
    a.epi_add_f32 = epi && epi->add_f32 ? 1 : 0;
    a.epi_as = epi ? epi->as : 0;
    a.epi_alpha = epi ? epi->alpha : 0.0;
    a.nsweeps = nsweeps;
}

int launch_halfsweep(const Batch& B, int bin, int dir, int prec, int colour, const void* rhs,
                     long long bstride, void* x, long long xstride, bool own_zero, bool nbr_zero,
                     const int* mask, const int* counts = nullptr, int sweep_idx = 0, const Epi* epi = nullptr,
                     int nsweeps = 0) {
    if (B.grp) {  // operators per slot group; the direction is the tables'
        SweepArgs a{};
        a.b = rhs;
        a.bstride = bstride;
        a.x = x;
        a.xstride = xstride;
        a.tmap = B.tmap_dev;
        a.tmap_identity = 0;
        a.active = mask;
        a.own_zero = own_zero;
        a.nbr_zero = nbr_zero;
        a.counts = counts;
        a.sweep_idx = sweep_idx;
        a.colour = colour;
        a.nslots = B.nslots;
        set_epi(a, epi, nsweeps);
        if (prec == SP_F64) return launch_group_kernel<double, double, double, double>(B, bin, prec, a, own_zero, nbr_zero);
        if (prec == SP_ACC_MIXED)
            return launch_group_kernel<double, double, uint16_t, double>(B, bin, prec, a, own_zero, nbr_zero);
        return launch_group_kernel<float, float, uint16_t, float>(B, bin, prec, a, own_zero, nbr_zero);
    }
    orbit_opset* op = B.op;
    const BinOps& O = op->bins[bin];
    SweepArgs a;
    a.ncols = op->ncols_c[colour];
    if (a.ncols == 0) return ORBIT_OK;
    a.nx = op->nx;
    a.ny = op->ny;
    a.nxc = op->nxc;
    a.px = op->px;
    a.omega = (float)op->omega;
    a.vstride = op->ngroups == 1 ? 0 : (long long)op->ncol * 256;
    a.hstride = op->ngroups == 1 ? 0 : (long long)op->ncol * 64;
    a.dstride = op->n16;
    a.pstride = (long long)op->ncol * 256;
    a.b = rhs;
    a.bstride = bstride;
    a.x = x;
    a.xstride = xstride;
    a.tmap = B.tmap_dev;
    a.tmap_identity = B.tmap_identity ? 1 : 0;
    a.active = mask;
    a.own_zero = own_zero;
    a.nbr_zero = nbr_zero;
    a.counts = counts;
    a.sweep_idx = sweep_idx;
    a.colour = colour;
    a.nslots = B.nslots;
    a.gops = nullptr;
    a.btab = nullptr;
    set_epi(a, epi, nsweeps);
    const bool T = dir == ADJ;
    // launch shape of this precision and slot count (the four-slot mapping reads
    // one shared column block, so it needs a single transport group); the
    // BF16-inverse sweeps of a shared-transport set run the block product on the
    // matrix cores, always in the four-slot mapping, so a slot's arithmetic is
    // the same in every batch
    SweepShape sh = op->shapes[prec][slot_bucket(B.nslots)];
    if (op->ngroups != 1) sh.cs = 0;
    const bool mf = op->mfma && prec != SP_F64 && op->ngroups == 1;
    if (prec == SP_F64) {
        a.V = T ? O.VT64 : O.V64; a.Vt = T ? O.V64 : O.VT64;
        a.H = T ? O.HT64 : O.H64; a.d = O.d64; a.P = T ? O.PT64 : O.P64;
        return launch_sweep_kernel<double, double, double, double>(B.ctx, prec, a, sh, own_zero, nbr_zero, B.ctx->stream);
    } else if (prec == SP_ACC_MIXED) {
        a.V = T ? O.VT64 : O.V64; a.Vt = T ? O.V64 : O.VT64;
        a.H = T ? O.HT64 : O.H64; a.d = O.d64; a.P = T ? O.PT16 : O.P16;
        return launch_sweep_kernel<double, double, uint16_t, double>(B.ctx, prec, a, sh, own_zero, nbr_zero, B.ctx->stream,
                                                                     mf);
    }
    a.V = T ? O.VT32 : O.V32; a.Vt = T ? O.V32 : O.VT32;
    a.H = T ? O.HT32 : O.H32; a.d = O.d32; a.P = T ? O.PT16 : O.P16;
    return launch_sweep_kernel<float, float, uint16_t, float>(B.ctx, prec, a, sh, own_zero, nbr_zero, B.ctx->stream, mf);
}

// The first FP32 sweep from x = 0 of a two-colour, shared-transport set in one
// launch over column patches (k_first_sweep): the same bits as its two
// half-sweeps.  The patch shapes (orbit_opset_set_patch; 0: two half-sweep
// launches): columns per patch in x and y, waves per workgroup.
struct PatchShape { int pa, pb, waves; };
constexpr PatchShape PATCH_SHAPES[] = {{0, 0, 0}, {4, 4, 8}, {4, 2, 4}, {8, 2, 8}, {8, 4, 16}};
constexpr int N_PATCH_SHAPES = (int)(sizeof(PATCH_SHAPES) / sizeof(PATCH_SHAPES[0]));

template <int PA, int PB, int W, bool GRP = false, bool HO = false>
void first_sweep_launch(hipStream_t st, const FirstSweepArgs& a, dim3 grid, hipEvent_t e0, hipEvent_t e1) {
    if (e0) hipExtLaunchKernelGGL((k_first_sweep<PA, PB, W, GRP, HO>), grid, dim3(W * 64), 0, st, e0, e1, 0, a);
    else hipLaunchKernelGGL((k_first_sweep<PA, PB, W, GRP, HO>), grid, dim3(W * 64), 0, st, a);
}

// Group launches use one patch shape (4 x 2 columns, 4 waves), whose block
// table the group solve builds.
constexpr int GROUP_PATCH_SHAPE = 2;

// The previous bin's last black half-sweep, held back to run in the next
// bin's first-sweep launch (k_first_sweep's hand-off form).
struct Handoff {
    bool pending = false;
    int bin = 0;
    const void* rhs = nullptr;
    long long bstride = 0;
    void* x = nullptr;
    long long xstride = 0;
    bool has_epi = false;
    Epi epi;
};

int launch_first_sweep(const Batch& B, int bin, int dir, const void* rhs, long long bstride, void* x,
                       long long xstride, const int* mask, const Epi* epi, const int* counts, int nsweeps,
                       const Handoff* ho = nullptr) {
    if (ho && !ho->pending) return fail(ORBIT_ERR_INVALID, "first sweep: no hand-off to take");
    // the previous bin's last black half-sweep: its value at a black column is
    // this bin's right-hand side there only if this bin reads it from where
    // that half-sweep (or its epilogue) would have stored it
    auto set_handoff = [&](FirstSweepArgs& a) {
        a.hb = (const float*)ho->rhs;
        a.hbstride = ho->bstride;
        a.hx = (float*)ho->x;
        a.hxstride = ho->xstride;
        a.hepi_out = ho->has_epi ? (float*)ho->epi.out : nullptr;
        a.hepi_os = ho->has_epi ? ho->epi.os : 0;
        a.hepi_add = ho->has_epi ? ho->epi.add : nullptr;
        a.hepi_add_f32 = ho->has_epi && ho->epi.add_f32 ? 1 : 0;
        a.hepi_as = ho->has_epi ? ho->epi.as : 0;
        const void* feeds = ho->has_epi ? ho->epi.out : ho->x;
        const long long fs = ho->has_epi ? ho->epi.os : ho->xstride;
        return feeds == rhs && fs == bstride;
    };
    if (B.grp) {
        const GroupSpec& G = *B.grp;
        FirstSweepArgs a{};
        a.gops = G.gops + ((size_t)bin * N_SWEEP_PRECS + SP_F32) * G.ngroups;
        a.btab = G.ptab;
        a.b = (const float*)rhs;
        a.bstride = bstride;
        a.x = (float*)x;
        a.xstride = xstride;
        a.tmap = B.tmap_dev;
        a.tmap_identity = 0;
        a.active = mask;
        a.nslots = B.nslots;
        a.epi_out = epi ? (float*)epi->out : nullptr;
        a.epi_os = epi ? epi->os : 0;
        a.epi_add = epi ? epi->add : nullptr;
        a.epi_add_f32 = epi && epi->add_f32 ? 1 : 0;
        a.epi_as = epi ? epi->as : 0;
        a.counts = counts;
        a.nsweeps = nsweeps;
        if (ho) {
            if (!set_handoff(a)) return fail(ORBIT_ERR_INVALID, "first sweep: the hand-off does not feed this right-hand side");
            a.hgops = G.gops + ((size_t)ho->bin * N_SWEEP_PRECS + SP_F32) * G.ngroups;
        }
        hipStream_t st = B.ctx->stream;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (SweepTiming* tm = B.ctx->timing) {
            if (tm->next + 2 <= tm->ev.size()) {
                e0 = tm->ev[tm->next];
                e1 = tm->ev[tm->next + 1];
                tm->next += 2;
                tm->recs.push_back({(ho ? HANDOFF_VARIANT : FUSED_FIRST_VARIANT) + SP_F32 * 16 + 4 + 3,
                                    G.first_bytes + (ho ? G.bytes[SP_F32][1] : 0.0)});
            } else {
                ++tm->dropped;
            }
        }
        static_assert(PATCH_SHAPES[GROUP_PATCH_SHAPE].pa == 4 && PATCH_SHAPES[GROUP_PATCH_SHAPE].pb == 2 &&
                          PATCH_SHAPES[GROUP_PATCH_SHAPE].waves == 4,
                      "the group patch table is built for 4 x 2 patches");
        if (ho) first_sweep_launch<4, 2, 4, true, true>(st, a, dim3(G.npblocks), e0, e1);
        else first_sweep_launch<4, 2, 4, true>(st, a, dim3(G.npblocks), e0, e1);
        return ORBIT_OK;
    }
    orbit_opset* op = B.op;
    const BinOps& O = op->bins[bin];
    const bool T = dir == ADJ;
    const PatchShape sh = PATCH_SHAPES[op->patch];
    FirstSweepArgs a{};
    a.nx = op->nx;
    a.ny = op->ny;
    a.px = op->px;
    a.npx = (op->nx + sh.pa - 1) / sh.pa;
    a.npatches = a.npx * ((op->ny + sh.pb - 1) / sh.pb);
    a.tiles_per_xcd = (a.npatches + 7) / 8;
    a.H = T ? O.HT32 : O.H32;
    a.P = T ? O.PT16 : O.P16;
    a.pstride = (long long)op->ncol * 256;
    a.b = (const float*)rhs;
    a.bstride = bstride;
    a.x = (float*)x;
    a.xstride = xstride;
    a.tmap = B.tmap_dev;
    a.tmap_identity = B.tmap_identity ? 1 : 0;
    a.active = mask;
    a.nslots = B.nslots;
    a.omega = (float)op->omega;
    a.epi_out = epi ? (float*)epi->out : nullptr;
    a.epi_os = epi ? epi->os : 0;
    a.epi_add = epi ? epi->add : nullptr;
    a.epi_add_f32 = epi && epi->add_f32 ? 1 : 0;
    a.epi_as = epi ? epi->as : 0;
    a.counts = counts;
    a.nsweeps = nsweeps;
    if (ho) {
        if (counts || op->ncolours != 2 || op->ngroups != 1 || !set_handoff(a))
            return fail(ORBIT_ERR_INVALID, "first sweep: the hand-off does not feed this right-hand side");
        const BinOps& Q = op->bins[ho->bin];
        a.hVt = T ? Q.V32 : Q.VT32;
        a.hH = T ? Q.HT32 : Q.H32;
        a.hd = Q.d32;
        a.hP = T ? Q.PT16 : Q.P16;
        a.hdstride = (long long)op->n16;
        a.hepi_omega = (float)op->omega;
    }
    const int nsg = (B.nslots + 3) / 4;
    const dim3 grid((unsigned)(8 * a.tiles_per_xcd * nsg));
    hipStream_t st = B.ctx->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (SweepTiming* tm = B.ctx->timing) {
        if (tm->next + 2 <= tm->ev.size()) {
            e0 = tm->ev[tm->next];
            e1 = tm->ev[tm->next + 1];
            tm->next += 2;
            // each operand once: both colours' P, b and x written, the black columns' stencil
            double bytes = halfsweep_bytes(op->ncols_c[0], B.nslots, 4, 2, 4, true, true) +
                           halfsweep_bytes(op->ncols_c[1], B.nslots, 4, 2, 4, true, true) +
                           256.0 * op->ncols_c[1];
            // with the hand-off, the previous bin's black half-sweep on top (its
            // right-hand side and x of this bin's black columns are not read back)
            if (ho) bytes += halfsweep_bytes(op->ncols_c[1], B.nslots, 4, 2, 4, false, false);
            tm->recs.push_back({(ho ? HANDOFF_VARIANT : FUSED_FIRST_VARIANT) + SP_F32 * 16 + 4 + 3, bytes});
        } else {
            ++tm->dropped;
        }
    }
    if (ho) {
        switch (op->patch) {
            case 1: first_sweep_launch<4, 4, 8, false, true>(st, a, grid, e0, e1); break;
            case 2: first_sweep_launch<4, 2, 4, false, true>(st, a, grid, e0, e1); break;
            case 3: first_sweep_launch<8, 2, 8, false, true>(st, a, grid, e0, e1); break;
            case 4: first_sweep_launch<8, 4, 16, false, true>(st, a, grid, e0, e1); break;
            default: return fail(ORBIT_ERR_INVALID, "first sweep: unknown patch shape");
        }
        return ORBIT_OK;
    }
    switch (op->patch) {
        case 1: first_sweep_launch<4, 4, 8>(st, a, grid, e0, e1); break;
        case 2: first_sweep_launch<4, 2, 4>(st, a, grid, e0, e1); break;
        case 3: first_sweep_launch<8, 2, 8>(st, a, grid, e0, e1); break;
        case 4: first_sweep_launch<8, 4, 16>(st, a, grid, e0, e1); break;
        default: return fail(ORBIT_ERR_INVALID, "first sweep: unknown patch shape");
    }
    return ORBIT_OK;
}

// x = P_bin(rhs) approximately: `sweeps` red-black sweeps starting from x = 0.
// One full sweep: every colour in turn.  first: the sweep starts from x = 0
// (x is not read: own-zero for every colour, neighbour-zero for the first).
// With four colours the last column of each row (colours 2 and 3) is set to
// zero first, so that colours 1 and 2, whose neighbours include columns not
// yet written in this sweep, read zeros there; slots the mask leaves out are
// not touched.
int sweep_once(const Batch& B, int bin, int dir, int prec, const void* rhs, long long bstride, void* x,
               long long xstride, bool first, const int* mask, const int* counts, int sweep_idx,
               const Epi* epi = nullptr, int nsweeps = 0) {
    orbit_opset* op = B.op;
    const int ncolours = B.grp ? 2 : op->ncolours;
    // (the first sweep has no column-block product, so the matrix-core setting
    // does not enter; a group solve's sets all have one shared transport and
    // two colours)
    if (first && prec == SP_F32 &&
        (B.grp ? B.grp->ptab != nullptr : op->ncolours == 2 && !counts && op->ngroups == 1 && op->patch > 0))
        return launch_first_sweep(B, bin, dir, rhs, bstride, x, xstride, mask, epi, counts, nsweeps);
    if (!B.grp && first && op->ncolours == 4) {
        const dim3 g((unsigned)((op->ny * 16 + 255) / 256), B.nslots);
        if (prec == SP_F32)
            hipLaunchKernelGGL(k_zero_last_col<float>, g, dim3(256), 0, B.ctx->stream, (float*)x, xstride, op->nx,
                               op->ny, mask);
        else
            hipLaunchKernelGGL(k_zero_last_col<double>, g, dim3(256), 0, B.ctx->stream, (double*)x, xstride, op->nx,
                               op->ny, mask);
    }
    for (int c = 0; c < ncolours; ++c) {
        const int rc = launch_halfsweep(B, bin, dir, prec, c, rhs, bstride, x, xstride, first, first && c == 0, mask,
                                        counts, sweep_idx, epi, nsweeps);
        if (rc) return rc;
    }
    return ORBIT_OK;
}

// Whether consecutive bins' FP32 inner solves of `sweeps` sweeps can hand the
// last black half-sweep of one to the first-sweep launch of the next: one
// shared transport, two colours, the patch first sweep on, no per-slot counts,
// a last sweep that is not itself a first sweep.
// Up to 48 slots a launch: measured faster with the hand-off (per headline-
// size forward solve at 4 / 12 / 24 / 48 slots 2.63 / 4.41 / 7.87 / 13.98 ms
// against 2.85 / 4.52 / 8.17 / 14.32; group solves over three sub-domain
// sizes with 6 / 12 / 24 slots 2.76 / 3.49 / 5.17 against 3.00 / 3.71 /
// 5.34 ms), but slower for the 144 uneven orbits of one group solve (23.8
// against 22.9 ms), whose launches are bound by bandwidth, not by their count.
constexpr int HANDOFF_MAX_SLOTS = 48;
bool handoff_ok(const Batch& B, int prec, int sweeps, const int* counts) {
    if (B.nslots > HANDOFF_MAX_SLOTS) return false;
    if (B.grp) return prec == SP_F32 && sweeps >= 2 && B.grp->ptab && B.grp->handoff;
    const orbit_opset* op = B.op;
    return prec == SP_F32 && !B.grp && !counts && sweeps >= 2 && op->handoff && op->ncolours == 2 &&
           op->ngroups == 1 && op->patch > 0;
}

int inner_solve(const Batch& B, int bin, int dir, int prec, int sweeps, const void* rhs,
                long long bstride, void* x, long long xstride, const int* mask, bool warm = false,
                const int* counts = nullptr, const Epi* epi = nullptr, Handoff* hin = nullptr,
                Handoff* hout = nullptr) {
    // hin (pending): the previous bin's last black half-sweep runs in this
    // solve's first-sweep launch; hout: this solve's last black half-sweep is
    // held back for the next bin's (the caller checked handoff_ok for both)
    for (int s = 0; s < sweeps; ++s) {
        int rc;
        if (s == 0 && hin && hin->pending) {
            rc = launch_first_sweep(B, bin, dir, rhs, bstride, x, xstride, mask, epi, counts, sweeps, hin);
            hin->pending = false;
        } else if (s == sweeps - 1 && hout) {
            rc = launch_halfsweep(B, bin, dir, prec, 0, rhs, bstride, x, xstride, false, false, mask, counts, s, epi,
                                  sweeps);
            hout->pending = true;
            hout->bin = bin;
            hout->rhs = rhs;
            hout->bstride = bstride;
            hout->x = x;
            hout->xstride = xstride;
            hout->has_epi = epi != nullptr;
            if (epi) hout->epi = *epi;
        } else {
            rc = sweep_once(B, bin, dir, prec, rhs, bstride, x, xstride, s == 0 && !warm, mask, counts, s, epi, sweeps);
        }
        if (rc) return rc;
    }
    return ORBIT_OK;
}

}  // namespace

// Largest contraction of one red-black sweep with relaxation op->omega over
// every bin and the sample tracers tr[0..ns) (device list), with a fixed smooth
// right-hand side: rho = (||r_8|| / ||r_3||)^(1/5).  A bin / tracer whose
// residual after 3 sweeps is already at round-off (a single column, or no
// horizontal coupling: the column solve is exact) says nothing about the rate
// -- its r8/r3 would be round-off noise near 1 -- and is skipped; 0 when all are.
int measure_contraction(orbit_ctx* ctx, orbit_opset* op, int ns, int* tr, const double* b, double bnorm,
                        double* x, double* y, double* red, double* out, double& rmax) {
    hipStream_t st = ctx->stream;
    const size_t n16 = op->n16;
    const int nb = (int)((n16 + RED_CHUNK - 1) / RED_CHUNK);
    const int nbins = op->nbins;
    Batch B{ctx, op, ns, tr, nullptr, nullptr};
    const long long vstride = op->ngroups == 1 ? 0 : (long long)op->ncol * 256;
    const long long hstride = op->ngroups == 1 ? 0 : (long long)op->ncol * 64;
    const long long xs = (long long)n16;
    for (int bin = 0; bin < nbins; ++bin) {
        const BinOps& O = op->bins[bin];
        for (int k = 1; k <= 8; ++k) {
            const int rc = sweep_once(B, bin, FWD, SP_F64, b, 0, x, xs, k == 1, nullptr, nullptr, 0);
            if (rc) return rc;
            if (k == 3 || k == 8) {
                hipLaunchKernelGGL(k_apply_A, dim3((op->ncol + 15) / 16, ns), dim3(256), 0, st, O.VT64, vstride, O.H64,
                                   hstride, O.d64, (long long)n16, x, y, xs, tr, op->ncol, op->nx, op->ny, op->px);
                hipLaunchKernelGGL(k_axpby, dim3(grid1(n16).x, ns), dim3(256), 0, st, b, 0LL, y, xs, -1.0, y, xs,
                                   (int)n16, nullptr, SlotLen{});
                hipLaunchKernelGGL(k_norm2, dim3(nb, ns), dim3(RED_THREADS), 0, st, y, xs, (int)n16, nb, red, SlotLen{});
                hipLaunchKernelGGL(k_sum_partials, grid_waves(ns), dim3(256), 0, st, red, nb,
                                   out + (size_t)(2 * bin + (k == 8)) * ns, ns);
            }
        }
    }
    ORBIT_HIP_CHECK(hipGetLastError());
    std::vector<double> ho((size_t)2 * nbins * ns);
    ORBIT_HIP_CHECK(hipMemcpyAsync(ho.data(), out, ho.size() * 8, hipMemcpyDeviceToHost, st));
    ORBIT_HIP_CHECK(hipStreamSynchronize(st));
    rmax = 0.0;
    for (int bin = 0; bin < nbins; ++bin)
        for (int q = 0; q < ns; ++q) {
            const double r3 = std::sqrt(ho[(size_t)(2 * bin) * ns + q]), r8 = std::sqrt(ho[(size_t)(2 * bin + 1) * ns + q]);
            if (!(r3 > 1e-12 * bnorm) || !(r8 > 0.0)) continue;  // converged to round-off already
            rmax = std::max(rmax, std::pow(r8 / r3, 0.2));
        }
    return ORBIT_OK;
}

namespace {
constexpr int CAL_SLOTS = 8;  // tracers sampled by the sweep calibration

// Sweep counts from a measured contraction: 1e-12 for an accurate apply (plain
// FP64 path; 20% margin on the rate) and, for the mixed refinement, 5e-4 for an
// FP32 monodromy apply inside GMRES and 1e-3 (with the margin) for the residual
// sweep.  Neither sets the accuracy of the answer (every round's residual is
// exact in FP64); the first sets how much one round gains, the second also how
// honestly the residual is reported: its error is at most about n_bins times
// the per-bin error of the sweep, 0.4 % of the residual at 1e-3.  Measured on
// the 64x70x15 FCT passes (rho 0.078): 3 and 3 sweeps take 25.6 ms per step
// (3 rounds in the first passes, 2 in the later ones) against 30.0 with 4 and 4
// (2 rounds everywhere), 26.4 with 3 and 4, 25.7 with 2 and 3, 32.0 with 1 and 3.
void set_sweep_counts(orbit_opset* op, double rmax) {
    op->rho = std::min(0.97, std::max(0.02, 1.2 * rmax));
    op->def_acc = std::min(80, std::max(4, (int)std::ceil(std::log(1e-12) / std::log(op->rho))));
    op->def_inner = std::min(40, std::max(2, (int)std::ceil(std::log(5e-4) / std::log(std::min(0.97, rmax)))));
    op->def_check = std::min(40, std::max(2, (int)std::ceil(std::log(1e-3) / std::log(op->rho))));
}
}  // namespace

// Calibrates the sweeps of an opset.  omega > 0: that relaxation; omega = 0:
// choose it.  The correction-form sweep x += omega P (b - A x) keeps A x = b as
// its fixed point for any omega, so omega changes only the contraction: with
// the complex spectrum of the advective block Jacobi matrix it falls from
// omega = 1 (Gauss-Seidel) to a minimum a few hundredths above 1 and rises
// steeply after it (64x70x15: 0.099 at 1, 0.063 at 1.02, 0.088 at 1.03, 0.15 at
// 1.08).  The search measures 1, 1.02, 1.04, 1.06, 1.08, then +-0.01, +-0.005
// and +-0.0025 around the running best (48x40x15: 0.060 at 1.015 against 0.071
// at 1.02, one sweep fewer per inner solve), and keeps omega != 1 only when it
// contracts at least 5% better.
// The tracers the calibration measures (scratch slot 7): all of them when there are at most ns;
// else the ns/2 with the smallest diagonal on bin 0 (the weakest loss
// contracts slowest: its sweeps set the counts) and the rest spread evenly
// over the others.
int calibration_tracers(orbit_ctx* ctx, orbit_opset* op, int ns, std::vector<int>& ht) {
    const int T = op->ntr;
    ht.clear();
    if (T <= ns) {
        for (int t = 0; t < T; ++t) ht.push_back(t);
        return ORBIT_OK;
    }
    hipStream_t st = ctx->stream;
    const size_t n16 = op->n16;
    const int nb = (int)((n16 + RED_CHUNK - 1) / RED_CHUNK);
    int rc = ctx->ensure(7, ((size_t)T * nb + (size_t)T) * 8 + 64);
    if (rc) return rc;
    double* part = ctx->buf<double>(7);
    double* sums = part + (size_t)T * nb;
    hipLaunchKernelGGL(k_norm2, dim3(nb, T), dim3(RED_THREADS), 0, st, op->bins[0].d64, (long long)n16, (int)n16, nb,
                       part, SlotLen{});
    hipLaunchKernelGGL(k_sum_partials, grid_waves(T), dim3(256), 0, st, part, nb, sums, T);
    ORBIT_HIP_CHECK(hipGetLastError());
    std::vector<double> hs(T);
    ORBIT_HIP_CHECK(hipMemcpyAsync(hs.data(), sums, (size_t)T * 8, hipMemcpyDeviceToHost, st));
    ORBIT_HIP_CHECK(hipStreamSynchronize(st));
    std::vector<int> order(T);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return hs[a] < hs[b]; });
    const int weak = ns / 2;
    std::vector<char> taken(T, 0);
    for (int q = 0; q < weak; ++q) {
        ht.push_back(order[q]);
        taken[order[q]] = 1;
    }
    std::vector<int> rest;
    for (int t = 0; t < T; ++t)
        if (!taken[t]) rest.push_back(t);
    const int more = ns - weak;
    for (int q = 0; q < more; ++q) ht.push_back(rest[(size_t)q * rest.size() / more]);
    return ORBIT_OK;
}

int calibrate_opset(orbit_ctx* ctx, orbit_opset* op, double omega) {
    hipStream_t st = ctx->stream;
    const size_t n16 = op->n16;
    const int nb = (int)((n16 + RED_CHUNK - 1) / RED_CHUNK);
    const int nbins = op->nbins;
    const int ns = std::min(op->ntr, CAL_SLOTS);
    std::vector<int> ht;  // first: it uses the same scratch slot
    int rc = calibration_tracers(ctx, op, ns, ht);
    if (rc) return rc;
    rc = ctx->ensure(7, ((1 + 2 * (size_t)ns) * n16 + (size_t)ns * nb + 2 * (size_t)nbins * ns + 8) * 8 +
                            (size_t)ns * 4 + 64);
    if (rc) return rc;
    double* b = ctx->buf<double>(7);
    double* x = b + n16;
    double* y = x + (size_t)ns * n16;
    double* red = y + (size_t)ns * n16;
    double* out = red + (size_t)ns * nb;
    int* tr = (int*)(out + 2 * (size_t)nbins * ns + 4);
    std::vector<double> hb(n16, 0.0);
    for (int c = 0; c < op->ncol; ++c)
        for (int k = 0; k < op->nz; ++k) hb[(size_t)c * 16 + k] = 1.0 + 0.5 * std::sin(0.37 * c + 1.3 * k);
    double bnorm = 0.0;
    for (double v : hb) bnorm += v * v;
    bnorm = std::sqrt(bnorm);
    { int rc0 = h2d_sync(st, tr, ht.data(), ns * 4); if (rc0) return rc0; }
    ORBIT_HIP_CHECK(hipMemcpyAsync(b, hb.data(), n16 * 8, hipMemcpyHostToDevice, st));
    auto eval = [&](double w, double& r) {
        op->omega = w;
        return measure_contraction(ctx, op, ns, tr, b, bnorm, x, y, red, out, r);
    };
    double best_w = omega > 0.0 ? omega : 1.0, best = 0.0;
    rc = eval(best_w, best);
    if (rc) return rc;
    if (omega <= 0.0 && best < 1.0) {
        const double gs = best;
        for (double w : {1.02, 1.04, 1.06, 1.08}) {
            double r = 0.0;
            if ((rc = eval(w, r))) return rc;
            if (r < best) best = r, best_w = w;
        }
        for (double h : {0.01, 0.005, 0.0025}) {  // the minimum is sharp: narrow in on it
            const double c = best_w;
            for (double w : {c - h, c + h}) {
                if (w < 1.0) continue;
                double r = 0.0;
                if ((rc = eval(w, r))) return rc;
                if (r < best) best = r, best_w = w;
            }
        }
        if (!(best < 0.95 * gs)) best = gs, best_w = 1.0;
    }
    op->omega = best_w;
    if (!(best < 1.0)) return fail(ORBIT_ERR_INVALID, "opset: red-black sweeps do not contract on this operator");
    set_sweep_counts(op, best);
    return ORBIT_OK;
}

namespace {

struct Solver {
    Batch B;
    int prec = ORBIT_PREC_FP64;
    int dir = FWD;
    int inner = 6, acc = 14, m = 30;
    int chk = 6;              // sweeps per inner solve of the mixed residual sweep
    bool use_graph = false;
    size_t n16 = 0;
    int nbins = 8;
    double* E = nullptr;      // forward: dtau*e ; adjoint: R   [slot][bin][n16]
    double* orbit = nullptr;  // [slot][nbins+1][n16]
    double* c = nullptr;      // [slot][n16]
    double* r = nullptr;
    double* w = nullptr;
    double* ta = nullptr;
    double* tb = nullptr;
    double* Vk = nullptr;     // [slot][m+1][n16]
    float* fa = nullptr;
    float* fb = nullptr;
    float* fc = nullptr;
    float* fc2 = nullptr;     // the residual sweep's right-hand sides alternate between fc and fc2
    double rho = 0.1;
    double eps_acc = 1e-12;
    double* sv = nullptr;    // s given directly (rhs_is_s), else null
    double* partials = nullptr;
    double* h1 = nullptr;
    double* h2 = nullptr;
    double* scal = nullptr;
    double* resid = nullptr;
    GState* gst = nullptr;
    int* kcols = nullptr;
    double* host = nullptr;   // the context's pinned host staging
    int nb = 1;
    SlotLen sl;               // per-slot lengths (group solves), else none
    const int* inner_cnt = nullptr;  // per-slot sweep counts (group solves), else the uniform ones
    const int* acc_cnt = nullptr;
    const int* chk_cnt = nullptr;
    // orbit-state refinement (the mixed path): the states live in `orbit`
    bool states = false;
    float* Rc = nullptr;      // bin residuals [slot][nbins][n16] (FP32: the residual sweep's forcing)
    float* Z = nullptr;       // intermediate states of the residual sweep [slot][nbins-1][n16]
    float* W = nullptr;       // per Krylov vector j < kp: intermediate states of its M~ apply
    int kp = 0;               // Krylov vectors whose states are kept: a cycle's step cap
    const double* scl = nullptr;  // a solve given s: the closure offset (end state = free state - s)
};

// Steps of a mixed-precision GMRES cycle at most (each keeps the FP32 states
// of its monodromy apply; a slot that needs more continues in the next round).
constexpr int MIXED_CYCLE_CAP = 16;

int norms(Solver& S, const double* x, long long xslot, int n, std::vector<double>& out) {
    hipStream_t st = S.B.ctx->stream;
    const int nb = (n + RED_CHUNK - 1) / RED_CHUNK;
    hipLaunchKernelGGL(k_norm2, dim3(nb, S.B.nslots), dim3(RED_THREADS), 0, st, x, xslot, n, nb, S.partials, S.sl);
    hipLaunchKernelGGL(k_sum_partials, grid_waves(S.B.nslots), dim3(256), 0, st, S.partials, nb, S.scal, S.B.nslots);
    ORBIT_HIP_CHECK(hipMemcpyAsync(S.host, S.scal, S.B.nslots * 8, hipMemcpyDeviceToHost, st));
    ORBIT_HIP_CHECK(hipStreamSynchronize(st));
    out.assign(S.host, S.host + S.B.nslots);
    for (auto& v : out) v = std::sqrt(v);
    return ORBIT_OK;
}

// Squared norms per slot into dst[0..nslots) on the device, no host sync (the
// same reduction as norms).
void norms_dev(Solver& S, const double* x, long long xslot, int n, double* dst) {
    hipStream_t st = S.B.ctx->stream;
    const int nb = (n + RED_CHUNK - 1) / RED_CHUNK;
    hipLaunchKernelGGL(k_norm2, dim3(nb, S.B.nslots), dim3(RED_THREADS), 0, st, x, xslot, n, nb, S.partials, S.sl);
    hipLaunchKernelGGL(k_sum_partials, grid_waves(S.B.nslots), dim3(256), 0, st, S.partials, nb, dst, S.B.nslots);
}

// Forced accurate sweep.  Forward: orbit[0] = c (or 0), orbit[t+1] = P_t(orbit[t] + dtau e_t).
// Adjoint: orbit[nb] = c (or 0), orbit[t] = P_t^T orbit[t+1] + R_t for t = nb-1..0.
// Either way the outer residual s - (I - M)c is the last state minus the first.
// The plain FP64 path's check (FP64 inverse blocks).
int forced_sweep(Solver& S, bool zero_start, int sweeps) {
    hipStream_t st = S.B.ctx->stream;
    const int nbn = S.nbins;
    const size_t n16 = S.n16;
    const long long os = (long long)(nbn + 1) * n16, es = (long long)nbn * n16, vs = (long long)n16;
    const int sp = SP_F64;
    dim3 g(grid1(n16).x, S.B.nslots);
    const int* mask = S.B.part_dev;
    if (S.dir == FWD) {
        // orbit[0] and the first right-hand side in one pass; every later
        // right-hand side ta = orbit[t] + E[t], and at the end r = orbit[nb] -
        // orbit[0], are written by the previous bin's last sweep (its epilogue)
        // as each slot's final x comes out (every slot runs at least one sweep)
        hipLaunchKernelGGL(k_forced_start, g, dim3(256), 0, st, S.c, vs, zero_start ? 1 : 0, S.E, es, S.orbit, os,
                           S.ta, vs, (int)n16, mask, S.sl);
        for (int t = 0; t < nbn; ++t) {
            Epi epi;
            if (t + 1 < nbn) {
                epi.out = S.ta;
                epi.os = vs;
                epi.add = S.E + (size_t)(t + 1) * n16;
                epi.as = es;
                epi.alpha = 1.0;
            } else {
                epi.out = S.r;
                epi.os = vs;
                epi.add = S.orbit;
                epi.as = os;
                epi.alpha = -1.0;
            }
            int rc = inner_solve(S.B, t, FWD, sp, sweeps, S.ta, vs, S.orbit + (size_t)(t + 1) * n16, os, mask, false,
                                 S.acc_cnt, &epi);
            if (rc) return rc;
        }
        if (S.sv) hipLaunchKernelGGL(k_axpby, g, dim3(256), 0, st, S.r, vs, S.sv, vs, 1.0, S.r, vs, (int)n16, mask, S.sl);
    } else {
        if (zero_start)
            hipLaunchKernelGGL(k_zero, g, dim3(256), 0, st, S.orbit + (size_t)nbn * n16, os, (int)n16, S.sl);
        else
            hipLaunchKernelGGL(k_axpby, g, dim3(256), 0, st, S.c, vs, S.c, vs, 0.0,
                               S.orbit + (size_t)nbn * n16, os, (int)n16, mask, S.sl);
        for (int t = nbn - 1; t >= 0; --t) {
            // orbit[t] = P_t^T orbit[t+1] + R_t, written by the solve's last sweep
            Epi epi;
            epi.out = S.orbit + (size_t)t * n16;
            epi.os = os;
            epi.add = S.E + (size_t)t * n16;
            epi.as = es;
            epi.alpha = 1.0;
            int rc = inner_solve(S.B, t, ADJ, sp, sweeps, S.orbit + (size_t)(t + 1) * n16, os, S.ta, vs, mask, false,
                                 S.acc_cnt, &epi);
            if (rc) return rc;
        }
        hipLaunchKernelGGL(k_axpby, g, dim3(256), 0, st, S.orbit, os, S.orbit + (size_t)nbn * n16, os,
                           -1.0, S.r, vs, (int)n16, mask, S.sl);
        if (S.sv) hipLaunchKernelGGL(k_axpby, g, dim3(256), 0, st, S.r, vs, S.sv, vs, 1.0, S.r, vs, (int)n16, mask, S.sl);
    }
    ORBIT_HIP_CHECK(hipGetLastError());
    return ORBIT_OK;
}

// Exact residual of every bin's equation (k_cyc_residual) into S.Rc for the
// slots of the round (part mask); xzero: the states are taken as zero.
int cyc_residual(Solver& S, bool xzero) {
    const orbit_opset* op = S.B.op;
    const int nbn = S.nbins;
    const size_t n16 = S.n16;
    const bool T = S.dir == ADJ;
    CycArgs a{};
    if (S.B.grp) {
        a.gops = S.B.grp->gops;
        a.ngroups = S.B.grp->ngroups;
        a.gprec = SP_F64;  // (the FP64 column blocks, stencil and diagonals; transposed in the tables)
    } else {
        if (nbn > CYC_MAXB) return fail(ORBIT_ERR_INVALID, "solve: too many bins for the residual launch");
        for (int b = 0; b < nbn; ++b) {
            const BinOps& O = op->bins[b];
            a.V[b] = T ? O.V64 : O.VT64;  // the transpose of the applied block
            a.H[b] = T ? O.HT64 : O.H64;
            a.d[b] = O.d64;
        }
        a.vstride = op->ngroups == 1 ? 0 : (long long)op->ncol * 256;
        a.hstride = op->ngroups == 1 ? 0 : (long long)op->ncol * 64;
        a.dstride = op->n16;
        a.ncol = op->ncol;
        a.nx = op->nx;
        a.ny = op->ny;
        a.px = op->px;
    }
    a.nslots = S.B.nslots;
    a.tmap = S.B.tmap_dev;
    a.X = S.orbit;
    a.xs = (long long)(nbn + 1) * n16;
    a.E = S.E;
    a.es = (long long)nbn * n16;
    a.R = S.Rc;
    a.rs = (long long)nbn * n16;
    a.n16 = (int)n16;
    a.adj = T ? 1 : 0;
    a.xzero = xzero ? 1 : 0;
    a.active = S.B.part_dev;
    const int ncolmax = S.B.grp ? S.B.grp->ncolmax : op->ncol;
    const int ntiles = (ncolmax + 15) / 16;
    dim3 grid;
    if (S.B.grp) {
        grid = dim3((unsigned)ntiles, (unsigned)S.B.grp->ngroups, (unsigned)nbn);
    } else {
        a.tiles_per_xcd = (ntiles + 7) / 8;
        grid = dim3((unsigned)(8 * a.tiles_per_xcd * ((S.B.nslots + 3) / 4)), 1, (unsigned)nbn);
    }
    hipLaunchKernelGGL(k_cyc_residual, grid, dim3(256), 0, S.B.ctx->stream, a);
    ORBIT_HIP_CHECK(hipGetLastError());
    return ORBIT_OK;
}

// The residual sweep: the bin residuals S.Rc propagated from zero by the FP32
// sweeps, each bin's right-hand side the previous state plus its residual
// (forward z_{t+1} = P~_t (z_t + r_t); adjoint z_t = P~_t^T (z_{t+1} + r_t)).
// The intermediate states go to S.Z, the last one (forward state nbins,
// adjoint state 0) in FP64 to S.r: the periodic residual s - (I - M) c of the
// current states, to the FP32 sweeps' accuracy.  Slots of the round only.
int z_sweep(Solver& S) {
    hipStream_t st = S.B.ctx->stream;
    const int nbn = S.nbins;
    const size_t n16 = S.n16;
    const long long es = (long long)nbn * n16, vs = (long long)n16, zs = (long long)(nbn - 1) * n16;
    dim3 g(grid1(n16).x, S.B.nslots);
    const int* mask = S.B.part_dev;
    const bool fwd = S.dir == FWD;
    // the first bin's right-hand side is its residual itself; every later one
    // (state + residual) is written by the previous bin's last sweep, into fc
    // and fc2 in turn (a one-launch sweep reads the right-hand sides of the
    // columns around its patch, so none is overwritten while the bin runs)
    float* fcb[2] = {S.fc, S.fc2};
    const bool ho_ok = handoff_ok(S.B, SP_F32, S.chk, S.chk_cnt);
    Handoff ho;
    for (int k = 0; k < nbn; ++k) {
        const int t = fwd ? k : nbn - 1 - k;
        const bool last = k == nbn - 1;
        const int sout = fwd ? t + 1 : t;  // the state this bin writes
        float* out = last ? S.fb : S.Z + (size_t)(sout - 1) * n16;
        const float* rhs = k == 0 ? S.Rc + (size_t)t * n16 : fcb[(k - 1) & 1];
        const long long rs = k == 0 ? es : vs;
        Epi epi;
        if (!last) {
            epi.out = fcb[k & 1];
            epi.os = vs;
            epi.add = S.Rc + (size_t)(fwd ? t + 1 : t - 1) * n16;
            epi.add_f32 = true;
            epi.as = es;
        }
        int rc = inner_solve(S.B, t, S.dir, SP_F32, S.chk, rhs, rs, out, last ? vs : zs, mask, false, S.chk_cnt,
                             last ? nullptr : &epi, k > 0 && ho_ok ? &ho : nullptr, !last && ho_ok ? &ho : nullptr);
        if (rc) return rc;
    }
    hipLaunchKernelGGL((k_convert_slots<float, double>), g, dim3(256), 0, st, S.fb, vs, S.r, vs, (int)n16, mask, S.sl);
    ORBIT_HIP_CHECK(hipGetLastError());
    return ORBIT_OK;
}

// The inexact monodromy's sweeps on fixed FP32 buffers, fa -> result.  With
// the state refinement the intermediate states of Krylov vector j's apply go
// to S.W (they correct the states after the cycle) and the result to fb;
// else the bins alternate between fa and fb.
int mono_sweeps_f32(Solver& S, float** result, int j) {
    const long long vs = (long long)S.n16, zs = (long long)(S.nbins - 1) * S.n16;
    const int nbn = S.nbins;
    const bool fwd = S.dir == FWD;
    float* in = S.fa;
    long long ins = vs;
    float* pp[2] = {S.fa, S.fb};
    // bins hand their last black half-sweep to the next one's first-sweep
    // launch when each bin's states have buffers of their own
    const bool ho_ok = S.states && handoff_ok(S.B, SP_F32, S.inner, S.inner_cnt);
    Handoff ho;
    for (int k = 0; k < nbn; ++k) {
        const int t = fwd ? k : nbn - 1 - k;
        const bool last = k == nbn - 1;
        float* out;
        long long outs = vs;
        if (S.states) {
            const int sout = fwd ? t + 1 : t;
            out = last ? S.fb : S.W + ((size_t)j * S.B.nslots * (nbn - 1) + (size_t)(sout - 1)) * S.n16;
            if (!last) outs = zs;
        } else {
            out = pp[(k + 1) & 1];
        }
        int rc = inner_solve(S.B, t, S.dir, SP_F32, S.inner, in, ins, out, outs, S.B.active_dev, false, S.inner_cnt,
                             nullptr, k > 0 && ho_ok ? &ho : nullptr, !last && ho_ok ? &ho : nullptr);
        if (rc) return rc;
        in = out;
        ins = outs;
    }
    *result = in;
    return ORBIT_OK;
}

// M v with the monodromy used inside GMRES (inexact FP32 for mixed, accurate
// FP64 for plain): *res receives the buffer holding it (FP32 when *f32), the
// caller forms w = v - M v (in the first orthogonalisation pass).  j: the
// Krylov vector's index (where the state refinement keeps its states).
int apply_M(Solver& S, const double* v, long long vslot, const void** res_out, bool* f32, bool v_in_fa, int j) {
    hipStream_t st = S.B.ctx->stream;
    const size_t n16 = S.n16;
    const long long vs = (long long)n16;
    dim3 g(grid1(n16).x, S.B.nslots);
    const int* mask = S.B.active_dev;
    if (S.prec == ORBIT_PREC_MIXED) {
        if (!v_in_fa)
            hipLaunchKernelGGL((k_convert_slots<double, float>), g, dim3(256), 0, st, v, vslot, S.fa, vs, (int)n16, mask,
                               S.sl);
        float* res = nullptr;
        if (S.use_graph) {
            float* wj = S.states ? S.W + (size_t)j * S.B.nslots * (S.nbins - 1) * n16 : nullptr;
            GraphKey key{S.dir, (S.states ? 3 : 2) + (S.B.tmap_identity ? 4 : 0), S.inner, S.B.nslots, {}, S.fa,
                         S.states ? (const void*)wj : S.fb};
            key.tracers = {(int)(intptr_t)S.B.tmap_dev, (int)((intptr_t)S.B.tmap_dev >> 32),
                           (int)(intptr_t)S.B.active_dev, (int)((intptr_t)S.B.active_dev >> 32),
                           (int)(intptr_t)S.fb, (int)((intptr_t)S.fb >> 32)};
            auto& graphs = S.B.op->graphs;
            std::lock_guard<std::mutex> lock(S.B.op->graph_mu);
            auto it = graphs.find(key);
            if (it == graphs.end()) {
                hipGraph_t graph;
                ORBIT_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                int rc = mono_sweeps_f32(S, &res, j);
                hipError_t e = hipStreamEndCapture(st, &graph);
                if (rc) return rc;
                if (e != hipSuccess) return fail(ORBIT_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
                hipGraphExec_t exec;
                e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                hipGraphDestroy(graph);
                if (e != hipSuccess) return fail(ORBIT_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
                it = graphs.emplace(key, exec).first;
            } else {
                res = (S.states || S.nbins % 2 == 1) ? S.fb : S.fa;
            }
            ORBIT_HIP_CHECK(hipGraphLaunch(it->second, st));
        } else {
            int rc = mono_sweeps_f32(S, &res, j);
            if (rc) return rc;
        }
        *res_out = res;
        *f32 = true;
    } else {
        const double* in = v;
        long long ins = vslot;
        double* bufs[2] = {S.ta, S.tb};
        int which = 0;
        const int nbn = S.nbins;
        for (int k = 0; k < nbn; ++k) {
            const int t = S.dir == FWD ? k : nbn - 1 - k;
            double* out = bufs[which];
            int rc = inner_solve(S.B, t, S.dir, SP_F64, S.acc, in, ins, out, vs, mask, false, S.acc_cnt);
            if (rc) return rc;
            in = out;
            ins = vs;
            which ^= 1;
        }
        *res_out = in;
        *f32 = false;
    }
    ORBIT_HIP_CHECK(hipGetLastError());
    return ORBIT_OK;
}

// One GMRES cycle (at most m Arnoldi steps) on (I - M~) delta = r, c += delta,
// for the slots with part[s] = 1.  Stops a slot once its residual estimate is
// at or below target[s].  Returns the Arnoldi steps taken per slot.
int gmres_cycle(Solver& S, const std::vector<int>& part, const std::vector<double>& target,
                std::vector<int>& steps) {
    hipStream_t st = S.B.ctx->stream;
    const int ns = S.B.nslots;
    const size_t n16 = S.n16;
    // the state refinement keeps every Krylov vector's FP32 states: at most kp steps
    const int mk = S.states ? std::min(S.m, S.kp) : S.m;
    const long long vs = (long long)n16, kslot = (long long)(mk + 1) * n16;
    const int nb = S.nb;
    dim3 g(grid1(n16).x, ns);
    std::vector<int> act(part);
    { int rc0 = set_ints(st, S.B.active_dev, act.data(), ns); if (rc0) return rc0; }
    // beta = ||r||, V0 = r / beta
    hipLaunchKernelGGL(k_norm2, dim3(nb, ns), dim3(RED_THREADS), 0, st, S.r, vs, (int)n16, nb, S.partials, S.sl);
    // the mixed path's monodromy reads its input in FP32 from fa: the kernels
    // that write a basis vector also write that copy
    float* fcopy = S.prec == ORBIT_PREC_MIXED ? S.fa : nullptr;
    hipLaunchKernelGGL(k_gmres_start_scale, g, dim3(256), 0, st, S.gst, S.partials, nb, S.B.active_dev, S.r, vs, S.Vk,
                       kslot, (int)n16, S.sl, fcopy, vs);
    steps.assign(ns, 0);
    std::vector<int> kc(ns, 0);
    for (int j = 0; j < mk; ++j) {
        const void* res = nullptr;
        bool f32 = false;
        int rc = apply_M(S, S.Vk + (size_t)j * n16, kslot, &res, &f32, fcopy != nullptr, j);
        if (rc) return rc;
        // w = v - M v with the first pass of CGS2, then the second pass; the
        // coefficients are summed inside the updates, ||w|| and the Givens
        // step inside the scaling of V_{j+1}
        const double* vj = S.Vk + (size_t)j * n16;
        if (f32)
            hipLaunchKernelGGL((k_mdot_sub<float>), dim3(nb, ns), dim3(RED_THREADS), 0, st, S.Vk, kslot, j + 1, vj,
                               kslot, (const float*)res, vs, S.w, vs, (int)n16, nb, S.partials, S.B.active_dev, S.sl);
        else
            hipLaunchKernelGGL((k_mdot_sub<double>), dim3(nb, ns), dim3(RED_THREADS), 0, st, S.Vk, kslot, j + 1, vj,
                               kslot, (const double*)res, vs, S.w, vs, (int)n16, nb, S.partials, S.B.active_dev,
                               S.sl);
        // CGS2 (the plain FP64 path); one pass for the mixed refinement, whose
        // short cycles stay orthogonal enough and whose every round's residual
        // is recomputed exactly
        double* np = S.partials + (size_t)ns * MAXM1 * nb;
        hipLaunchKernelGGL(k_mupdate_fin, dim3(nb, ns), dim3(RED_THREADS), 0, st, S.Vk, kslot, j + 1, S.partials,
                           S.h1, S.w, vs, (int)n16, nb, S.states ? np : (double*)nullptr, S.B.active_dev, S.sl);
        if (!S.states) {
            hipLaunchKernelGGL(k_mdot, dim3(nb, ns), dim3(RED_THREADS), 0, st, S.Vk, kslot, j + 1, S.w, vs,
                               (int)n16, nb, S.partials, S.B.active_dev, S.sl);
            hipLaunchKernelGGL(k_mupdate_fin, dim3(nb, ns), dim3(RED_THREADS), 0, st, S.Vk, kslot, j + 1,
                               S.partials, S.h2, S.w, vs, (int)n16, nb, np, S.B.active_dev, S.sl);
        }
        hipLaunchKernelGGL(k_arnoldi_tail, g, dim3(256), 0, st, S.gst, S.h1, S.states ? (const double*)nullptr : S.h2,
                           np, nb, j, S.B.active_dev, S.resid, S.w, vs, S.Vk + (size_t)(j + 1) * n16, kslot,
                           (int)n16, S.sl, fcopy, vs);
        ORBIT_HIP_CHECK(hipGetLastError());
        ORBIT_HIP_CHECK(hipMemcpyAsync(S.host, S.resid, ns * 8, hipMemcpyDeviceToHost, st));
        ORBIT_HIP_CHECK(hipStreamSynchronize(st));
        bool any = false, changed = false;
        for (int s = 0; s < ns; ++s) {
            if (!act[s]) continue;
            steps[s] = j + 1;
            kc[s] = j + 1;
            if (!(S.host[s] > target[s]) || j + 1 == mk) {
                act[s] = 0;
                changed = true;
            } else {
                any = true;
            }
        }
        if (!any) break;
        if (changed)
            { int rc0 = set_ints(st, S.B.active_dev, act.data(), ns); if (rc0) return rc0; }
    }
    // c += V y for the participating slots
    { int rc0 = set_ints(st, S.kcols, kc.data(), ns); if (rc0) return rc0; }
    hipLaunchKernelGGL(k_gmres_solve, grid1(ns), dim3(256), 0, st, S.gst, S.kcols, ns, S.h1);
    hipLaunchKernelGGL(k_mupdate, dim3(nb, ns), dim3(RED_THREADS), 0, st, S.Vk, kslot, mk, S.h1, -1.0, S.c, vs,
                       (int)n16, nb, (double*)nullptr, S.B.part_dev, (const int*)S.kcols, S.sl);
    if (S.states) {
        // every state: the free one is c, the intermediate ones get z_t + sum_i y_i W_i[t]
        const int nbn = S.nbins;
        const long long ws = (long long)(nbn - 1) * n16;
        hipLaunchKernelGGL(k_cyc_combine, dim3((unsigned)((n16 + 1023) / 1024), ns, nbn + 1), dim3(256), 0, st,
                           S.orbit, (long long)(nbn + 1) * n16, S.Z, ws, S.W, (long long)ns * ws, ws, S.h1, S.kcols,
                           S.c, vs, S.scl, vs, (const double*)nullptr, (int)n16, nbn, S.dir == ADJ ? 1 : 0,
                           S.B.part_dev, S.sl);
    }
    ORBIT_HIP_CHECK(hipGetLastError());
    return ORBIT_OK;
}

// Device scratch of one periodic solve in N_SOLVE_PARTS parts: 0 the forcing,
// 1 the orbit, 2 the Krylov basis, 3 six FP64 work vectors, 4 three FP32 sweep
// buffers, 5 reduction partials, Hessenberg coefficients and per-slot ints,
// 6 the GMRES states; the state refinement of the mixed path adds 7 the bin
// residuals, 8 the residual sweep's FP32 states and 9 those of every Krylov
// vector's monodromy apply.  Sizes per slot scale with n16 (the padded vector).
constexpr int N_SOLVE_PARTS = 10;
// the context scratch slot of each part (slots 7-9 hold other work)
constexpr int SOLVE_PART_SLOT[N_SOLVE_PARTS] = {0, 1, 2, 3, 4, 5, 6, 10, 11, 12};
inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
// Krylov vectors a cycle may take: the restart, capped for the mixed path
inline int cycle_steps(int m, bool mixed) { return mixed ? std::min(m, MIXED_CYCLE_CAP) : m; }
void solve_scratch_parts(size_t nt, int nbn, size_t n16, int m, bool mixed, size_t* b) {
    const size_t nb = (n16 + RED_CHUNK - 1) / RED_CHUNK;
    const size_t mk = (size_t)cycle_steps(m, mixed);
    b[0] = nt * nbn * n16 * 8;
    b[1] = nt * (nbn + 1) * n16 * 8;
    b[2] = nt * (mk + 1) * n16 * 8;
    b[3] = nt * 6 * n16 * 8;
    b[4] = nt * 4 * n16 * 4;
    b[5] = nt * MAXM1 * nb * 8 * 2 + nt * MAXM1 * 8 * 2 + nt * 8 * 2 + nt * 4 * 6 + 1024;
    b[6] = nt * sizeof(GState);
    b[7] = mixed ? nt * nbn * n16 * 4 : 0;
    b[8] = mixed ? nt * (nbn - 1) * n16 * 4 : 0;
    b[9] = mixed ? mk * nt * (nbn - 1) * n16 * 4 : 0;
}
size_t solve_scratch_bytes(size_t nt, int nbn, size_t n16, int m, bool mixed) {
    size_t b[N_SOLVE_PARTS], t = 0;
    solve_scratch_parts(nt, nbn, n16, m, mixed, b);
    for (int q = 0; q < N_SOLVE_PARTS; ++q) t += align256(b[q]);
    return t;
}

int solve_periodic(orbit_ctx* ctx, orbit_opset* op, const int* tracers, int nt, const double* in_dev,
                   const double* warm_dev, const orbit_opts* opts_in, double* out_dev, orbit_info* info,
                   int dir, PassDelta* pd = nullptr, void* arena = nullptr, size_t arena_bytes = 0,
                   const GroupSpec* G = nullptr) {
    if (!ctx || !op || !tracers || nt < 1 || !in_dev || !out_dev || !info)
        return fail(ORBIT_ERR_INVALID, "solve: NULL argument or empty tracer list");
    // the operator set of slot s (a group solve's slots come from several)
    auto op_of = [&](int s) { return G ? G->slot_op[s] : op; };
    orbit_opts o;
    orbit_opts_default(&o);
    if (opts_in) o = *opts_in;
    if (!(o.tol > 0.0)) return fail(ORBIT_ERR_INVALID, "solve: tol must be positive");
    if (o.restart < 1 || o.restart > MAXM) return fail(ORBIT_ERR_INVALID, "solve: restart must be in [1, 64]");
    if (o.maxiter < 1) return fail(ORBIT_ERR_INVALID, "solve: maxiter must be >= 1");
    if (o.precision != ORBIT_PREC_FP64 && o.precision != ORBIT_PREC_MIXED)
        return fail(ORBIT_ERR_INVALID, "solve: unknown precision");
    for (int s = 0; s < nt; ++s)
        if (tracers[s] < 0 || tracers[s] >= op_of(s)->ntr)
            return fail(ORBIT_ERR_INVALID, "solve: tracer index out of range");
    if (G && (pd || o.rhs_is_s || warm_dev))
        return fail(ORBIT_ERR_INVALID, "solve: a group solve starts cold from e (no s, warm start or pass state)");
    if (pd && (dir != FWD || o.precision != ORBIT_PREC_MIXED || o.rhs_is_s || !pd->s_state))
        return fail(ORBIT_ERR_INVALID, "solve: pass-to-pass state needs a mixed-precision forward solve and s_state");
    // a pass after the first: the previous pass's orbit starts the states and
    // pd->s_state holds its s~
    const bool next_pass = pd && warm_dev && o.warm_orbit;
    if (pd && warm_dev && !o.warm_orbit)
        return fail(ORBIT_ERR_INVALID, "solve: a later pass starts from the previous pass's whole orbit");
    hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;

    Solver S;
    S.prec = o.precision;
    S.dir = dir;
    S.rho = op->rho;
    S.eps_acc = std::min(1e-12, 0.01 * o.tol);
    const bool mixed = S.prec == ORBIT_PREC_MIXED;
    // per slot (its set's calibration): sweep contraction, inner and accurate
    // sweep counts, the relative error of one inexact (FP32) P apply
    std::vector<double> rho_s(nt), eps_in_s(nt);
    std::vector<int> inner_s(nt), acc_s(nt), chk_s(nt);
    for (int s = 0; s < nt; ++s) {
        const orbit_opset* os_ = op_of(s);
        rho_s[s] = os_->rho;
        inner_s[s] = o.inner_sweeps > 0 ? o.inner_sweeps : os_->def_inner;
        chk_s[s] = o.check_sweeps > 0 ? o.check_sweeps : os_->def_check;
        // what one round can gain: the less accurate of the two inexact applies
        eps_in_s[s] = std::pow(rho_s[s], std::min(inner_s[s], chk_s[s]));
        acc_s[s] = o.acc_sweeps > 0 ? o.acc_sweeps
                                    : std::min(80, std::max(4, (int)std::ceil(std::log(S.eps_acc) / std::log(rho_s[s]))));
    }
    S.inner = *std::max_element(inner_s.begin(), inner_s.end());
    S.acc = *std::max_element(acc_s.begin(), acc_s.end());
    S.chk = *std::max_element(chk_s.begin(), chk_s.end());
    S.m = o.restart;
    S.use_graph = o.use_graph != 0 && mixed && !ctx->timing && !G;
    S.n16 = G ? G->n16max : op->n16;
    S.nbins = op->nbins;
    const size_t n16 = S.n16;
    const int nbn = op->nbins;
    S.nb = (int)((n16 + RED_CHUNK - 1) / RED_CHUNK);

    // scratch: the context's growable buffers, or carved from a caller's arena
    size_t part_bytes[N_SOLVE_PARTS];
    solve_scratch_parts(nt, nbn, n16, S.m, mixed, part_bytes);
    void* part_ptr[N_SOLVE_PARTS];
    int rc = ORBIT_OK;
    if (arena) {
        const size_t need = solve_scratch_bytes(nt, nbn, n16, S.m, mixed);
        if (need > arena_bytes)
            return fail(ORBIT_ERR_INVALID, "solve: " + std::to_string(nt) + " tracers of this operator set with restart " +
                                               std::to_string(S.m) + " need " + std::to_string(need) +
                                               " bytes of scratch; the fixed scratch array holds " +
                                               std::to_string(arena_bytes));
        char* a = (char*)arena;
        for (int q = 0; q < N_SOLVE_PARTS; ++q) {
            part_ptr[q] = a;
            a += align256(part_bytes[q]);
        }
    } else {
        for (int q = 0; q < N_SOLVE_PARTS && !rc; ++q) {
            if (!part_bytes[q]) {
                part_ptr[q] = nullptr;
                continue;
            }
            rc = ctx->ensure(SOLVE_PART_SLOT[q], part_bytes[q]);
            part_ptr[q] = ctx->scratch[SOLVE_PART_SLOT[q]].p;
        }
        if (rc) return rc;
    }
    S.E = (double*)part_ptr[0];
    S.orbit = (double*)part_ptr[1];
    S.Vk = (double*)part_ptr[2];
    double* v5 = (double*)part_ptr[3];
    S.c = v5;
    S.r = v5 + (size_t)nt * n16;
    S.w = v5 + (size_t)2 * nt * n16;
    S.ta = v5 + (size_t)3 * nt * n16;
    S.tb = v5 + (size_t)4 * nt * n16;
    if (o.rhs_is_s) S.sv = v5 + (size_t)5 * nt * n16;
    S.fa = (float*)part_ptr[4];
    S.fb = S.fa + (size_t)nt * n16;
    S.fc = S.fa + (size_t)2 * nt * n16;
    S.fc2 = S.fa + (size_t)3 * nt * n16;
    char* sm = (char*)part_ptr[5];
    S.partials = (double*)sm; sm += (size_t)nt * MAXM1 * S.nb * 8 * 2;
    S.h1 = (double*)sm; sm += (size_t)nt * MAXM1 * 8;
    S.h2 = (double*)sm; sm += (size_t)nt * MAXM1 * 8;
    S.scal = (double*)sm; sm += (size_t)nt * 8;
    S.resid = (double*)sm; sm += (size_t)nt * 8;
    S.kcols = (int*)sm; sm += (size_t)nt * 4;
    S.B.tmap_dev = (int*)sm; sm += (size_t)nt * 4;
    S.B.active_dev = (int*)sm; sm += (size_t)nt * 4;
    S.B.part_dev = (int*)sm;
    S.gst = (GState*)part_ptr[6];
    if (mixed) {
        S.states = true;
        S.Rc = (float*)part_ptr[7];
        S.Z = (float*)part_ptr[8];
        S.W = (float*)part_ptr[9];
        S.kp = cycle_steps(S.m, true);
        S.scl = S.sv;
    }
    S.B.ctx = ctx;
    S.B.op = op;
    S.B.nslots = nt;
    if (G) {
        S.B.grp = G;
        S.sl = SlotLen{G->slen, (int)S.n16};
        S.inner_cnt = G->inner_cnt;
        S.acc_cnt = G->acc_cnt;
        S.chk_cnt = G->chk_cnt;
        int rc0 = set_ints(ctx->stream, G->inner_cnt, inner_s.data(), nt);
        if (!rc0) rc0 = set_ints(ctx->stream, G->acc_cnt, acc_s.data(), nt);
        if (!rc0) rc0 = set_ints(ctx->stream, G->chk_cnt, chk_s.data(), nt);
        if (rc0) return rc0;
    }
    // host staging: the context's pinned buffer; an arena solve (no allocation
    // of any kind) stages through ordinary host memory
    std::vector<double> host_pageable;
    if (arena) {
        host_pageable.resize((size_t)std::max(nt, 1) * 3 + 8);
        S.host = host_pageable.data();
    } else {
        const int rc0 = ctx->ensure_pinned((size_t)std::max(nt, 1) * 3 + 8);
        if (rc0) return rc0;
        S.host = ctx->pinned;
    }

    const long long vs = (long long)n16;
    const long long os = (long long)(nbn + 1) * n16;
    dim3 g(grid1(n16).x, nt);
    std::vector<int> ones(nt, 1), tv(tracers, tracers + nt);
    S.B.tmap_identity = !G;
    for (int s = 0; s < nt; ++s)
        if (tracers[s] != s) S.B.tmap_identity = false;
    { int rc0 = set_ints(st, S.B.tmap_dev, tv.data(), nt); if (rc0) return rc0; }
    { int rc0 = set_ints(st, S.B.part_dev, ones.data(), nt); if (rc0) return rc0; }
    { int rc0 = set_ints(st, S.B.active_dev, ones.data(), nt); if (rc0) return rc0; }
    // inputs -> device layout (forward: dtau*e; adjoint: R)
    const double scale = dir == FWD ? op->dtau : 1.0;
    if (o.rhs_is_s) {
        // homogeneous sweeps: no forcing; s itself enters every residual
        hipLaunchKernelGGL(k_zero, dim3(grid1((size_t)nbn * n16).x, nt), dim3(256), 0, st, S.E, (long long)nbn * n16,
                           (int)(nbn * n16), S.sl);
        hipLaunchKernelGGL(k_ext_to_int, grid_permute(op->ncol, nt), dim3(256), 0, st, in_dev, (long long)op->N, S.sv,
                           (long long)n16, 1, op->nz, op->ncol, 1.0);
    } else if (G) {
        hipLaunchKernelGGL(k_ext_to_int_g, grid_permute(G->ncolmax, nt * nbn), dim3(256), 0, st, G->in_ptrs, G->sncol,
                           G->snz, S.E, (long long)nbn * n16, (long long)n16, nbn, scale);
    } else {
        hipLaunchKernelGGL(k_ext_to_int, grid_permute(op->ncol, nt * nbn), dim3(256), 0, st, in_dev,
                           (long long)nbn * op->N, S.E, (long long)nbn * n16, nbn, op->nz, op->ncol, scale);
    }
    ORBIT_HIP_CHECK(hipGetLastError());

    std::vector<double> snorm(nt, 0.0);
    std::vector<int> part(nt, 1), conv(nt, 0), zero(nt, 0), iters(nt, 0), rounds(nt, 0);
    std::vector<double> rnorm(nt, 0.0), last_rel(nt, 1.0);
    const int max_cycles = o.max_rounds > 0 ? o.max_rounds : o.maxiter;
    int cycles = 0;
    // the part mask on the device, uploaded only when it changes
    std::vector<int> part_up(ones);
    auto upload_part = [&]() {
        if (part_up == part) return ORBIT_OK;
        part_up = part;
        return set_ints(st, S.B.part_dev, part.data(), nt);
    };
    auto mark_zero_sources = [&]() {
        for (int s = 0; s < nt; ++s)
            if (snorm[s] < 1e-30) {
                zero[s] = 1;
                part[s] = 0;
                conv[s] = 1;
            }
    };

    if (mixed) {
        // Orbit-state refinement.  The states x_0 .. x_nb live in S.orbit (FP64).
        // A round: the exact residual of every bin's equation (one launch), the
        // residual sweep (FP32: the periodic residual s - (I - M)c, the check and
        // the right-hand side), one GMRES cycle on (I - M~) delta = that residual
        // whose monodromy applies keep their FP32 states, and the correction of
        // every state from them.  No accurate inner solve is run: the residual is
        // exact, and the residual sweep reports it to the FP32 sweeps' accuracy.
        const int tfree = dir == FWD ? 0 : nbn;  // the state GMRES solves for
        const int tend = nbn - tfree;            // tied to it: x_end = x_free - s (s: given s, else 0)
        bool xzero = false;
        bool need_s = true;                      // ||s~|| still to be measured
        if (warm_dev && o.warm_orbit) {
            // a whole earlier orbit starts every state
            hipLaunchKernelGGL(k_ext_to_int, grid_permute(op->ncol, nt * (nbn + 1)), dim3(256), 0, st, warm_dev,
                               (long long)(nbn + 1) * op->N, S.orbit, os, nbn + 1, op->nz, op->ncol, 1.0);
        } else {
            hipLaunchKernelGGL(k_zero, dim3(grid1((size_t)(nbn + 1) * n16).x, nt), dim3(256), 0, st, S.orbit, os,
                               (int)((nbn + 1) * n16), S.sl);
            if (warm_dev)  // c0 only: the first residual sweep builds the other states from it
                hipLaunchKernelGGL(k_ext_to_int, grid_permute(op->ncol, nt), dim3(256), 0, st, warm_dev,
                                   (long long)op->N, S.orbit + (size_t)tfree * n16, os, 1, op->nz, op->ncol, 1.0);
            else
                xzero = !S.sv;
        }
        // the free state in c and the closure of the other end
        hipLaunchKernelGGL(k_axpby, g, dim3(256), 0, st, S.orbit + (size_t)tfree * n16, os,
                           S.orbit + (size_t)tfree * n16, os, 0.0, S.c, vs, (int)n16, nullptr, S.sl);
        if (S.sv)
            hipLaunchKernelGGL(k_axpby, g, dim3(256), 0, st, S.c, vs, S.sv, vs, -1.0, S.orbit + (size_t)tend * n16, os,
                               (int)n16, nullptr, S.sl);
        else
            hipLaunchKernelGGL(k_axpby, g, dim3(256), 0, st, S.c, vs, S.c, vs, 0.0, S.orbit + (size_t)tend * n16, os,
                               (int)n16, nullptr, S.sl);
        if (S.sv) {
            rc = norms(S, S.sv, vs, (int)n16, snorm);
            if (rc) return rc;
            need_s = false;
        } else if (warm_dev && !next_pass) {
            // a warm start without the previous s~: one residual sweep of the
            // zero states gives it
            rc = cyc_residual(S, true);
            if (!rc) rc = z_sweep(S);
            if (!rc) rc = norms(S, S.r, vs, (int)n16, snorm);
            if (rc) return rc;
            need_s = false;
        }
        if (!need_s) mark_zero_sources();
        bool first = true;
        while (true) {
            { int rc0 = upload_part(); if (rc0) return rc0; }
            rc = cyc_residual(S, first && xzero);
            if (!rc) rc = z_sweep(S);
            if (rc) return rc;
            if (first && pd) {
                // s~ of this pass: the previous pass's s~ plus the residual of the
                // previous orbit under the new forcing (off by that orbit's own
                // residual, at most tol ||s~||); the first pass's is the residual
                // of the zero states.  ||r|| and ||s~|| with one host round trip.
                if (next_pass)
                    hipLaunchKernelGGL(k_axpby, g, dim3(256), 0, st, pd->s_state, vs, S.r, vs, 1.0, pd->s_state, vs,
                                       (int)n16, nullptr, S.sl);
                else
                    ORBIT_HIP_CHECK(hipMemcpyAsync(pd->s_state, S.r, (size_t)nt * n16 * 8, hipMemcpyDeviceToDevice, st));
                norms_dev(S, S.r, vs, (int)n16, S.h1);
                norms_dev(S, pd->s_state, vs, (int)n16, S.h1 + nt);
                ORBIT_HIP_CHECK(hipMemcpyAsync(S.host, S.h1, (size_t)2 * nt * 8, hipMemcpyDeviceToHost, st));
                ORBIT_HIP_CHECK(hipStreamSynchronize(st));
                for (int s = 0; s < nt; ++s) {
                    rnorm[s] = std::sqrt(S.host[s]);
                    snorm[s] = std::sqrt(S.host[nt + s]);
                }
            } else {
                std::vector<double> rn;
                rc = norms(S, S.r, vs, (int)n16, rn);
                if (rc) return rc;
                for (int s = 0; s < nt; ++s)
                    if (part[s]) rnorm[s] = rn[s];
                if (first && need_s) snorm = rn;  // the residual of the zero states is s~
            }
            if (first && need_s) mark_zero_sources();
            bool any = false;
            std::vector<int> done(nt, 0);
            bool any_done = false;
            for (int s = 0; s < nt; ++s) {
                if (!part[s]) continue;
                last_rel[s] = rnorm[s] / snorm[s];
                if (!std::isfinite(last_rel[s]) || !std::isfinite(snorm[s])) {
                    part[s] = 0;  // a non-finite input or state: stops, not converged
                } else if (rnorm[s] <= o.tol * snorm[s]) {
                    conv[s] = 1;
                    part[s] = 0;
                    done[s] = 1;
                    any_done = true;
                } else {
                    any = true;
                }
            }
            if (any_done) {
                // A converged slot's states take the residual sweep's (the
                // correction with delta = 0, the end state its last one): every
                // bin's equation then holds to the FP32 sweeps' accuracy, and the
                // end state differs from the free one by the periodic residual, as
                // when the states are propagated from c.  The free state and the
                // periodic residual do not change.
                std::vector<int> zk(nt, 0);
                { int rc0 = set_ints(st, S.kcols, zk.data(), nt); if (rc0) return rc0; }
                { int rc0 = set_ints(st, S.B.part_dev, done.data(), nt); if (rc0) return rc0; }
                part_up = done;
                const long long zs = (long long)(nbn - 1) * n16;
                hipLaunchKernelGGL(k_cyc_combine, dim3((unsigned)((n16 + 1023) / 1024), nt, nbn + 1), dim3(256), 0, st,
                                   S.orbit, os, S.Z,
                                   zs, S.W, (long long)nt * zs, zs, S.h1, S.kcols, S.c, vs, S.scl, vs,
                                   (const double*)S.r, (int)n16, nbn, dir == ADJ ? 1 : 0, S.B.part_dev, S.sl);
                ORBIT_HIP_CHECK(hipGetLastError());
            }
            if (o.verbose) {
                fprintf(stderr, "[orbit] round %d:", cycles);
                for (int s = 0; s < nt; ++s) fprintf(stderr, " %.2e", last_rel[s]);
                fprintf(stderr, "\n");
            }
            first = false;
            if (!any || cycles >= max_cycles) break;
            // A round cannot gain more than the FP32 applies' error allows (the
            // residual sweep's and the monodromy's, a few rho^sweeps relative): a
            // cycle stops at 3 rho^sweeps (or 1e-6, or well below tol).  Measured at
            // 64x70x15 with FCT: 25.4 ms per step against 25.8 at 1x, 26.5 at 0.3x,
            // 26.0 at 10x (one round more).
            std::vector<double> target(nt, 0.0);
            for (int s = 0; s < nt; ++s)
                target[s] = std::max(0.3 * o.tol * snorm[s], std::max(3.0 * eps_in_s[s], 1e-6) * rnorm[s]);
            { int rc0 = upload_part(); if (rc0) return rc0; }
            std::vector<int> steps;
            rc = gmres_cycle(S, part, target, steps);
            if (rc) return rc;
            for (int s = 0; s < nt; ++s)
                if (part[s]) {
                    iters[s] += steps[s];
                    rounds[s] += 1;
                }
            ++cycles;
        }
    } else {
        // Plain FP64: the reference's restarted GMRES on (I - M)c = s with the
        // accurate sweeps as M; every check re-propagates the orbit from c.
        // s (forward) / s_adj (adjoint): accurate forced sweep from zero; x0 = s
        const double* sptr;
        long long sstride;
        if (S.sv) {
            sptr = S.sv;
            sstride = vs;
        } else {
            rc = forced_sweep(S, true, S.acc);
            if (rc) return rc;
            sptr = dir == FWD ? S.orbit + (size_t)nbn * n16 : S.orbit;
            sstride = os;
        }
        rc = norms(S, sptr, sstride, (int)n16, snorm);
        if (rc) return rc;
        rnorm = snorm;
        if (warm_dev) {
            // c0, or the free state of a whole earlier orbit
            const int tfree = dir == FWD ? 0 : nbn;
            const long long wst = o.warm_orbit ? (long long)(nbn + 1) * op->N : (long long)op->N;
            const double* w0 = warm_dev + (o.warm_orbit ? (size_t)tfree * op->N : 0);
            hipLaunchKernelGGL(k_ext_to_int, grid_permute(op->ncol, nt), dim3(256), 0, st, w0, wst, S.c, vs, 1,
                               op->nz, op->ncol, 1.0);
        } else {
            hipLaunchKernelGGL(k_axpby, g, dim3(256), 0, st, sptr, sstride, sptr, sstride, 0.0, S.c, vs, (int)n16,
                               nullptr, S.sl);
        }
        mark_zero_sources();
        while (true) {
            { int rc0 = upload_part(); if (rc0) return rc0; }
            rc = forced_sweep(S, false, S.acc);
            if (rc) return rc;
            rc = norms(S, S.r, vs, (int)n16, rnorm);
            if (rc) return rc;
            bool any = false;
            for (int s = 0; s < nt; ++s) {
                if (!part[s]) continue;
                last_rel[s] = rnorm[s] / snorm[s];
                if (!std::isfinite(last_rel[s]) || !std::isfinite(snorm[s])) {
                    part[s] = 0;  // a non-finite input or state: stops, not converged
                } else if (rnorm[s] <= o.tol * snorm[s]) {
                    conv[s] = 1;
                    part[s] = 0;
                } else {
                    any = true;
                }
            }
            if (o.verbose) {
                fprintf(stderr, "[orbit] cycle %d (%d sweeps):", cycles, S.acc);
                for (int s = 0; s < nt; ++s) fprintf(stderr, " %.2e", last_rel[s]);
                fprintf(stderr, "\n");
            }
            if (!any || cycles >= max_cycles) break;
            std::vector<double> target(nt, 0.0);
            for (int s = 0; s < nt; ++s) target[s] = o.tol * snorm[s];
            { int rc0 = upload_part(); if (rc0) return rc0; }
            std::vector<int> steps;
            rc = gmres_cycle(S, part, target, steps);
            if (rc) return rc;
            for (int s = 0; s < nt; ++s)
                if (part[s]) {
                    iters[s] += steps[s];
                    rounds[s] += 1;
                }
            ++cycles;
        }
    }
    // zero-source slots: zero orbit (a slot that converged earlier kept its orbit)
    std::vector<int> zmask(nt, 0);
    bool anyzero = false;
    for (int s = 0; s < nt; ++s)
        if (zero[s]) {
            zmask[s] = 1;
            anyzero = true;
        }
    if (anyzero) {
        { int rc0 = set_ints(st, S.B.part_dev, zmask.data(), nt); if (rc0) return rc0; }
        for (int k = 0; k <= nbn; ++k)
            hipLaunchKernelGGL(k_axpby, g, dim3(256), 0, st, S.orbit + (size_t)k * n16, os, S.orbit + (size_t)k * n16,
                               os, -1.0, S.orbit + (size_t)k * n16, os, (int)n16, S.B.part_dev, S.sl);
        hipLaunchKernelGGL(k_axpby, g, dim3(256), 0, st, S.c, vs, S.c, vs, -1.0, S.c, vs, (int)n16, S.B.part_dev, S.sl);
    }
    // adjoint output convention: state 0 and state nb are the GMRES solution lam0
    norms_dev(S, S.c, vs, (int)n16, S.h1);  // ||c||^2, read back with the negative counts
    if (dir == ADJ)
        hipLaunchKernelGGL(k_axpby, g, dim3(256), 0, st, S.orbit + (size_t)nbn * n16, os, S.orbit + (size_t)nbn * n16,
                           os, 0.0, S.orbit, os, (int)n16, nullptr, S.sl);
    // negative fraction over the 9 states
    {
        const int n = (int)((nbn + 1) * n16);
        const int nb = (n + RED_CHUNK - 1) / RED_CHUNK;
        hipLaunchKernelGGL(k_negcount, dim3(nb, nt), dim3(RED_THREADS), 0, st, S.orbit, os, n, nb, S.partials, S.sl);
        hipLaunchKernelGGL(k_sum_partials, grid_waves(nt), dim3(256), 0, st, S.partials, nb, S.scal, nt);
        ORBIT_HIP_CHECK(hipMemcpyAsync(S.host, S.scal, nt * 8, hipMemcpyDeviceToHost, st));
        ORBIT_HIP_CHECK(hipMemcpyAsync(S.host + nt, S.h1, nt * 8, hipMemcpyDeviceToHost, st));
    }
    if (G)
        hipLaunchKernelGGL(k_int_to_ext_g, grid_permute(G->ncolmax, nt * (nbn + 1)), dim3(256), 0, st, S.orbit, os,
                           (long long)n16, G->out_ptrs, G->sncol, G->snz, nbn + 1);
    else
        hipLaunchKernelGGL(k_int_to_ext, grid_permute(op->ncol, nt * (nbn + 1)), dim3(256), 0, st, S.orbit, os,
                           out_dev, (long long)(nbn + 1) * op->N, nbn + 1, op->nz, op->ncol);
    ORBIT_HIP_CHECK(hipGetLastError());
    ORBIT_HIP_CHECK(hipStreamSynchronize(st));
    bool all_conv = true;
    for (int s = 0; s < nt; ++s) {
        orbit_info& I = info[s];
        I.gmres_iters = iters[s];
        I.rounds = rounds[s];
        I.converged = conv[s];
        I.zero_source = zero[s];
        I.s_norm = snorm[s];
        I.residual = zero[s] ? 0.0 : last_rel[s];
        I.periodicity = zero[s] ? 0.0 : (last_rel[s] * snorm[s]) / (std::sqrt(S.host[nt + s]) + 1e-30);
        I.neg_frac = S.host[s] / ((double)(nbn + 1) * op_of(s)->N);
        if (!conv[s]) all_conv = false;
    }
    return all_conv ? ORBIT_OK : ORBIT_WARN_NOT_CONVERGED;
}

int host_solve(orbit_ctx* ctx, orbit_opset* op, const int* tracers, int nt, const double* in,
               const double* warm, const orbit_opts* opts, double* out, orbit_info* info, int dir) {
    if (!ctx || !op || !in || !out || nt < 1) return fail(ORBIT_ERR_INVALID, "solve: NULL argument");
    hipSetDevice(ctx->device);
    const size_t nin = (size_t)nt * ((opts && opts->rhs_is_s) ? 1 : op->nbins) * op->N,
                 nout = (size_t)nt * (op->nbins + 1) * op->N,
                 nw = (size_t)nt * ((opts && opts->warm_orbit) ? op->nbins + 1 : 1) * op->N;
    double *din = nullptr, *dout = nullptr, *dw = nullptr;
    ORBIT_HIP_CHECK(hipMalloc(&din, nin * 8));
    if (hipMalloc(&dout, nout * 8) != hipSuccess) {
        hipFree(din);
        return fail(ORBIT_ERR_NOMEM, "solve: device allocation failed");
    }
    if (warm && hipMalloc(&dw, nw * 8) != hipSuccess) {
        hipFree(din);
        hipFree(dout);
        return fail(ORBIT_ERR_NOMEM, "solve: device allocation failed");
    }
    int rc = ORBIT_OK;
    if (hipMemcpyAsync(din, in, nin * 8, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        (warm && hipMemcpyAsync(dw, warm, nw * 8, hipMemcpyHostToDevice, ctx->stream) != hipSuccess))
        rc = fail(ORBIT_ERR_HIP, "solve: host to device copy failed");
    if (rc == ORBIT_OK) rc = solve_periodic(ctx, op, tracers, nt, din, dw, opts, dout, info, dir);
    if (rc >= 0) {
        if (hipMemcpyAsync(out, dout, nout * 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = fail(ORBIT_ERR_HIP, "solve: device to host copy failed");
    }
    hipStreamSynchronize(ctx->stream);
    hipFree(din);
    hipFree(dout);
    if (dw) hipFree(dw);
    return rc;
}

}  // namespace

extern "C" int orbit_solve_forward_dev(orbit_ctx* ctx, orbit_opset* op, const int* tracers, int nt,
                                       const double* e_dev, const double* c_warm_dev, const orbit_opts* opts,
                                       double* orbit_out_dev, orbit_info* info) {
    return solve_periodic(ctx, op, tracers, nt, e_dev, c_warm_dev, opts, orbit_out_dev, info, FWD);
}

// The forward solve with no device allocation: scratch from this device's
// fixed array, no graph (instantiating one allocates).  Calls that overlap on
// one GPU share the array and race.
extern "C" int orbit_solve_forward_dev_noalloc(orbit_ctx* ctx, orbit_opset* op, const int* tracers, int nt,
                                               const double* e_dev, const double* c_warm_dev,
                                               const orbit_opts* opts, double* orbit_out_dev, orbit_info* info) {
    if (!ctx || !op) return fail(ORBIT_ERR_INVALID, "solve (no allocation): NULL context or operator set");
    hipSetDevice(ctx->device);
    if (!ctx->fixed_scratch) {
        const int rc = fixed_scratch_address(ctx->device, &ctx->fixed_scratch);
        if (rc) return rc;
    }
    orbit_opts o;
    orbit_opts_default(&o);
    if (opts) o = *opts;
    o.use_graph = 0;
    return solve_periodic(ctx, op, tracers, nt, e_dev, c_warm_dev, &o, orbit_out_dev, info, FWD, nullptr,
                          ctx->fixed_scratch, (size_t)ORBIT_FIXED_SCRATCH_BYTES);
}

int orbit::solve_forward_pass(orbit_ctx* ctx, orbit_opset* op, const int* tracers, int nt, const double* e_dev,
                              const double* prev_orbit_dev, PassDelta* pd, const orbit_opts* opts, double* out_dev,
                              orbit_info* info) {
    return solve_periodic(ctx, op, tracers, nt, e_dev, prev_orbit_dev, opts, out_dev, info, FWD, pd);
}

extern "C" int orbit_solve_adjoint_dev(orbit_ctx* ctx, orbit_opset* op, const int* tracers, int nt,
                                       const double* R_dev, const double* lam_warm_dev, const orbit_opts* opts,
                                       double* lam_out_dev, orbit_info* info) {
    return solve_periodic(ctx, op, tracers, nt, R_dev, lam_warm_dev, opts, lam_out_dev, info, ADJ);
}

extern "C" int orbit_solve_forward(orbit_ctx* ctx, orbit_opset* op, const int* tracers, int nt, const double* e,
                                   const double* c_warm, const orbit_opts* opts, double* orbit_out,
                                   orbit_info* info) {
    return host_solve(ctx, op, tracers, nt, e, c_warm, opts, orbit_out, info, FWD);
}

extern "C" int orbit_solve_adjoint(orbit_ctx* ctx, orbit_opset* op, const int* tracers, int nt, const double* R,
                                   const double* lam_warm, const orbit_opts* opts, double* lam_out,
                                   orbit_info* info) {
    return host_solve(ctx, op, tracers, nt, R, lam_warm, opts, lam_out, info, ADJ);
}

// Periodic solves of slots from several operator sets in one batch: every
// sweep, vector and reduction kernel covers all slots in one launch.  Slot s
// solves tracer tracers[s] of sets[slot_set[s]] for the device input in_dev[s]
// ((n_bins, N) of its set: e, or R with adjoint = 1) into out_dev[s].
extern "C" int orbit_solve_group_dev(orbit_ctx* ctx, int nsets, orbit_opset* const* sets, int nt, const int* slot_set,
                                     const int* tracers, const double* const* in_dev, int adjoint,
                                     const orbit_opts* opts, double* const* out_dev, orbit_info* info) {
    if (!ctx || !sets || nsets < 1 || nt < 1 || !slot_set || !tracers || !in_dev || !out_dev || !info)
        return fail(ORBIT_ERR_INVALID, "group solve: NULL argument or no slots");
    if (adjoint != 0 && adjoint != 1) return fail(ORBIT_ERR_INVALID, "group solve: adjoint must be 0 or 1");
    for (int q = 0; q < nsets; ++q) {
        const orbit_opset* so = sets[q];
        if (!so) return fail(ORBIT_ERR_INVALID, "group solve: NULL operator set");
        if (so->ngroups != 1)
            return fail(ORBIT_ERR_INVALID, "group solve: every operator set needs one shared transport per bin "
                                           "(orbit_opset_create_shared or orbit_opset_create_grids)");
        if (so->ncolours != 2)
            return fail(ORBIT_ERR_INVALID, "group solve: a periodic longitude with an odd nx (four sweep colours) "
                                           "is not supported here; solve that set alone");
        if (so->nbins != sets[0]->nbins || so->dtau != sets[0]->dtau)
            return fail(ORBIT_ERR_INVALID, "group solve: every operator set needs the same n_bins and dtau");
    }
    for (int s = 0; s < nt; ++s) {
        if (slot_set[s] < 0 || slot_set[s] >= nsets) return fail(ORBIT_ERR_INVALID, "group solve: slot_set out of range");
        if (tracers[s] < 0 || tracers[s] >= sets[slot_set[s]]->ntr)
            return fail(ORBIT_ERR_INVALID, "group solve: tracer index out of range");
        if (!in_dev[s] || !out_dev[s]) return fail(ORBIT_ERR_INVALID, "group solve: NULL input or output pointer");
    }
    orbit_opts o;
    orbit_opts_default(&o);
    if (opts) o = *opts;
    if (o.rhs_is_s || o.warm_orbit) return fail(ORBIT_ERR_INVALID, "group solve: starts cold from e (or R)");
    hipSetDevice(ctx->device);
    const int nbn = sets[0]->nbins;
    // slots by set (stable): up to four consecutive slots of one set form a slot group
    std::vector<int> perm(nt);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return slot_set[a] < slot_set[b]; });
    GroupSpec G;
    struct Grp { int set, first, count; };
    std::vector<Grp> grps;
    std::vector<int> tmap(nt);
    G.mf = true;
    G.handoff = true;
    for (int k = 0; k < nt; ++k) {
        const int s = perm[k], q = slot_set[s];
        orbit_opset* so = sets[q];
        G.slot_op.push_back(so);
        tmap[k] = tracers[s];
        G.n16max = std::max(G.n16max, (size_t)so->n16);
        G.ncolmax = std::max(G.ncolmax, so->ncol);
        G.mf = G.mf && so->mfma;
        G.handoff = G.handoff && so->handoff;
        if (grps.empty() || grps.back().set != q || grps.back().count == 4) grps.push_back({q, k, 1});
        else grps.back().count += 1;
    }
    G.ngroups = (int)grps.size();
    if (G.ngroups >= 32768) return fail(ORBIT_ERR_INVALID, "group solve: too many slot groups");
    const bool T = adjoint != 0;
    std::vector<GroupOp> hg((size_t)nbn * N_SWEEP_PRECS * G.ngroups);
    for (int b = 0; b < nbn; ++b)
        for (int p = 0; p < N_SWEEP_PRECS; ++p)
            for (int g = 0; g < G.ngroups; ++g) {
                const orbit_opset* so = sets[grps[g].set];
                const BinOps& O = so->bins[b];
                GroupOp& go = hg[((size_t)b * N_SWEEP_PRECS + p) * G.ngroups + g];
                if (p == SP_F32) {
                    go.V = T ? O.VT32 : O.V32; go.Vt = T ? O.V32 : O.VT32; go.H = T ? O.HT32 : O.H32; go.d = O.d32;
                    go.P = T ? O.PT16 : O.P16;
                } else {
                    go.V = T ? O.VT64 : O.V64; go.Vt = T ? O.V64 : O.VT64; go.H = T ? O.HT64 : O.H64; go.d = O.d64;
                    go.P = p == SP_F64 ? (T ? (const void*)O.PT64 : (const void*)O.P64)
                                       : (T ? (const void*)O.PT16 : (const void*)O.P16);
                }
                go.dstride = so->n16;
                go.pstride = (long long)so->ncol * 256;
                for (int c = 0; c < 4; ++c) go.ncols[c] = so->ncols_c[c];
                go.nx = so->nx; go.ny = so->ny; go.nxc = so->nxc; go.px = so->px;
                go.first = grps[g].first;
                go.count = grps[g].count;
                go.omega = (float)so->omega;
            }
    // per colour, the (group, tile) of every block: by set, then tile, then the
    // set's groups, so a tile's groups are neighbours; XCD-contiguous ranges
    std::vector<int> tabs[2];
    for (int c = 0; c < 2; ++c) {
        std::vector<int> pairs;
        for (int g0 = 0; g0 < G.ngroups;) {
            int g1 = g0;
            while (g1 < G.ngroups && grps[g1].set == grps[g0].set) ++g1;
            const int ntile = (sets[grps[g0].set]->ncols_c[c] + 3) / 4;
            if (ntile >= 65536) return fail(ORBIT_ERR_INVALID, "group solve: grid too large for the block table");
            for (int t = 0; t < ntile; ++t)
                for (int g = g0; g < g1; ++g) pairs.push_back(g * 65536 + t);
            g0 = g1;
        }
        const size_t P = pairs.size(), per = (P + 7) / 8;
        tabs[c].assign(8 * per, -1);
        for (size_t x = 0; x < 8; ++x)
            for (size_t r = 0; r < per; ++r)
                if (x * per + r < P) tabs[c][8 * r + x] = pairs[x * per + r];
        G.nblocks[c] = (int)(8 * per);
        for (int p = 0; p < N_SWEEP_PRECS; ++p) {
            const size_t ae = p == SP_F32 ? 4 : 8, pe = p == SP_F64 ? 8 : 2, xe = p == SP_F32 ? 4 : 8;
            double bytes = 0.0;
            for (const Grp& gr : grps)
                bytes += halfsweep_bytes(sets[gr.set]->ncols_c[c], gr.count, ae, pe, xe, false, false);
            G.bytes[p][c] = bytes;
        }
    }
    // the first sweep's (group, patch) of every block: 4 x 2 column patches,
    // ordered like the half-sweeps' tiles
    std::vector<int> ptab;
    {
        std::vector<int> pairs;
        for (int g0 = 0; g0 < G.ngroups;) {
            int g1 = g0;
            while (g1 < G.ngroups && grps[g1].set == grps[g0].set) ++g1;
            const orbit_opset* so = sets[grps[g0].set];
            const int npat = ((so->nx + 3) / 4) * ((so->ny + 1) / 2);
            if (npat >= 65536) return fail(ORBIT_ERR_INVALID, "group solve: grid too large for the block table");
            for (int t = 0; t < npat; ++t)
                for (int g = g0; g < g1; ++g) pairs.push_back(g * 65536 + t);
            g0 = g1;
        }
        const size_t P = pairs.size(), per = (P + 7) / 8;
        ptab.assign(8 * per, -1);
        for (size_t x = 0; x < 8; ++x)
            for (size_t r = 0; r < per; ++r)
                if (x * per + r < P) ptab[8 * r + x] = pairs[x * per + r];
        G.npblocks = (int)(8 * per);
        // each operand once: both colours' P, b and x written, the black columns' stencil
        for (const Grp& gr : grps) {
            const orbit_opset* so = sets[gr.set];
            G.first_bytes += halfsweep_bytes(so->ncols_c[0], gr.count, 4, 2, 4, true, true) +
                             halfsweep_bytes(so->ncols_c[1], gr.count, 4, 2, 4, true, true) + 256.0 * so->ncols_c[1];
        }
    }
    // one upload of every table
    std::vector<int> slen(nt), sncol(nt), snz(nt);
    std::vector<const double*> inp(nt);
    std::vector<double*> outp(nt);
    for (int k = 0; k < nt; ++k) {
        slen[k] = G.slot_op[k]->n16;
        sncol[k] = G.slot_op[k]->ncol;
        snz[k] = G.slot_op[k]->nz;
        inp[k] = in_dev[perm[k]];
        outp[k] = out_dev[perm[k]];
    }
    size_t off = 0;
    auto place = [&](size_t bytes) { const size_t o0 = off; off += align256(bytes); return o0; };
    const size_t o_gops = place(hg.size() * sizeof(GroupOp)), o_t0 = place(tabs[0].size() * 4),
                 o_t1 = place(tabs[1].size() * 4), o_pt = place(ptab.size() * 4), o_slen = place((size_t)nt * 4),
                 o_ncol = place((size_t)nt * 4),
                 o_nz = place((size_t)nt * 4), o_in = place((size_t)nt * 8), o_out = place((size_t)nt * 8),
                 o_ic = place((size_t)nt * 4), o_ac = place((size_t)nt * 4), o_cc = place((size_t)nt * 4);
    std::vector<char> blob(off, 0);
    auto put = [&](size_t o0, const void* src, size_t bytes) { if (bytes) memcpy(blob.data() + o0, src, bytes); };
    put(o_gops, hg.data(), hg.size() * sizeof(GroupOp));
    put(o_t0, tabs[0].data(), tabs[0].size() * 4);
    put(o_t1, tabs[1].data(), tabs[1].size() * 4);
    put(o_pt, ptab.data(), ptab.size() * 4);
    put(o_slen, slen.data(), (size_t)nt * 4);
    put(o_ncol, sncol.data(), (size_t)nt * 4);
    put(o_nz, snz.data(), (size_t)nt * 4);
    put(o_in, inp.data(), (size_t)nt * 8);
    put(o_out, outp.data(), (size_t)nt * 8);
    int rc = ctx->ensure(9, off);
    if (rc) return rc;
    char* dev = (char*)ctx->scratch[9].p;
    rc = h2d_sync(ctx->stream, dev, blob.data(), off);
    if (rc) return rc;
    G.gops = (const GroupOp*)(dev + o_gops);
    G.btab[0] = (const int*)(dev + o_t0);
    G.btab[1] = (const int*)(dev + o_t1);
    G.ptab = (const int*)(dev + o_pt);
    G.slen = (const int*)(dev + o_slen);
    G.sncol = (const int*)(dev + o_ncol);
    G.snz = (const int*)(dev + o_nz);
    G.in_ptrs = (const double* const*)(dev + o_in);
    G.out_ptrs = (double* const*)(dev + o_out);
    G.inner_cnt = (int*)(dev + o_ic);
    G.acc_cnt = (int*)(dev + o_ac);
    G.chk_cnt = (int*)(dev + o_cc);
    std::vector<orbit_info> inf(nt);
    rc = solve_periodic(ctx, G.slot_op[0], tmap.data(), nt, inp[0], nullptr, &o, outp[0], inf.data(),
                        adjoint ? ADJ : FWD, nullptr, nullptr, 0, &G);
    if (rc < 0) return rc;
    for (int k = 0; k < nt; ++k) info[perm[k]] = inf[k];
    return rc;
}

// LU-protocol shim: accurate single solve, sweeps until ||b - A x|| <= 1e-14 ||b||.
extern "C" int orbit_apply_P(orbit_ctx* ctx, orbit_opset* op, int bin, int tracer, int trans,
                             const double* b, double* x) {
    if (!ctx || !op || !b || !x) return fail(ORBIT_ERR_INVALID, "apply_P: NULL argument");
    if (bin < 0 || bin >= op->nbins) return fail(ORBIT_ERR_INVALID, "apply_P: bin out of range");
    if (tracer < 0 || tracer >= op->ntr) return fail(ORBIT_ERR_INVALID, "apply_P: tracer out of range");
    if (trans != 0 && trans != 1) return fail(ORBIT_ERR_INVALID, "apply_P: trans must be 0 or 1");
    hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    const size_t n16 = op->n16;
    const int nb = (int)((n16 + RED_CHUNK - 1) / RED_CHUNK);
    int rc = ctx->ensure(7, (n16 * 4 + (size_t)nb + 8 + (size_t)op->N) * 8 + 64);
    if (rc) return rc;
    double* base = ctx->buf<double>(7);
    double* rhs = base;
    double* xs = base + n16;
    double* y = base + 2 * n16;
    double* rr = base + 3 * n16;
    double* red = base + 4 * n16;       // nb partials, then the sum at red[nb]
    double* ext = red + nb + 8;
    int* tmap = (int*)(ext + op->N);
    { int rc0 = h2d_sync(st, tmap, &tracer, 4); if (rc0) return rc0; }
    ORBIT_HIP_CHECK(hipMemcpyAsync(ext, b, (size_t)op->N * 8, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_ext_to_int, grid_permute(op->ncol, 1), dim3(256), 0, st, ext, (long long)op->N, rhs,
                       (long long)n16, 1, op->nz, op->ncol, 1.0);
    Batch B{ctx, op, 1, tmap, nullptr, nullptr};
    const int dir = trans ? ADJ : FWD;
    const BinOps& O = op->bins[bin];
    const long long vstride = op->ngroups == 1 ? 0 : (long long)op->ncol * 256;
    const long long hstride = op->ngroups == 1 ? 0 : (long long)op->ncol * 64;
    double hb[2];
    hipLaunchKernelGGL(k_norm2, dim3(nb, 1), dim3(RED_THREADS), 0, st, rhs, (long long)n16, (int)n16, nb, red, SlotLen{});
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, st, red, nb, red + nb, 1);
    ORBIT_HIP_CHECK(hipMemcpyAsync(&hb[0], red + nb, 8, hipMemcpyDeviceToHost, st));
    ORBIT_HIP_CHECK(hipStreamSynchronize(st));
    const double bn = std::sqrt(hb[0]);
    int total = 0;
    bool first = true;
    while (total < 200) {
        const int chunk = 4;
        for (int s = 0; s < chunk; ++s) {
            rc = sweep_once(B, bin, dir, SP_F64, rhs, (long long)n16, xs, (long long)n16, first, nullptr, nullptr, 0);
            if (rc) return rc;
            first = false;
        }
        total += chunk;
        hipLaunchKernelGGL(k_apply_A, dim3((op->ncol + 15) / 16, 1), dim3(256), 0, st, trans ? O.V64 : O.VT64, vstride,
                           trans ? O.HT64 : O.H64, hstride, O.d64, (long long)n16, xs, y, (long long)n16, tmap,
                           op->ncol, op->nx, op->ny, op->px);
        hipLaunchKernelGGL(k_axpby, dim3(grid1(n16).x, 1), dim3(256), 0, st, rhs, (long long)n16, y, (long long)n16, -1.0,
                           rr, (long long)n16, (int)n16, nullptr, SlotLen{});
        hipLaunchKernelGGL(k_norm2, dim3(nb, 1), dim3(RED_THREADS), 0, st, rr, (long long)n16, (int)n16, nb, red, SlotLen{});
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, st, red, nb, red + nb, 1);
        ORBIT_HIP_CHECK(hipMemcpyAsync(&hb[1], red + nb, 8, hipMemcpyDeviceToHost, st));
        ORBIT_HIP_CHECK(hipStreamSynchronize(st));
        if (!(std::sqrt(hb[1]) > 1e-14 * bn)) break;
    }
    hipLaunchKernelGGL(k_int_to_ext, grid_permute(op->ncol, 1), dim3(256), 0, st, xs, (long long)n16, ext,
                       (long long)op->N, 1, op->nz, op->ncol);
    ORBIT_HIP_CHECK(hipMemcpyAsync(x, ext, (size_t)op->N * 8, hipMemcpyDeviceToHost, st));
    ORBIT_HIP_CHECK(hipStreamSynchronize(st));
    return ORBIT_OK;
}

extern "C" int orbit_bench_halfsweep(orbit_ctx* ctx, orbit_opset* op, int bin, int precision, int nslots,
                                     int reps, double* ms_per_launch, double* bytes_per_launch) {
    if (!ctx || !op || !ms_per_launch || !bytes_per_launch) return fail(ORBIT_ERR_INVALID, "bench: NULL argument");
    if (bin < 0 || bin >= op->nbins || nslots < 1 || nslots > op->ntr || reps < 1)
        return fail(ORBIT_ERR_INVALID, "bench: bad bin, slot count or reps");
    hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    if (precision != ORBIT_PREC_FP64 && precision != ORBIT_PREC_MIXED && precision != ORBIT_SWEEP_ACC_MIXED)
        return fail(ORBIT_ERR_INVALID, "bench: unknown precision");
    const size_t n16 = op->n16;
    const bool mixed = precision == ORBIT_PREC_MIXED;
    const bool bf16 = precision != ORBIT_PREC_FP64;
    const size_t xe = mixed ? 4 : 8;
    int rc = ctx->ensure(7, 2 * (size_t)nslots * n16 * xe + (size_t)nslots * 4 + 256);
    if (rc) return rc;
    char* base = (char*)ctx->scratch[7].p;
    void* b = base;
    void* x = base + (size_t)nslots * n16 * xe;
    int* tmap = (int*)(base + 2 * (size_t)nslots * n16 * xe);
    std::vector<int> tv(nslots);
    std::iota(tv.begin(), tv.end(), 0);
    { int rc0 = h2d_sync(st, tmap, tv.data(), nslots * 4); if (rc0) return rc0; }
    ORBIT_HIP_CHECK(hipMemsetAsync(base, 0, 2 * (size_t)nslots * n16 * xe, st));
    Batch B{ctx, op, nslots, tmap, nullptr, nullptr};
    const int sp = mixed ? SP_F32 : (bf16 ? SP_ACC_MIXED : SP_F64);
    // warm up both colours once
    launch_halfsweep(B, bin, FWD, sp, 0, b, (long long)n16, x, (long long)n16, false, false, nullptr);
    launch_halfsweep(B, bin, FWD, sp, 1, b, (long long)n16, x, (long long)n16, false, false, nullptr);
    hipEvent_t e0, e1;
    ORBIT_HIP_CHECK(hipEventCreate(&e0));
    ORBIT_HIP_CHECK(hipEventCreate(&e1));
    ORBIT_HIP_CHECK(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r)
        launch_halfsweep(B, bin, FWD, sp, r & 1, b, (long long)n16, x, (long long)n16, false, false, nullptr);
    ORBIT_HIP_CHECK(hipEventRecord(e1, st));
    ORBIT_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    ORBIT_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    ORBIT_HIP_CHECK(hipGetLastError());
    *ms_per_launch = (double)ms / reps;
    // algorithmic bytes of one (average) colour launch: shared V and H of the
    // colour's columns, per slot d, P, b, x (read+write) of the colour, and the
    // other colour's x once.
    const double cols = 0.5 * op->ncol;
    const double ae = mixed ? 4.0 : 8.0, pe = bf16 ? 2.0 : 8.0;
    const double shared = cols * (256.0 * ae + 64.0 * ae);
    const double per_slot = cols * (16.0 * ae /*d*/ + 256.0 * pe /*P*/ + 16.0 * xe /*b*/ + 32.0 * xe /*x rw*/ +
                                    16.0 * xe /*nbr x*/);
    *bytes_per_launch = shared + nslots * per_slot;
    return ORBIT_OK;
}


extern "C" int orbit_timing_start(orbit_ctx* ctx, int max_launches) {
    if (!ctx) return fail(ORBIT_ERR_INVALID, "timing_start: NULL context");
    if (ctx->timing) return fail(ORBIT_ERR_INVALID, "timing_start: kernel timing is already on for this context");
    if (max_launches < 1 || max_launches > 1000000)
        return fail(ORBIT_ERR_INVALID, "timing_start: max_launches must be in [1, 1000000]");
    hipSetDevice(ctx->device);
    auto* tm = new SweepTiming();
    tm->ev.resize(2 * (size_t)max_launches, nullptr);
    for (auto& e : tm->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            for (auto& f : tm->ev)
                if (f) hipEventDestroy(f);
            delete tm;
            return fail(ORBIT_ERR_HIP, "timing_start: hipEventCreate failed");
        }
    }
    tm->recs.reserve(max_launches);
    ctx->timing = tm;
    return ORBIT_OK;
}

extern "C" int orbit_timing_stop(orbit_ctx* ctx, orbit_sweep_timing* out, int max_out, int* n_out,
                                 long long* untimed) {
    if (!ctx || !ctx->timing) return fail(ORBIT_ERR_INVALID, "timing_stop: kernel timing is not on");
    hipSetDevice(ctx->device);
    SweepTiming* tm = ctx->timing;
    ctx->timing = nullptr;
    int rc = ORBIT_OK;
    std::map<int, orbit_sweep_timing> agg;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(ORBIT_ERR_HIP, "timing_stop: stream sync failed");
    for (size_t i = 0; i < tm->recs.size() && rc == ORBIT_OK; ++i) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, tm->ev[2 * i], tm->ev[2 * i + 1]) != hipSuccess) {
            rc = fail(ORBIT_ERR_HIP, "timing_stop: hipEventElapsedTime failed");
            break;
        }
        const int v = tm->recs[i].variant;
        auto it = agg.find(v);
        if (it == agg.end()) {
            const int vf = v >= HANDOFF_VARIANT ? v - HANDOFF_VARIANT : v >= FUSED_FIRST_VARIANT ? v - FUSED_FIRST_VARIANT : v;
            const int p = vf / 16;  // internal sweep precision -> ABI constant
            orbit_sweep_timing t{};
            t.precision = p == SP_F64 ? ORBIT_PREC_FP64 : p == SP_ACC_MIXED ? ORBIT_SWEEP_ACC_MIXED : ORBIT_PREC_MIXED;
            t.mfma = (vf & 8) ? 1 : 0;
            t.slots_per_wave = (vf & 4) ? 4 : 1;
            t.own_zero = (vf & 2) ? 1 : 0;
            t.nbr_zero = (vf & 1) ? 1 : 0;
            t.first_sweep = v >= HANDOFF_VARIANT ? 2 : v >= FUSED_FIRST_VARIANT ? 1 : 0;
            it = agg.emplace(v, t).first;
        }
        it->second.launches += 1;
        it->second.total_ms += ms;
        it->second.total_bytes += tm->recs[i].bytes;
    }
    const long long dropped = tm->dropped;
    for (auto& e : tm->ev) hipEventDestroy(e);
    delete tm;
    if (rc) return rc;
    std::vector<orbit_sweep_timing> v;
    for (auto& kv : agg) v.push_back(kv.second);
    std::stable_sort(v.begin(), v.end(),
                     [](const orbit_sweep_timing& a, const orbit_sweep_timing& b) { return a.total_ms > b.total_ms; });
    const int n = std::min<int>((int)v.size(), std::max(0, max_out));
    for (int i = 0; i < n && out; ++i) out[i] = v[i];
    if (n_out) *n_out = (int)v.size();
    if (untimed) *untimed = dropped;
    return ORBIT_OK;
}

namespace {

bool shape_ok(const orbit_sweep_shape& t) {
    return (t.slots_per_wave == 1 || t.slots_per_wave == 4) &&
           (t.threads == 64 || t.threads == 128 || t.threads == 256) && (t.xcd_map == 0 || t.xcd_map == 1);
}

void drop_graphs(orbit_opset* op) {
    std::lock_guard<std::mutex> lock(op->graph_mu);
    for (auto& kv : op->graphs) hipGraphExecDestroy(kv.second);
    op->graphs.clear();
}

}  // namespace

extern "C" int orbit_opset_set_patch(orbit_opset* op, int shape) {
    if (!op) return fail(ORBIT_ERR_INVALID, "set_patch: NULL operator set");
    if (shape < 0 || shape >= N_PATCH_SHAPES) return fail(ORBIT_ERR_INVALID, "set_patch: unknown patch shape");
    op->patch = shape;
    drop_graphs(op);  // captured sweeps carry the old kernels
    return ORBIT_OK;
}

extern "C" int orbit_opset_set_handoff(orbit_opset* op, int on) {
    if (!op) return fail(ORBIT_ERR_INVALID, "set_handoff: NULL operator set");
    op->handoff = on != 0;
    drop_graphs(op);  // captured sweeps carry the old launch sequence
    return ORBIT_OK;
}

extern "C" int orbit_opset_set_mfma(orbit_opset* op, int on) {
    if (!op) return fail(ORBIT_ERR_INVALID, "set_mfma: NULL operator set");
    op->mfma = on != 0;
    drop_graphs(op);  // captured sweeps carry the old kernels
    return ORBIT_OK;
}

extern "C" int orbit_opset_get_shapes(orbit_opset* op, orbit_sweep_shape* table) {
    if (!op || !table) return fail(ORBIT_ERR_INVALID, "get_shapes: NULL argument");
    for (int p = 0; p < N_SWEEP_PRECS; ++p)
        for (int b = 0; b < N_SLOT_BUCKETS; ++b) {
            const SweepShape& sh = op->shapes[p][b];
            table[p * N_SLOT_BUCKETS + b] = {sh.cs ? 4 : 1, sh.threads, sh.xcd};
        }
    return ORBIT_OK;
}

extern "C" int orbit_opset_set_shapes(orbit_opset* op, const orbit_sweep_shape* table) {
    if (!op || !table) return fail(ORBIT_ERR_INVALID, "set_shapes: NULL argument");
    for (int q = 0; q < N_SWEEP_PRECS * N_SLOT_BUCKETS; ++q)
        if (!shape_ok(table[q]))
            return fail(ORBIT_ERR_INVALID, "set_shapes: slots_per_wave must be 1 or 4, threads 64, 128 or 256, "
                                           "xcd_map 0 or 1");
    for (int p = 0; p < N_SWEEP_PRECS; ++p)
        for (int b = 0; b < N_SLOT_BUCKETS; ++b) {
            const orbit_sweep_shape& t = table[p * N_SLOT_BUCKETS + b];
            SweepShape& sh = op->shapes[p][b];
            sh.cs = (int8_t)(t.slots_per_wave == 4 && op->ngroups == 1);
            sh.threads = (int16_t)t.threads;
            sh.xcd = (int8_t)t.xcd_map;
        }
    drop_graphs(op);  // captured sweeps carry the old launch shapes
    return ORBIT_OK;
}

// Time every launch shape of the half-sweep for each sweep precision and slot
// bucket (representative slot counts 1, 2, 4, 8, 16; slot s runs tracer
// s mod n_tracers) on bin 0.  Each candidate: two untimed launches, then `reps`
// launches alternating colours between HIP events; every candidate is timed in
// two rounds and keeps its faster round.  The shape in place before the call
// (the incumbent) is kept unless another is faster by more than
// AUTOTUNE_MARGIN: shapes differ by a few percent, and a pick within the timing
// noise can be slower in the real workload.
static constexpr double AUTOTUNE_MARGIN = 0.03;

extern "C" int orbit_opset_autotune(orbit_ctx* ctx, orbit_opset* op, int reps, orbit_sweep_shape* table,
                                    double* us_table) {
    if (!ctx || !op) return fail(ORBIT_ERR_INVALID, "autotune: NULL argument");
    if (reps < 1 || reps > 100000) return fail(ORBIT_ERR_INVALID, "autotune: reps must be in [1, 100000]");
    hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    const size_t n16 = op->n16;
    const int slots_of[N_SLOT_BUCKETS] = {1, 2, 4, 8, 16};
    const int maxs = 16;
    int rc = ctx->ensure(7, 2 * (size_t)maxs * n16 * 8 + (size_t)maxs * 4 + 256);
    if (rc) return rc;
    char* base = (char*)ctx->scratch[7].p;
    void* b = base;
    void* x = base + (size_t)maxs * n16 * 8;
    int* tmap = (int*)(base + 2 * (size_t)maxs * n16 * 8);
    std::vector<int> tv(maxs);
    for (int q = 0; q < maxs; ++q) tv[q] = q % op->ntr;
    { int rc0 = h2d_sync(st, tmap, tv.data(), maxs * 4); if (rc0) return rc0; }
    ORBIT_HIP_CHECK(hipMemsetAsync(base, 0, 2 * (size_t)maxs * n16 * 8, st));
    hipEvent_t e0, e1;
    ORBIT_HIP_CHECK(hipEventCreate(&e0));
    ORBIT_HIP_CHECK(hipEventCreate(&e1));
    int err = ORBIT_OK;
    for (int p = 0; p < N_SWEEP_PRECS && !err; ++p) {
        const long long vs = (long long)n16;  // slot stride in elements of the precision's type
        for (int bk = 0; bk < N_SLOT_BUCKETS && !err; ++bk) {
            const int ns = slots_of[bk];
            Batch B{ctx, op, ns, tmap, nullptr, nullptr};
            // candidate 0 is the incumbent
            std::vector<SweepShape> cand{op->shapes[p][bk]};
            for (int cs = 0; cs <= 1; ++cs) {
                if (cs && (op->ngroups != 1 || ns < 2)) continue;
                for (int W : {64, 128, 256})
                    for (int xcd = 0; xcd <= 1; ++xcd) cand.push_back(SweepShape{(int8_t)cs, (int16_t)W, (int8_t)xcd});
            }
            std::vector<double> us(cand.size(), 1e300);
            for (int round = 0; round < 2 && !err; ++round)
                for (size_t c = 0; c < cand.size() && !err; ++c) {
                    op->shapes[p][bk] = cand[c];
                    for (int w = 0; w < 2 && !err; ++w)
                        err = launch_halfsweep(B, 0, FWD, p, w & 1, b, vs, x, vs, false, false, nullptr);
                    if (err) break;
                    hipEventRecord(e0, st);
                    for (int r = 0; r < reps && !err; ++r)
                        err = launch_halfsweep(B, 0, FWD, p, r & 1, b, vs, x, vs, false, false, nullptr);
                    hipEventRecord(e1, st);
                    hipError_t he = hipEventSynchronize(e1);
                    if (he != hipSuccess) {
                        err = fail(ORBIT_ERR_HIP, std::string("autotune: ") + hipGetErrorString(he));
                        break;
                    }
                    float ms = 0.f;
                    hipEventElapsedTime(&ms, e0, e1);
                    us[c] = std::min(us[c], 1e3 * (double)ms / reps);
                }
            if (err) break;
            size_t best = 0;
            for (size_t c = 1; c < cand.size(); ++c)
                if (us[c] < us[best]) best = c;
            if (us[best] > (1.0 - AUTOTUNE_MARGIN) * us[0]) best = 0;
            op->shapes[p][bk] = cand[best];
            if (table) table[p * N_SLOT_BUCKETS + bk] = {cand[best].cs ? 4 : 1, cand[best].threads, cand[best].xcd};
            if (us_table) us_table[p * N_SLOT_BUCKETS + bk] = us[best];
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (err) {
        default_shapes(op);
        drop_graphs(op);
        return err;
    }
    ORBIT_HIP_CHECK(hipGetLastError());
    drop_graphs(op);
    return ORBIT_OK;
}

namespace orbit {
template <class T>
__device__ __forceinline__ bool finite_of(T v) { return isfinite((double)v); }
__device__ __forceinline__ bool finite_of(uint16_t v) { return isfinite(bf16_to_f32(v)); }

template <class T>
__global__ void k_count_nonfinite(const T* __restrict__ a, size_t n, unsigned long long* out) {
    size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    unsigned long long c = 0;
    for (; g < n; g += stride)
        if (!finite_of(a[g])) ++c;
    if (c) atomicAdd(out, c);
}
}  // namespace orbit

extern "C" int orbit_opset_check(orbit_ctx* ctx, orbit_opset* op, long long* counts) {
    if (!ctx || !op || !counts) return fail(ORBIT_ERR_INVALID, "opset_check: NULL argument");
    hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    unsigned long long* dc = nullptr;
    ORBIT_HIP_CHECK(hipMalloc(&dc, 14 * sizeof(unsigned long long)));
    {
        const hipError_t e0 = hipMemsetAsync(dc, 0, 14 * sizeof(unsigned long long), st);
        if (e0 != hipSuccess) {
            hipFree(dc);
            return fail(ORBIT_ERR_HIP, std::string("opset_check: ") + hipGetErrorString(e0));
        }
    }
    const size_t ncol = op->ncol, G = op->ngroups, T = op->ntr;
    const size_t nV = G * ncol * 256, nH = G * ncol * 64, nd = T * op->n16, nP = T * ncol * 256;
    for (const BinOps& B : op->bins) {
        dim3 g(1024), b(256);
        hipLaunchKernelGGL(k_count_nonfinite<double>, g, b, 0, st, B.V64, nV, dc + 0);
        hipLaunchKernelGGL(k_count_nonfinite<float>, g, b, 0, st, B.V32, nV, dc + 1);
        hipLaunchKernelGGL(k_count_nonfinite<double>, g, b, 0, st, B.VT64, nV, dc + 2);
        hipLaunchKernelGGL(k_count_nonfinite<float>, g, b, 0, st, B.VT32, nV, dc + 3);
        hipLaunchKernelGGL(k_count_nonfinite<double>, g, b, 0, st, B.H64, nH, dc + 4);
        hipLaunchKernelGGL(k_count_nonfinite<float>, g, b, 0, st, B.H32, nH, dc + 5);
        hipLaunchKernelGGL(k_count_nonfinite<double>, g, b, 0, st, B.HT64, nH, dc + 6);
        hipLaunchKernelGGL(k_count_nonfinite<float>, g, b, 0, st, B.HT32, nH, dc + 7);
        hipLaunchKernelGGL(k_count_nonfinite<double>, g, b, 0, st, B.d64, nd, dc + 8);
        hipLaunchKernelGGL(k_count_nonfinite<float>, g, b, 0, st, B.d32, nd, dc + 9);
        hipLaunchKernelGGL(k_count_nonfinite<double>, g, b, 0, st, B.P64, nP, dc + 10);
        hipLaunchKernelGGL(k_count_nonfinite<uint16_t>, g, b, 0, st, B.P16, nP, dc + 11);
        hipLaunchKernelGGL(k_count_nonfinite<double>, g, b, 0, st, B.PT64, nP, dc + 12);
        hipLaunchKernelGGL(k_count_nonfinite<uint16_t>, g, b, 0, st, B.PT16, nP, dc + 13);
    }
    unsigned long long hc[14];
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(hc, dc, sizeof(hc), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    hipFree(dc);
    if (e != hipSuccess) return fail(ORBIT_ERR_HIP, std::string("opset_check: ") + hipGetErrorString(e));
    for (int i = 0; i < 14; ++i) counts[i] = (long long)hc[i];
    return ORBIT_OK;
}
