// Regular-grid multilinear lookup table on the device: the 7-D ISORROPIA
// partitioning table of the closure (reference IsorropiaLUT,
// orbit/core/isorropia_lut.py:12-147), queried at every cell and bin of the
// orbit state.
//
// The reference evaluates scipy's RegularGridInterpolator(method="linear",
// bounds_error=False, fill_value=None): per axis the cell index is
// searchsorted(axis, x) - 1 clipped to [0, n-2] and the weight
// y = (x - a[i]) / (a[i+1] - a[i]) is not clipped, so points outside the table
// are extrapolated linearly; the value is the sum over the 2^d cell corners of
// table[corner] * prod(y or 1 - y).  Outputs are then clamped per field
// (f_nh4, f_no3 to [0, 1], aerosol water to >= 0).
//
// One thread per point: the axes (a few hundred values in all) are staged in
// LDS, the cell is found by binary search on each axis, and the 2^d corners are
// gathered from the FP32 table in HBM (the last axis is contiguous, so corner
// pairs share a cache line).  Sums are FP64 in a fixed corner order.

#include <cmath>

#include "orbit_internal.h"

using namespace orbit;

constexpr int LUT_MAXD = 8;
constexpr int LUT_MAXAX = 1024;  // total axis values staged in LDS
constexpr int LUT_MAXF = 8;

struct orbit_lut {
    orbit_ctx* ctx = nullptr;
    int nd = 0, nf = 0;
    int len[LUT_MAXD] = {0};
    long long stride[LUT_MAXD] = {0};
    int off[LUT_MAXD] = {0};
    double* axes = nullptr;     // concatenated axis values
    float* table[LUT_MAXF] = {nullptr};
    size_t n_entries = 0;
};

namespace {

struct LutArgs {
    int nd, nf;
    int len[LUT_MAXD];
    int off[LUT_MAXD];
    long long stride[LUT_MAXD];
    int n_ax;
    const double* axes;
    const float* table[LUT_MAXF];
    double lo[LUT_MAXF], hi[LUT_MAXF];
    long long npts;
    const double* x[LUT_MAXD];
    double* out[LUT_MAXF];
};

// searchsorted(a, x, side="left") - 1, clipped to [0, n-2]
__device__ __forceinline__ int cell_index(const double* a, int n, double x) {
    int lo = 0, hi = n;  // first index with a[idx] >= x
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    int i = lo - 1;
    i = i < 0 ? 0 : i;
    return i > n - 2 ? n - 2 : i;
}

template <int ND>
__global__ __launch_bounds__(256) void k_lut_query(LutArgs A) {
    __shared__ double sax[LUT_MAXAX];
    for (int q = threadIdx.x; q < A.n_ax; q += blockDim.x) sax[q] = A.axes[q];
    __syncthreads();
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.npts) return;
    long long base = 0;
    double y[ND], ym[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        const double* a = sax + A.off[d];
        const double x = A.x[d][p];
        const int i = cell_index(a, A.len[d], x);
        y[d] = (x - a[i]) / (a[i + 1] - a[i]);
        ym[d] = 1.0 - y[d];
        base += (long long)i * A.stride[d];
    }
    for (int f = 0; f < A.nf; ++f) {
        const float* T = A.table[f];
        double acc = 0.0;
        // corners in the order of the first axis varying slowest
        for (int c = 0; c < (1 << ND); ++c) {
            long long idx = base;
            double w = 1.0;
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                const bool up = (c >> (ND - 1 - d)) & 1;
                idx += up ? A.stride[d] : 0;
                w *= up ? y[d] : ym[d];
            }
            acc += (double)T[idx] * w;
        }
        acc = acc < A.lo[f] ? A.lo[f] : acc;  // NaN passes both tests unchanged
        acc = acc > A.hi[f] ? A.hi[f] : acc;
        A.out[f][p] = acc;
    }
}

}  // namespace

extern "C" int orbit_lut_create(orbit_ctx* ctx, int ndim, const int* axis_len, const double* const* axes,
                                int nfields, const float* const* fields, orbit_lut** out) {
    if (!ctx || !axis_len || !axes || !fields || !out) return fail(ORBIT_ERR_INVALID, "lut_create: NULL argument");
    *out = nullptr;
    if (ndim < 1 || ndim > LUT_MAXD) return fail(ORBIT_ERR_INVALID, "lut_create: ndim must be in [1, 8]");
    if (nfields < 1 || nfields > LUT_MAXF) return fail(ORBIT_ERR_INVALID, "lut_create: nfields must be in [1, 8]");
    int total = 0;
    size_t n = 1;
    for (int d = 0; d < ndim; ++d) {
        if (axis_len[d] < 2 || !axes[d]) return fail(ORBIT_ERR_INVALID, "lut_create: every axis needs >= 2 values");
        for (int q = 0; q + 1 < axis_len[d]; ++q)
            if (!(axes[d][q] < axes[d][q + 1]))
                return fail(ORBIT_ERR_INVALID, "lut_create: axis values must be strictly ascending");
        total += axis_len[d];
        n *= (size_t)axis_len[d];
    }
    if (total > LUT_MAXAX) return fail(ORBIT_ERR_INVALID, "lut_create: more than 1024 axis values in all");
    for (int f = 0; f < nfields; ++f)
        if (!fields[f]) return fail(ORBIT_ERR_INVALID, "lut_create: NULL field");
    hipSetDevice(ctx->device);
    auto* L = new orbit_lut();
    L->ctx = ctx;
    L->nd = ndim;
    L->nf = nfields;
    L->n_entries = n;
    long long s = 1;
    for (int d = ndim - 1; d >= 0; --d) {  // C order: the last axis is contiguous
        L->len[d] = axis_len[d];
        L->stride[d] = s;
        s *= axis_len[d];
    }
    std::vector<double> ax;
    for (int d = 0; d < ndim; ++d) {
        L->off[d] = (int)ax.size();
        ax.insert(ax.end(), axes[d], axes[d] + axis_len[d]);
    }
    hipStream_t st = ctx->stream;
    bool ok = hipMalloc(&L->axes, ax.size() * 8) == hipSuccess;
    for (int f = 0; f < nfields && ok; ++f) ok = hipMalloc(&L->table[f], n * 4) == hipSuccess;
    if (!ok) {
        orbit_lut_destroy(L);
        return fail(ORBIT_ERR_NOMEM, "lut_create: device allocation failed");
    }
    hipError_t e = hipMemcpyAsync(L->axes, ax.data(), ax.size() * 8, hipMemcpyHostToDevice, st);
    for (int f = 0; f < nfields && e == hipSuccess; ++f)
        e = hipMemcpyAsync(L->table[f], fields[f], n * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);  // `ax` and the host fields may go now
    if (e != hipSuccess) {
        orbit_lut_destroy(L);
        return fail(ORBIT_ERR_HIP, std::string("lut_create: ") + hipGetErrorString(e));
    }
    *out = L;
    return ORBIT_OK;
}

extern "C" int orbit_lut_query_dev(orbit_ctx* ctx, orbit_lut* L, long long npts, const double* const* coords_dev,
                                   const double* clamp_lo, const double* clamp_hi, double* const* out_dev) {
    if (!ctx || !L || !coords_dev || !out_dev) return fail(ORBIT_ERR_INVALID, "lut_query: NULL argument");
    if (npts < 0) return fail(ORBIT_ERR_INVALID, "lut_query: npts < 0");
    if (npts == 0) return ORBIT_OK;
    LutArgs A{};
    A.nd = L->nd;
    A.nf = L->nf;
    A.n_ax = 0;
    for (int d = 0; d < L->nd; ++d) {
        if (!coords_dev[d]) return fail(ORBIT_ERR_INVALID, "lut_query: NULL coordinate array");
        A.len[d] = L->len[d];
        A.off[d] = L->off[d];
        A.stride[d] = L->stride[d];
        A.x[d] = coords_dev[d];
        A.n_ax += L->len[d];
    }
    for (int f = 0; f < L->nf; ++f) {
        if (!out_dev[f]) return fail(ORBIT_ERR_INVALID, "lut_query: NULL output array");
        A.table[f] = L->table[f];
        A.out[f] = out_dev[f];
        A.lo[f] = clamp_lo ? clamp_lo[f] : -HUGE_VAL;
        A.hi[f] = clamp_hi ? clamp_hi[f] : HUGE_VAL;
    }
    A.axes = L->axes;
    A.npts = npts;
    hipSetDevice(ctx->device);
    const dim3 g((unsigned)((npts + 255) / 256)), b(256);
    switch (L->nd) {
        case 1: hipLaunchKernelGGL(k_lut_query<1>, g, b, 0, ctx->stream, A); break;
        case 2: hipLaunchKernelGGL(k_lut_query<2>, g, b, 0, ctx->stream, A); break;
        case 3: hipLaunchKernelGGL(k_lut_query<3>, g, b, 0, ctx->stream, A); break;
        case 4: hipLaunchKernelGGL(k_lut_query<4>, g, b, 0, ctx->stream, A); break;
        case 5: hipLaunchKernelGGL(k_lut_query<5>, g, b, 0, ctx->stream, A); break;
        case 6: hipLaunchKernelGGL(k_lut_query<6>, g, b, 0, ctx->stream, A); break;
        case 7: hipLaunchKernelGGL(k_lut_query<7>, g, b, 0, ctx->stream, A); break;
        default: hipLaunchKernelGGL(k_lut_query<8>, g, b, 0, ctx->stream, A); break;
    }
    ORBIT_HIP_CHECK(hipGetLastError());
    return ORBIT_OK;
}

extern "C" int orbit_lut_destroy(orbit_lut* L) {
    if (!L) return ORBIT_OK;
    if (L->ctx) {
        hipSetDevice(L->ctx->device);
        hipStreamSynchronize(L->ctx->stream);
    }
    if (L->axes) hipFree(L->axes);
    for (int f = 0; f < LUT_MAXF; ++f)
        if (L->table[f]) hipFree(L->table[f]);
    delete L;
    return ORBIT_OK;
}
