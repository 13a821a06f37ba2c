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

// ------------------------------------------------------------------------- //
// ISORROPIA partitioning update (reference partitioning_per_bin,
// orbit/core/isorropia_closure.py:75-417): per cell and bin, 19 queries of the
// two partitioning fields around the end-of-bin state and the difference
// quotients built from them, in the reference's order of operations.
// ------------------------------------------------------------------------- //
constexpr int PART_FIELDS = 8;  // f_nh_eq, f_nh_marg, f_no3_eq, f_no3_marg, f_nh_dno3, f_no3_dnh, f_nh_dso4, f_no3_dso4

struct PartArgs {
    int len[7], off[7];
    long long stride[7];
    int n_ax;
    const double* axes;
    const float* tnh;   // f_nh4 table
    const float* tno3;  // f_no3 table
    long long ncell;
    int nbins;
    const double *so4, *nh, *no3;  // end-of-bin states; bin b at + b * sstride
    long long sstride;
    const double *T, *RH, *dust, *seasalt;  // met fields; bin b at + b * mstride (dust / seasalt nullable)
    long long mstride;
    int nout;      // 4, or 8 with the cross partials
    int blend;     // 1: out = alpha * new + (1 - alpha) * prev
    double alpha;
    const double* prev[PART_FIELDS];
    double* out[PART_FIELDS];
};

// f_nh4 and f_no3 at one point, clamped to [0, 1] (IsorropiaLUT.query), with the
// cell search, weights and corner order of k_lut_query.
__device__ void part_query(const double* sax, const PartArgs& A, double so4, double nh, double no3, double ca,
                           double na, double T, double RH, double& fnh, double& fno3) {
    const double x[7] = {so4, nh, no3, ca, na, T, RH};
    long long base = 0;
    double y[7], ym[7];
#pragma unroll
    for (int d = 0; d < 7; ++d) {
        const double* a = sax + A.off[d];
        const int i = cell_index(a, A.len[d], x[d]);
        y[d] = (x[d] - a[i]) / (a[i + 1] - a[i]);
        ym[d] = 1.0 - y[d];
        base += (long long)i * A.stride[d];
    }
    double acc_a = 0.0, acc_b = 0.0;
    for (int c = 0; c < 128; ++c) {
        long long idx = base;
        double w = 1.0;
#pragma unroll
        for (int d = 0; d < 7; ++d) {
            const bool up = (c >> (6 - d)) & 1;
            idx += up ? A.stride[d] : 0;
            w *= up ? y[d] : ym[d];
        }
        acc_a += (double)A.tnh[idx] * w;
        acc_b += (double)A.tno3[idx] * w;
    }
    fnh = acc_a < 0.0 ? 0.0 : (acc_a > 1.0 ? 1.0 : acc_a);
    fno3 = acc_b < 0.0 ? 0.0 : (acc_b > 1.0 ? 1.0 : acc_b);
}

// clip(0.5 * (up + down), 0, 1) of the one-sided differences of f*c, a side
// with no step giving f itself (isorropia_closure.py:163-176)
__device__ double part_marg(double fc, double fp, double fm, double c, double cp, double cm) {
#pragma clang fp contract(off)
    const double base = fc * c;
    const double dp = cp - c, dm = c - cm;
    const double up = dp > 0.0 ? (fp * cp - base) / fmax(dp, 1e-30) : fc;
    const double dn = dm > 0.0 ? (base - fm * cm) / fmax(dm, 1e-30) : fc;
    const double v = 0.5 * (up + dn);
    return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
}

__global__ __launch_bounds__(256) void k_partitioning(PartArgs A) {
#pragma clang fp contract(off)
    __shared__ double sax[LUT_MAXAX];
    for (int q = threadIdx.x; q < A.n_ax; q += blockDim.x) sax[q] = A.axes[q];
    __syncthreads();
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= A.ncell * A.nbins) return;
    const long long b = g / A.ncell, n = g - b * A.ncell;
    const long long si = b * A.sstride + n, mi = b * A.mstride + n;
    const double so4 = fmax(A.so4[si], 0.0), nh = fmax(A.nh[si], 0.0), no3 = fmax(A.no3[si], 0.0);
    const double ca = A.dust ? A.dust[mi] * 0.036 : 0.0;          // isorropia_closure.py:40, 130-133
    const double na = A.seasalt ? A.seasalt[mi] * 0.306 : 0.0;    // isorropia_closure.py:42, 134-137
    const double T = A.T[mi], RH = A.RH[mi] / 100.0;
    const double d = 0.01;                                        // isorropia_closure.py:48
    double r[PART_FIELDS];
    double f_nh, f_no3;
    part_query(sax, A, so4, nh, no3, ca, na, T, RH, f_nh, f_no3);
    const double nh_p = nh * (1.0 + d) + 1e-12, nh_m = fmax(nh * (1.0 - d), 0.0);
    const double no3_p = no3 * (1.0 + d) + 1e-12, no3_m = fmax(no3 * (1.0 - d), 0.0);
    double nh_at_nhp, no3_at_nhp, nh_at_nhm, no3_at_nhm, nh_at_no3p, no3_at_no3p, nh_at_no3m, no3_at_no3m;
    part_query(sax, A, so4, nh_p, no3, ca, na, T, RH, nh_at_nhp, no3_at_nhp);
    part_query(sax, A, so4, nh_m, no3, ca, na, T, RH, nh_at_nhm, no3_at_nhm);
    part_query(sax, A, so4, nh, no3_p, ca, na, T, RH, nh_at_no3p, no3_at_no3p);
    part_query(sax, A, so4, nh, no3_m, ca, na, T, RH, nh_at_no3m, no3_at_no3m);
    r[0] = f_nh;
    r[1] = part_marg(f_nh, nh_at_nhp, nh_at_nhm, nh, nh_p, nh_m);
    r[2] = f_no3;
    r[3] = part_marg(f_no3, no3_at_no3p, no3_at_no3m, no3, no3_p, no3_m);
    if (A.nout > 4) {
        const double so4_p = so4 * (1.0 + d) + 1e-12, so4_m = fmax(so4 * (1.0 - d), 0.0);
        double nh_pp, no3_pp, nh_pm, no3_pm, nh_mp, no3_mp, nh_mm, no3_mm;  // NH-NO3 plane
        part_query(sax, A, so4, nh_p, no3_p, ca, na, T, RH, nh_pp, no3_pp);
        part_query(sax, A, so4, nh_p, no3_m, ca, na, T, RH, nh_pm, no3_pm);
        part_query(sax, A, so4, nh_m, no3_p, ca, na, T, RH, nh_mp, no3_mp);
        part_query(sax, A, so4, nh_m, no3_m, ca, na, T, RH, nh_mm, no3_mm);
        const double mnh_no3p = part_marg(nh_at_no3p, nh_pp, nh_mp, nh, nh_p, nh_m);
        const double mnh_no3m = part_marg(nh_at_no3m, nh_pm, nh_mm, nh, nh_p, nh_m);
        r[4] = nh * ((mnh_no3p - mnh_no3m) / fmax(no3_p - no3_m, 1e-30));
        const double mno3_nhp = part_marg(no3_at_nhp, no3_pp, no3_pm, no3, no3_p, no3_m);
        const double mno3_nhm = part_marg(no3_at_nhm, no3_mp, no3_mm, no3, no3_p, no3_m);
        r[5] = no3 * ((mno3_nhp - mno3_nhm) / fmax(nh_p - nh_m, 1e-30));
        double nh_pp_s, nh_pm_s, nh_mp_s, nh_mm_s, no3_pp_s, no3_pm_s, no3_mp_s, no3_mm_s, unused;
        part_query(sax, A, so4_p, nh_p, no3, ca, na, T, RH, nh_pp_s, unused);
        part_query(sax, A, so4_m, nh_p, no3, ca, na, T, RH, nh_pm_s, unused);
        part_query(sax, A, so4_p, nh_m, no3, ca, na, T, RH, nh_mp_s, unused);
        part_query(sax, A, so4_m, nh_m, no3, ca, na, T, RH, nh_mm_s, unused);
        part_query(sax, A, so4_p, nh, no3_p, ca, na, T, RH, unused, no3_pp_s);
        part_query(sax, A, so4_m, nh, no3_p, ca, na, T, RH, unused, no3_pm_s);
        part_query(sax, A, so4_p, nh, no3_m, ca, na, T, RH, unused, no3_mp_s);
        part_query(sax, A, so4_m, nh, no3_m, ca, na, T, RH, unused, no3_mm_s);
        double nh_c_sp, no3_c_sp, nh_c_sm, no3_c_sm;
        part_query(sax, A, so4_p, nh, no3, ca, na, T, RH, nh_c_sp, no3_c_sp);
        part_query(sax, A, so4_m, nh, no3, ca, na, T, RH, nh_c_sm, no3_c_sm);
        const double dso4 = fmax(so4_p - so4_m, 1e-30);
        const double mnh_sp = part_marg(nh_c_sp, nh_pp_s, nh_mp_s, nh, nh_p, nh_m);
        const double mnh_sm = part_marg(nh_c_sm, nh_pm_s, nh_mm_s, nh, nh_p, nh_m);
        r[6] = nh * ((mnh_sp - mnh_sm) / dso4);
        const double mno3_sp = part_marg(no3_c_sp, no3_pp_s, no3_mp_s, no3, no3_p, no3_m);
        const double mno3_sm = part_marg(no3_c_sm, no3_pm_s, no3_mm_s, no3, no3_p, no3_m);
        r[7] = no3 * ((mno3_sp - mno3_sm) / dso4);
    }
    const long long oi = b * A.ncell + n;
    for (int f = 0; f < A.nout; ++f) {
        const double v = A.blend ? A.alpha * r[f] + (1.0 - A.alpha) * A.prev[f][oi] : r[f];
        A.out[f][oi] = v;
    }
}

}  // namespace

extern "C" int orbit_lut_partitioning_dev(orbit_ctx* ctx, orbit_lut* L, int field_nh4, int field_no3, int n_bins,
                                          long long ncell, const double* const* state_dev, long long state_stride,
                                          const double* const* met_dev, long long met_stride, int cross,
                                          double alpha, const double* const* prev_dev, double* const* out_dev) {
    if (!ctx || !L || !state_dev || !met_dev || !out_dev)
        return fail(ORBIT_ERR_INVALID, "partitioning: NULL argument");
    if (L->nd != 7) return fail(ORBIT_ERR_INVALID, "partitioning: the table must have the 7 ISORROPIA axes");
    if (field_nh4 < 0 || field_nh4 >= L->nf || field_no3 < 0 || field_no3 >= L->nf)
        return fail(ORBIT_ERR_INVALID, "partitioning: field index out of range");
    if (n_bins < 1 || ncell < 0) return fail(ORBIT_ERR_INVALID, "partitioning: n_bins >= 1 and ncell >= 0 needed");
    if (state_stride < ncell || met_stride < ncell)
        return fail(ORBIT_ERR_INVALID, "partitioning: bin strides must be >= ncell");
    if (ncell == 0) return ORBIT_OK;
    PartArgs A{};
    A.n_ax = 0;
    for (int d = 0; d < 7; ++d) {
        A.len[d] = L->len[d];
        A.off[d] = L->off[d];
        A.stride[d] = L->stride[d];
        A.n_ax += L->len[d];
    }
    A.axes = L->axes;
    A.tnh = L->table[field_nh4];
    A.tno3 = L->table[field_no3];
    A.ncell = ncell;
    A.nbins = n_bins;
    A.so4 = state_dev[0];
    A.nh = state_dev[1];
    A.no3 = state_dev[2];
    A.sstride = state_stride;
    A.T = met_dev[0];
    A.RH = met_dev[1];
    A.dust = met_dev[2];
    A.seasalt = met_dev[3];
    A.mstride = met_stride;
    if (!A.so4 || !A.nh || !A.no3 || !A.T || !A.RH)
        return fail(ORBIT_ERR_INVALID, "partitioning: the states and T, RH are required");
    A.nout = cross ? 8 : 4;
    A.blend = prev_dev && alpha > 0.0 && alpha < 1.0;
    A.alpha = alpha;
    for (int f = 0; f < A.nout; ++f) {
        if (!out_dev[f]) return fail(ORBIT_ERR_INVALID, "partitioning: NULL output field");
        A.out[f] = out_dev[f];
        if (A.blend) {
            if (!prev_dev[f]) return fail(ORBIT_ERR_INVALID, "partitioning: NULL previous field");
            A.prev[f] = prev_dev[f];
        }
    }
    hipSetDevice(ctx->device);
    const long long total = ncell * n_bins;
    hipLaunchKernelGGL(k_partitioning, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream, A);
    ORBIT_HIP_CHECK(hipGetLastError());
    return ORBIT_OK;
}

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
