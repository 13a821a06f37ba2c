// Species coupling blocks on the device: sparse N x N matrices K applied as
// y += alpha * K x or y += alpha * K^T x on vectors in the boundary (k-major)
// layout.  They carry the chemistry-DAG and isorropia couplings between the
// species of a month: the forward RHS term e_s -= K_(s,src) c_src
// (reference orbit/core/orbit.py:263-307) and the adjoint receptor term
// R_src -= dtau K_(r,src)^T (lam_r - R_r) (orbit/modes/adjoint.py:650-676).
// K is stored as CSR and, for the transposed apply, as the CSR of K^T built
// at creation, so both applies are one thread per output row with no atomics
// (results are deterministic).  Production blocks are diagonal or nearly so
// (rate constants times a baseline state), which this row-per-thread form
// reads with coalesced loads.

#include "orbit_internal.h"

using namespace orbit;

struct orbit_spmat {
    orbit_ctx* ctx = nullptr;
    int n = 0;
    int nnz = 0;
    int32_t* rp[2] = {nullptr, nullptr};  // K, K^T
    int32_t* ci[2] = {nullptr, nullptr};
    double* va[2] = {nullptr, nullptr};
};

namespace {

__global__ void k_csr_axpy(int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                           const double* __restrict__ va, double alpha, const double* __restrict__ x,
                           double* __restrict__ y) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double acc = 0.0;
    for (int j = rp[r]; j < rp[r + 1]; ++j) acc += va[j] * x[ci[j]];
    y[r] += alpha * acc;
}

void release(orbit_spmat* K) {
    for (int t = 0; t < 2; ++t) {
        if (K->rp[t]) hipFree(K->rp[t]);
        if (K->ci[t]) hipFree(K->ci[t]);
        if (K->va[t]) hipFree(K->va[t]);
    }
    delete K;
}

}  // namespace

extern "C" int orbit_spmat_create(orbit_ctx* ctx, int n, const int32_t* rowptr, const int32_t* colidx,
                                  const double* vals, orbit_spmat** out) {
    if (!ctx || !rowptr || !out || n < 1) return fail(ORBIT_ERR_INVALID, "spmat: NULL argument or n < 1");
    *out = nullptr;
    const int nnz = rowptr[n];
    if (rowptr[0] != 0 || nnz < 0) return fail(ORBIT_ERR_INVALID, "spmat: row pointers must start at 0");
    if (nnz > 0 && (!colidx || !vals)) return fail(ORBIT_ERR_INVALID, "spmat: NULL column indices or values");
    for (int r = 0; r < n; ++r)
        if (rowptr[r + 1] < rowptr[r]) return fail(ORBIT_ERR_INVALID, "spmat: row pointers decrease");
    for (int j = 0; j < nnz; ++j)
        if (colidx[j] < 0 || colidx[j] >= n) return fail(ORBIT_ERR_INVALID, "spmat: column index out of range");
    // CSR of K^T by counting columns
    std::vector<int32_t> trp(n + 1, 0), tci(nnz);
    std::vector<double> tva(nnz);
    for (int j = 0; j < nnz; ++j) ++trp[colidx[j] + 1];
    for (int r = 0; r < n; ++r) trp[r + 1] += trp[r];
    {
        std::vector<int32_t> pos(trp.begin(), trp.end() - 1);
        for (int r = 0; r < n; ++r)
            for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
                const int p = pos[colidx[j]]++;
                tci[p] = r;
                tva[p] = vals[j];
            }
    }
    hipSetDevice(ctx->device);
    orbit_spmat* K = new orbit_spmat;
    K->ctx = ctx;
    K->n = n;
    K->nnz = nnz;
    const int32_t* hrp[2] = {rowptr, trp.data()};
    const int32_t* hci[2] = {colidx, tci.data()};
    const double* hva[2] = {vals, tva.data()};
    const size_t nz1 = std::max(nnz, 1);
    for (int t = 0; t < 2; ++t) {
        if (hipMalloc(&K->rp[t], (size_t)(n + 1) * 4) != hipSuccess ||
            hipMalloc(&K->ci[t], nz1 * 4) != hipSuccess || hipMalloc(&K->va[t], nz1 * 8) != hipSuccess) {
            release(K);
            return fail(ORBIT_ERR_NOMEM, "spmat: device allocation failed");
        }
        hipError_t e = hipMemcpy(K->rp[t], hrp[t], (size_t)(n + 1) * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess && nnz > 0) e = hipMemcpy(K->ci[t], hci[t], (size_t)nnz * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess && nnz > 0) e = hipMemcpy(K->va[t], hva[t], (size_t)nnz * 8, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            release(K);
            return fail(ORBIT_ERR_HIP, std::string("spmat: upload: ") + hipGetErrorString(e));
        }
    }
    *out = K;
    return ORBIT_OK;
}

extern "C" int orbit_spmat_apply_dev(orbit_ctx* ctx, orbit_spmat* K, int trans, double alpha,
                                     const double* x_dev, double* y_dev) {
    if (!ctx || !K || !x_dev || !y_dev) return fail(ORBIT_ERR_INVALID, "spmat apply: NULL argument");
    if (trans != 0 && trans != 1) return fail(ORBIT_ERR_INVALID, "spmat apply: trans must be 0 or 1");
    if (K->ctx->device != ctx->device) return fail(ORBIT_ERR_INVALID, "spmat apply: matrix is on another device");
    if (x_dev == y_dev) return fail(ORBIT_ERR_INVALID, "spmat apply: x and y must not alias");
    hipSetDevice(ctx->device);
    const int t = trans;
    hipLaunchKernelGGL(k_csr_axpy, dim3((K->n + 255) / 256), dim3(256), 0, ctx->stream, K->n, K->rp[t], K->ci[t],
                       K->va[t], alpha, x_dev, y_dev);
    ORBIT_HIP_CHECK(hipGetLastError());
    return ORBIT_OK;
}

extern "C" int orbit_spmat_destroy(orbit_spmat* K) {
    if (!K) return ORBIT_OK;
    hipSetDevice(K->ctx->device);
    hipStreamSynchronize(K->ctx->stream);
    release(K);
    return ORBIT_OK;
}
