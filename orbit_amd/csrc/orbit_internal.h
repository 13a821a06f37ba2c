// Internal declarations shared by the HIP translation units of liborbit_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "orbit_hip.h"

namespace orbit {

// Column blocks are padded to 16 levels: a column is 16 contiguous values in the
// device layout idx = col*16 + k (k-major vectors at the boundary are permuted).
constexpr int NZP = 16;
constexpr int MAXM = 64;  // largest GMRES restart supported

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define ORBIT_HIP_CHECK(expr)                                                         \
    do {                                                                              \
        hipError_t _e = (expr);                                                       \
        if (_e != hipSuccess)                                                         \
            return ::orbit::fail(ORBIT_ERR_HIP, std::string(#expr) + ": " +          \
                                                    hipGetErrorString(_e));           \
    } while (0)

// bf16 storage as raw 16-bit words: round-to-nearest-even on the way in.
__host__ __device__ inline uint16_t f32_to_bf16(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf/nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint16_t h) {
    uint32_t u = ((uint32_t)h) << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

// Per-launch kernel timing of the half-sweeps (a diagnostic): while `on`,
// every half-sweep launch on the context's stream records its own start and
// stop events (hipExtLaunchKernelGGL) from a pool made beforehand, with its
// variant and algorithmic bytes; solves do not replay captured graphs then.
// Timing variant offset of the first sweep from x = 0 with both colours in one launch.
constexpr int FUSED_FIRST_VARIANT = 64;
// ... and of that launch carrying the previous bin's last black half-sweep too.
constexpr int HANDOFF_VARIANT = 128;
struct SweepTiming {
    struct Rec {
        int variant;   // prec * 16 + mfma * 8 + cs * 4 + own_zero * 2 + nbr_zero (+ FUSED_FIRST_VARIANT or HANDOFF_VARIANT)
        double bytes;  // algorithmic bytes of the launch
    };
    std::vector<hipEvent_t> ev;  // 2 per launch
    size_t next = 0;
    long long dropped = 0;  // launches past the pool (not timed)
    std::vector<Rec> recs;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> owned;  // scratch owned by the context
    // reusable scratch, grown on demand: 0-6 and 10-12 the periodic solve, 7
    // calibration and benchmarks, 8 the forward FCT driver's emissions, c0 and
    // s~, 9 the tables of a group solve
    DevBuf scratch[13];
    // this device's copy of the fixed scratch array of the allocation-free
    // solve (its device address, looked up at context creation)
    void* fixed_scratch = nullptr;
    SweepTiming* timing = nullptr;  // non-null while kernel timing is on
    // pinned host staging of the solves' small device-to-host reads (residuals,
    // norms): a direct copy instead of one staged through a bounce buffer
    double* pinned = nullptr;
    size_t pinned_n = 0;
    int ensure(int slot, size_t bytes);
    int ensure_pinned(size_t n);
    template <class T>
    T* buf(int slot) { return reinterpret_cast<T*>(scratch[slot].p); }
};

// Per-bin device operator data.
struct BinOps {
    // shared (per group) column blocks (identity + dtau*offdiag of the column; the
    // diagonal lives in d) and 4-point horizontal stencil (W,E,S,N per level)
    double* V64 = nullptr;   float* V32 = nullptr;
    double* VT64 = nullptr;  float* VT32 = nullptr;
    double* H64 = nullptr;   float* H32 = nullptr;
    double* HT64 = nullptr;  float* HT32 = nullptr;
    // per tracer: dtau * diagonal of L, padded layout
    double* d64 = nullptr;   float* d32 = nullptr;
    // per tracer: inverse column blocks
    double* P64 = nullptr;   uint16_t* P16 = nullptr;
    double* PT64 = nullptr;  uint16_t* PT16 = nullptr;
    // shared transport diagonal (dtau*T_nn) for the mass-balance check
    double* tdiag = nullptr;
};

// West / east neighbour column of column `col` at x index i: the adjacent
// column, the column across the wrap face when x is periodic (px: couplings
// between i = 0 and i = nx - 1), or -1 at an open edge.  With nx = 2
// both neighbours of a column are the other column.
__host__ __device__ inline int west_col(int col, int i, int nx, int px) {
    return i > 0 ? col - 1 : (px ? col + nx - 1 : -1);
}
__host__ __device__ inline int east_col(int col, int i, int nx, int px) {
    return i < nx - 1 ? col + 1 : (px ? col - nx + 1 : -1);
}

// The first red-black sweep from x = 0 in one launch over column patches
// (k_first_sweep).  A workgroup owns the PA x PB columns i0 .. i0+PA-1,
// j0 .. j0+PB-1 of its patch (PA, PB even, so i0 + j0 is even; columns past
// the grid's edge belong to no patch) and first computes the new values of
// its red columns and of the red columns just outside it that its black
// columns touch (the ring reds), each exactly as the red half-sweep does, then
// the new values of its black columns from those.  Positions in the
// (PA+2) x (PB+2) ext grid: own column (a, c) at (a+1, c+1); the ring west /
// east of row c at (0, c+1) / (PA+1, c+1), south / north of column a at
// (a+1, 0) / (a+1, PB+1).  The east ring is the column after the patch's last
// one in the grid (i = min(i0+PA, nx), across the wrap face when x is
// periodic), so a patch cut by the grid's edge still finds its east
// neighbours.  Red items: PA*PB/2 own reds, then PB/2 west, PB/2 east, PA/2
// south and PA/2 north ring reds; black items: PA*PB/2 own blacks.
struct PatchCol {
    int col;     // column id j * nx + i (a valid column of the patch's own if !valid)
    int i, j;
    int pos;     // position in the ext grid
    bool valid;  // a column of the grid
};
template <int PA, int PB>
__host__ __device__ inline PatchCol patch_red(int t, int i0, int j0, int nx, int ny, int px) {
    constexpr int HA = PA / 2, HB = PB / 2, EW = PA + 2;
    PatchCol r;
    int i, j, a, c;
    bool ok = true;
    if (t < HA * PB) {  // own red: row c holds the reds a = (c & 1), (c & 1) + 2, ...
        c = t / HA;
        a = (c & 1) + 2 * (t - c * HA);
        i = i0 + a;
        j = j0 + c;
        ok = i < nx && j < ny;
        r.pos = (a + 1) + (c + 1) * EW;
    } else if ((t -= HA * PB) < HB) {  // west ring, odd rows
        c = 2 * t + 1;
        i = i0 - 1;
        j = j0 + c;
        if (i < 0) {
            if (px) i += nx; else ok = false;
        }
        ok = ok && j < ny;
        r.pos = (c + 1) * EW;
    } else if ((t -= HB) < HB) {  // east ring, even rows
        c = 2 * t;
        i = i0 + PA < nx ? i0 + PA : nx;
        j = j0 + c;
        if (i >= nx) {
            if (px) i -= nx; else ok = false;
        }
        ok = ok && j < ny;
        r.pos = (PA + 1) + (c + 1) * EW;
    } else if ((t -= HB) < HA) {  // south ring, odd columns
        a = 2 * t + 1;
        i = i0 + a;
        j = j0 - 1;
        ok = i < nx && j >= 0;
        r.pos = a + 1;
    } else {  // north ring, even columns
        t -= HA;
        a = 2 * t;
        i = i0 + a;
        j = j0 + PB;
        ok = i < nx && j < ny;
        r.pos = (a + 1) + (PB + 1) * EW;
    }
    r.valid = ok;
    r.i = ok ? i : i0;
    r.j = ok ? j : j0;
    r.col = r.j * nx + r.i;
    return r;
}
template <int PA, int PB>
__host__ __device__ inline PatchCol patch_black(int t, int i0, int j0, int nx, int ny) {
    constexpr int HA = PA / 2, EW = PA + 2;
    const int c = t / HA, a = 1 - (c & 1) + 2 * (t - c * HA);
    PatchCol r;
    r.valid = i0 + a < nx && j0 + c < ny;
    r.i = r.valid ? i0 + a : i0;
    r.j = r.valid ? j0 + c : j0;
    r.col = r.j * nx + r.i;
    r.pos = (a + 1) + (c + 1) * EW;
    return r;
}
// Ext-grid positions of the four neighbours of an own black column at pos
// (grid column i): west, east (the east ring when i + 1 is past the patch's
// last grid column), south, north.
template <int PA>
__host__ __device__ inline void patch_black_nbrs(int pos, int i, int i0, int nx, int* q) {
    constexpr int EW = PA + 2;
    const int iend = i0 + PA < nx ? i0 + PA : nx;
    q[0] = pos - 1;
    q[1] = i + 1 == iend ? (PA + 1) + (pos / EW) * EW : pos + 1;
    q[2] = pos - EW;
    q[3] = pos + EW;
}

// Launch shape of the half-sweep kernel: slots per wave (1 or 4), threads per
// workgroup (64, 128, 256) and block order (XCD-contiguous or linear).  Chosen
// per sweep precision and slot-count bucket; every shape computes the same
// per-lane arithmetic, so results do not depend on it.
struct SweepShape {
    int8_t cs = 0;
    int16_t threads = 256;
    int8_t xcd = 1;
};
constexpr int N_SWEEP_PRECS = 3;   // FP64, accurate mixed (FP64 + BF16 blocks), FP32 inner
constexpr int N_SLOT_BUCKETS = 5;  // 1, 2-3, 4-7, 8-15, 16+
inline int slot_bucket(int nslots) {
    return nslots <= 1 ? 0 : nslots < 4 ? 1 : nslots < 8 ? 2 : nslots < 16 ? 3 : 4;
}

struct GraphKey {
    int dir, prec, sweeps, nslots;
    std::vector<int> tracers;
    const void* in;
    const void* out;
    bool operator<(const GraphKey& o) const {
        return std::tie(dir, prec, sweeps, nslots, tracers, in, out) <
               std::tie(o.dir, o.prec, o.sweeps, o.nslots, o.tracers, o.in, o.out);
    }
};

struct Opset {
    Ctx* ctx = nullptr;
    int nz = 0, ny = 0, nx = 0, nbins = 0, ntr = 0, ngroups = 1;
    int ncol = 0, n16 = 0, N = 0;
    int px = 0;  // periodic x: the operators couple i = 0 and i = nx - 1
    double dtau = 0.0;
    bool has_tdiag = false;
    int def_inner = 6, def_acc = 14, def_check = 6;
    double rho = 0.1;  // calibrated contraction of one red-black sweep (with margin)
    double omega = 1.0;  // relaxation of the sweeps' correction (chosen by the calibration)
    bool mfma = false;   // block product of the BF16-inverse sweeps on the matrix cores (opt-in)
    int patch = 2;       // first FP32 sweep of an inner solve in one launch over column patches of this shape (0: two half-sweeps)
    bool handoff = true; // a bin's last black FP32 half-sweep runs in the next bin's first-sweep launch (needs patch > 0)
    std::vector<BinOps> bins;
    // Column colours of the sweeps: red / black ((i + j) even / odd) over the
    // first nxc columns of every row; with a periodic x and an odd nx the wrap
    // face would join two columns of one colour, so the last column of each
    // row (i = nx - 1) is left out of the red-black pattern (nxc = nx - 1) and
    // takes colours 2 (j even) and 3 (j odd), neither of which touches itself.
    int ncolours = 2;
    int nxc = 0;
    int ncols_c[4] = {0, 0, 0, 0};
    std::vector<void*> owned;
    double bytes = 0.0;
    SweepShape shapes[N_SWEEP_PRECS][N_SLOT_BUCKETS];  // set at creation, by autotune or by the caller
    std::map<GraphKey, hipGraphExec_t> graphs;
    std::mutex graph_mu;  // contexts on several host threads share an opset
    template <class T>
    T* alloc(size_t n);
};

}  // namespace orbit

struct orbit_ctx : orbit::Ctx {};
struct orbit_opset : orbit::Opset {};

namespace orbit {
// Opset construction steps shared by the CSR and the grid-field builders:
// allocate every bin's arrays; derive the FP32 copies, transposes and factored
// column blocks of a bin whose V64, H64 and d64 are filled; then the colour
// counts, default launch shapes and sweep calibration.
int alloc_bins(orbit_opset* op);
int derive_bin(orbit_ctx* ctx, orbit_opset* op, int b);
int finalize_opset(orbit_ctx* ctx, orbit_opset* op);

// Pass-to-pass state of forward solves whose forcing changes a little from one
// solve to the next (the FCT deferred-correction passes), mixed precision.  The
// first solve (no previous orbit) leaves s~ (the FP32 forced sweep of its
// forcing) in s_state.  A later solve, given the previous orbit as its warm
// whole orbit, starts its states there: its first residual is that of the
// previous orbit under the new forcing, and s~ = s~_prev + that residual's
// periodic part (returned in s_state).
struct PassDelta {
    double* s_state = nullptr;  // [nt][n16] device layout, FP64: in s~ of the previous pass, out s~ of this one
};
int solve_forward_pass(orbit_ctx* ctx, orbit_opset* op, const int* tracers, int nt, const double* e_dev,
                       const double* prev_orbit_dev, PassDelta* pd, const orbit_opts* opts, double* out_dev,
                       orbit_info* info);
}  // namespace orbit
