// Horizontal Zalesak FCT on the device: the anti-diffusive deferred-correction
// source d_AD = rho * (-div(C*A)) and its frozen-coefficient tangent-linear
// operator L_AD.  Restates SumilThakr/orbit orbit/core/fct.py:235-572:
//   van Leer phi with guarded ratio        fct.py:235-247
//   interior face fluxes A, beta, tr, 1/dx  fct.py:256-357
//   Zalesak bounds, P+-, Q+-, R+-, C        fct.py:381-447
//   divergence into cells                   fct.py:360-378
//   L_AD gains w = C*beta/dx                fct.py:495-572
// One thread per cell or face; FP64.  x = c/rho is formed once per cell; face
// quantities are recomputed where they are needed (each needs x at most two
// cells away), so the only other intermediates in HBM are R+ and R- per cell.
#include <algorithm>
#include <cmath>
#include <vector>

#include "orbit_internal.h"

namespace orbit {

constexpr double kGrav = 9.80665;
constexpr double kEps = 1e-30;

struct FctGrid {
    int nz, ny, nx, split, conc;
    const double* up;  // UAvg or UAvg_plus
    const double* um;  // UAvg_minus (split only)
    const double* vp;
    const double* vm;
    const double* dx;
    double dy;
    const double* dP;
    const double* Dz;
};

__device__ __forceinline__ double rho_at(const FctGrid& g, size_t n) {
    if (!g.conc) return 1.0;
    const double p = g.dP[n], z = g.Dz[n];
    return (p > 0.0 && z > 0.0) ? p / (kGrav * z) : 1.0;
}

// van Leer limiter phi(r) = (r + |r|) / (1 + |r|) of the ratio r = num / den of
// successive differences (with the reference's guard: r = 1 or 1e6 when
// |den| < 1e-30).  With one division: phi = 0 for r <= 0 and 2r / (1 + r) =
// 2 num / (num + den) for r > 0 (num and den of one sign, so the sum does not
// cancel); the same value to round-off.
__device__ __forceinline__ double vl_phi(double up, double far, double down) {
    const double num = up - far, den = down - up;
    if (fabs(den) < kEps) return fabs(num) < kEps ? 1.0 : 2.0e6 / (1.0 + 1.0e6);
    return ((num > 0.0 && den > 0.0) || (num < 0.0 && den < 0.0)) ? 2.0 * num / (num + den) : 0.0;
}

struct Face {
    double A;      // anti-diffusive flux (x units * speed)
    double beta;   // frozen reconstruction gain
    double inv;    // 1/dx or 1/dy
    double tr_lo, tr_hi;
    bool valid;
};

// x = c / rho for nvec vectors (the transported variable), once per cell: the
// face fluxes read x at up to two cells away, so forming it once saves the
// stencil's repeated loads of dP and Dz and its divisions.
__global__ void k_fct_x(FctGrid g, const double* __restrict__ c, long long cs, double* __restrict__ x,
                        long long xs, int N) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    x[(size_t)blockIdx.y * xs + n] = c[(size_t)blockIdx.y * cs + n] / rho_at(g, n);
}

// The tracer-independent part of a face: its terrain weights (fct.py:330-345),
// whether both cells have air, and 1 / its width.
struct FaceGeo {
    double tr_lo, tr_hi, inv;
    bool valid;
};

// x face between (k,j,i) [lo] and (k,j,i+1) [hi]
__device__ __forceinline__ FaceGeo face_geo_x(const FctGrid& g, int k, int j, int i) {
    const size_t row = ((size_t)k * g.ny + j) * g.nx;
    FaceGeo f;
    const double dl = g.dP[row + i], dh = g.dP[row + i + 1];
    const double df = 0.5 * (dl + dh);
    f.valid = dl > 0.0 && dh > 0.0;
    f.tr_lo = f.valid ? df / dl : 1.0;
    f.tr_hi = f.valid ? df / dh : 1.0;
    const double dxf = g.dx[j];
    f.inv = 1.0 / (dxf > 0.0 ? dxf : 1.0);
    return f;
}

// y face between (k,j,i) [lo = south] and (k,j+1,i) [hi = north]
__device__ __forceinline__ FaceGeo face_geo_y(const FctGrid& g, int k, int j, int i) {
    const size_t nl = ((size_t)k * g.ny + j) * g.nx + i, nh = nl + g.nx;
    FaceGeo f;
    const double dl = g.dP[nl], dh = g.dP[nh];
    const double df = 0.5 * (dl + dh);
    f.valid = dl > 0.0 && dh > 0.0;
    const double dxl = g.dx[j], dxh = g.dx[j + 1], dxf = 0.5 * (dxl + dxh);
    const double wl = (dxf / (dxl > 0.0 ? dxl : 1.0)) * (df / (dl > 0.0 ? dl : 1.0));
    const double wh = (dxf / (dxh > 0.0 ? dxh : 1.0)) * (df / (dh > 0.0 ? dh : 1.0));
    f.tr_lo = f.valid ? wl : 1.0;
    f.tr_hi = f.valid ? wh : 1.0;
    f.inv = g.dy > 0.0 ? 1.0 / g.dy : 1.0;
    return f;
}

// The flux part of the x face (lo = (k,j,i)): van Leer reconstruction gain
// beta and anti-diffusive flux A.  xat(j, i): x of an in-domain cell of level
// k; gat(j, i): x of the ghost there (read only when has_g).  The per-cell
// kernels and the LDS-tiled one share this code, so a face's arithmetic is the
// same expression whichever kernel evaluates it.
template <class XA, class GA>
__device__ __forceinline__ void face_flux_x(const FctGrid& g, XA xat, GA gat, bool has_g, int k, int j, int i,
                                            double& beta, double& A) {
    const int nx = g.nx;
    const double xl = xat(j, i), xh = xat(j, i + 1);
    double xll, xhh;
    if (i > 0) xll = xat(j, i - 1);
    else xll = has_g ? gat(j, 0) : xl;
    if (i + 2 < nx) xhh = xat(j, i + 2);
    else xhh = has_g ? gat(j, nx - 1) : xh;
    double phl = vl_phi(xl, xll, xh), phh = vl_phi(xh, xhh, xl);
    if (!has_g) {
        if (i == 0) phl = 0.0;
        if (i == nx - 2) phh = 0.0;
    }
    const size_t fi = ((size_t)k * g.ny + j) * nx + i + 1;  // UAvg index of the face west of cell i+1
    double sp, sm;
    if (g.split) {
        sp = g.up[fi];
        sm = g.um[fi];
    } else {
        sp = fmax(g.up[fi], 0.0);
        sm = fmax(-g.up[fi], 0.0);
    }
    beta = 0.5 * (sp * phl + sm * phh);
    A = beta * (xh - xl);
}

template <class XA, class GA>
__device__ __forceinline__ void face_flux_y(const FctGrid& g, XA xat, GA gat, bool has_g, int k, int j, int i,
                                            double& beta, double& A) {
    const int nx = g.nx, ny = g.ny;
    const double xl = xat(j, i), xh = xat(j + 1, i);
    double xll, xhh;
    if (j > 0) xll = xat(j - 1, i);
    else xll = has_g ? gat(0, i) : xl;
    if (j + 2 < ny) xhh = xat(j + 2, i);
    else xhh = has_g ? gat(ny - 1, i) : xh;
    double phl = vl_phi(xl, xll, xh), phh = vl_phi(xh, xhh, xl);
    if (!has_g) {
        if (j == 0) phl = 0.0;
        if (j == ny - 2) phh = 0.0;
    }
    const size_t fi = ((size_t)k * ny + j + 1) * nx + i;
    double sp, sm;
    if (g.split) {
        sp = g.vp[fi];
        sm = g.vm[fi];
    } else {
        sp = fmax(g.vp[fi], 0.0);
        sm = fmax(-g.vp[fi], 0.0);
    }
    beta = 0.5 * (sp * phl + sm * phh);
    A = beta * (xh - xl);
}

template <class XA, class GA>
__device__ __forceinline__ Face face_x_t(const FctGrid& g, XA xat, GA gat, bool has_g, int k, int j, int i) {
    Face f;
    face_flux_x(g, xat, gat, has_g, k, j, i, f.beta, f.A);
    const FaceGeo q = face_geo_x(g, k, j, i);
    f.tr_lo = q.tr_lo;
    f.tr_hi = q.tr_hi;
    f.inv = q.inv;
    f.valid = q.valid;
    return f;
}

template <class XA, class GA>
__device__ __forceinline__ Face face_y_t(const FctGrid& g, XA xat, GA gat, bool has_g, int k, int j, int i) {
    Face f;
    face_flux_y(g, xat, gat, has_g, k, j, i, f.beta, f.A);
    const FaceGeo q = face_geo_y(g, k, j, i);
    f.tr_lo = q.tr_lo;
    f.tr_hi = q.tr_hi;
    f.inv = q.inv;
    f.valid = q.valid;
    return f;
}

// Per-bin geometry of the tiled source, tracer- and pass-independent: per x
// face (k, j, i < nx-1) and per y face (k, j < ny-1, i) tr_lo, tr_hi and
// valid (1 / 0); 1/dx per row; rho per cell.  Each value is face_geo_x /
// face_geo_y / rho_at itself, computed once instead of per tracer and pass.
struct FctGeo {
    const double* gx;
    const double* gy;
    const double* invx;
    const double* rho;
};

inline size_t geo_doubles(int nz, int ny, int nx) {
    const size_t nfx = nx > 1 ? (size_t)nz * ny * (nx - 1) : 0, nfy = ny > 1 ? (size_t)nz * (ny - 1) * nx : 0;
    return 3 * nfx + 3 * nfy + (size_t)ny + (size_t)nz * ny * nx;
}

__global__ void k_fct_geometry(FctGrid g, double* __restrict__ gx, double* __restrict__ gy,
                               double* __restrict__ invx, double* __restrict__ rho) {
    const size_t N = (size_t)g.nz * g.ny * g.nx;
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int i = (int)(n % g.nx), j = (int)((n / g.nx) % g.ny), k = (int)(n / ((size_t)g.nx * g.ny));
    rho[n] = rho_at(g, n);
    if (g.nx > 1 && i < g.nx - 1) {
        const FaceGeo q = face_geo_x(g, k, j, i);
        const size_t f = 3 * (((size_t)k * g.ny + j) * (g.nx - 1) + i);
        gx[f] = q.tr_lo;
        gx[f + 1] = q.tr_hi;
        gx[f + 2] = q.valid ? 1.0 : 0.0;
    }
    if (g.ny > 1 && j < g.ny - 1) {
        const FaceGeo q = face_geo_y(g, k, j, i);
        const size_t f = 3 * (((size_t)k * (g.ny - 1) + j) * g.nx + i);
        gy[f] = q.tr_lo;
        gy[f + 1] = q.tr_hi;
        gy[f + 2] = q.valid ? 1.0 : 0.0;
    }
    if (k == 0 && i == 0) {
        const double dxf = g.dx[j];
        invx[j] = 1.0 / (dxf > 0.0 ? dxf : 1.0);
    }
}

// Carve the geometry arrays out of one allocation of geo_doubles() doubles.
inline FctGeo geo_view(double* p, int nz, int ny, int nx) {
    const size_t nfx = nx > 1 ? (size_t)nz * ny * (nx - 1) : 0, nfy = ny > 1 ? (size_t)nz * (ny - 1) * nx : 0;
    FctGeo v;
    v.gx = p;
    v.gy = p + 3 * nfx;
    v.invx = p + 3 * nfx + 3 * nfy;
    v.rho = p + 3 * nfx + 3 * nfy + ny;
    return v;
}

// The same faces on x (and the ghost's x) stored per cell in global memory.
__device__ __forceinline__ Face face_x(const FctGrid& g, const double* x, const double* xg, int k, int j, int i) {
    const size_t lvl = (size_t)k * g.ny * g.nx;
    const int nx = g.nx;
    return face_x_t(g, [&](int jj, int ii) { return x[lvl + (size_t)jj * nx + ii]; },
                    [&](int jj, int ii) { return xg[lvl + (size_t)jj * nx + ii]; }, xg != nullptr, k, j, i);
}

__device__ __forceinline__ Face face_y(const FctGrid& g, const double* x, const double* xg, int k, int j, int i) {
    const size_t lvl = (size_t)k * g.ny * g.nx;
    const int nx = g.nx;
    return face_y_t(g, [&](int jj, int ii) { return x[lvl + (size_t)jj * nx + ii]; },
                    [&](int jj, int ii) { return xg[lvl + (size_t)jj * nx + ii]; }, xg != nullptr, k, j, i);
}

// Cell kernels run on a (64 x 4)-thread tile of (i, j) per block and one block
// row per (level, vector): blockIdx.z = k + nz * v.  No integer division.
struct CellIdx {
    int i, j, k, v;
    size_t n;
    bool ok;
};
__device__ __forceinline__ CellIdx cell_index(const FctGrid& g) {
    CellIdx c;
    c.i = blockIdx.x * 64 + (threadIdx.x & 63);
    c.j = blockIdx.y * 4 + (threadIdx.x >> 6);
    c.k = blockIdx.z % g.nz;
    c.v = blockIdx.z / g.nz;
    c.ok = c.i < g.nx && c.j < g.ny;
    c.n = ((size_t)c.k * g.ny + c.j) * g.nx + c.i;
    return c;
}

// R+ and R- per cell (Zalesak 1979 in the steady-state form of fct.py:381-447)
__global__ __launch_bounds__(256) void k_fct_R(FctGrid g, const double* __restrict__ x,
                                               const double* __restrict__ xg, double dt, double* __restrict__ Rp,
                                               double* __restrict__ Rm, long long xs, long long rs) {
    const CellIdx q = cell_index(g);
    if (!q.ok) return;
    x += (size_t)q.v * xs;
    if (xg) xg += (size_t)q.v * xs;
    Rp += (size_t)q.v * rs;
    Rm += (size_t)q.v * rs;
    const int nx = g.nx, ny = g.ny, i = q.i, j = q.j, k = q.k;
    const size_t n = q.n;
    const double xc = x[n];
    // neighbourhood bounds; beyond the edge the ghost (or the edge cell itself)
    const double* edge = xg ? xg : x;
    const double xw = i > 0 ? x[n - 1] : edge[n];
    const double xe = i < nx - 1 ? x[n + 1] : edge[n];
    const double xs_ = j > 0 ? x[n - nx] : edge[n];
    const double xn = j < ny - 1 ? x[n + nx] : edge[n];
    const double xmax = fmax(fmax(fmax(xc, xw), fmax(xe, xs_)), xn);
    const double xmin = fmin(fmin(fmin(xc, xw), fmin(xe, xs_)), xn);
    double pp = 0.0, pm = 0.0;
    auto add = [&](double v) {
        pp += fmax(v, 0.0);
        pm += fmax(-v, 0.0);
    };
    if (nx > 1) {
        if (i > 0) {  // west face: this cell is hi
            Face f = face_x(g, x, xg, k, j, i - 1);
            add(f.valid ? f.tr_hi * f.A * f.inv : 0.0);
        }
        if (i < nx - 1) {  // east face: this cell is lo
            Face f = face_x(g, x, xg, k, j, i);
            add(f.valid ? -f.tr_lo * f.A * f.inv : 0.0);
        }
    }
    if (ny > 1) {
        if (j > 0) {
            Face f = face_y(g, x, xg, k, j - 1, i);
            add(f.valid ? f.tr_hi * f.A * f.inv : 0.0);
        }
        if (j < ny - 1) {
            Face f = face_y(g, x, xg, k, j, i);
            add(f.valid ? -f.tr_lo * f.A * f.inv : 0.0);
        }
    }
    // R+- = min(1, Q+- / P+-) with Q+- = (x_max - x) / dt, (x - x_min) / dt: one division
    Rp[n] = pp > 0.0 ? fmin(1.0, (xmax - xc) / (dt * pp)) : 0.0;
    Rm[n] = pm > 0.0 ? fmin(1.0, (xc - xmin) / (dt * pm)) : 0.0;
}

__device__ __forceinline__ double face_C(const Face& f, double Rp_hi, double Rm_hi, double Rp_lo,
                                         double Rm_lo) {
    const double into_hi = f.valid ? f.tr_hi * f.A * f.inv : 0.0;
    const double C = into_hi >= 0.0 ? fmin(Rp_hi, Rm_lo) : fmin(Rm_hi, Rp_lo);
    return fmin(fmax(C, 0.0), 1.0);
}

// The source d = base + rho * sum over faces of +-tr*C*A/dx (C = 1 when not
// limited) in one pass per (64 x 8 tile of (i, j), level, vector): x = c/rho on
// the tile and a 3-cell halo, every face the tile and its 1-cell ring touch
// (each once), R+- on the tile and that ring, then the divergence of the
// tile's cells -- all through LDS, with nothing between the steps in HBM.
// Each face, R and cell value is the same expression as in face_x_t /
// face_y_t / k_fct_R, summed in the same order.
constexpr int FT_I = 64, FT_J = 8, FT_H = 3;
constexpr int FX_W = FT_I + 2 * FT_H, FX_H = FT_J + 2 * FT_H;  // x tile with halo
constexpr int XF_W = FT_I + 3, XF_H = FT_J + 2;  // x faces: lo i in [i0-2, i0+FT_I], j in [j0-1, j0+FT_J]
constexpr int YF_W = FT_I + 2, YF_H = FT_J + 3;  // y faces: lo j in [j0-2, j0+FT_J], i in [i0-1, i0+FT_I]
constexpr int RR_W = FT_I + 2, RR_H = FT_J + 2;  // R: the tile and its 1-cell ring

// One tile (blockIdx.x, blockIdx.y) of level k of vector v, for the kernels below.
__device__ __forceinline__ void fct_tile(const FctGrid& g, const FctGeo& geo, int k, int v,
                                         const double* __restrict__ c, long long cs,
                                         const double* __restrict__ cg, long long gs,
                                         const int* __restrict__ gmask, double dt, int limited,
                                         double* __restrict__ d, long long ds,
                                         const double* __restrict__ base, long long bs) {
    // per-tracer values in LDS; the faces' terrain weights and validity are the
    // bin's precomputed geometry, read where used (L2-resident, the same for
    // every tracer): less LDS per block, more blocks per CU
    __shared__ double sx[FX_H][FX_W];
    __shared__ double xfA[XF_H][XF_W];
    __shared__ double yfA[YF_H][YF_W];
    __shared__ double sRp[RR_H][RR_W], sRm[RR_H][RR_W];
    __shared__ double sinvx[XF_H];
    const int nx = g.nx, ny = g.ny;
    const int i0 = blockIdx.x * FT_I, j0 = blockIdx.y * FT_J;
    const size_t lvl = (size_t)k * ny * nx;
    const double* cv = c + (size_t)v * cs;
    // ghost of this vector: cg's row v, unless gmask says it has none (closed edges)
    const bool has_g = cg != nullptr && (gmask == nullptr || gmask[v] != 0);
    const double* gv = has_g ? cg + (size_t)v * gs : nullptr;
    for (int t = threadIdx.x; t < FX_H * FX_W; t += blockDim.x) {
        const int jj = t / FX_W, ii = t - jj * FX_W;
        const int j = j0 - FT_H + jj, i = i0 - FT_H + ii;
        if (j >= 0 && j < ny && i >= 0 && i < nx) {
            const size_t n = lvl + (size_t)j * nx + i;
            sx[jj][ii] = cv[n] / geo.rho[n];
        }
    }
    if (threadIdx.x < XF_H) {
        const int j = j0 - 1 + threadIdx.x;
        if (j >= 0 && j < ny) sinvx[threadIdx.x] = geo.invx[j];
    }
    __syncthreads();
    auto xat = [&](int j, int i) { return sx[j - j0 + FT_H][i - i0 + FT_H]; };
    auto gat = [&](int j, int i) {
        const size_t n = lvl + (size_t)j * nx + i;
        return gv[n] / geo.rho[n];
    };
    if (nx > 1)
        for (int t = threadIdx.x; t < XF_H * XF_W; t += blockDim.x) {
            const int fj = t / XF_W, fi = t - fj * XF_W;
            const int j = j0 - 1 + fj, i = i0 - 2 + fi;
            if (j >= 0 && j < ny && i >= 0 && i <= nx - 2) {
                double beta, A;
                face_flux_x(g, xat, gat, has_g, k, j, i, beta, A);
                xfA[fj][fi] = A;
            }
        }
    if (ny > 1)
        for (int t = threadIdx.x; t < YF_H * YF_W; t += blockDim.x) {
            const int fj = t / YF_W, fi = t - fj * YF_W;
            const int j = j0 - 2 + fj, i = i0 - 1 + fi;
            if (i >= 0 && i < nx && j >= 0 && j <= ny - 2) {
                double beta, A;
                face_flux_y(g, xat, gat, has_g, k, j, i, beta, A);
                yfA[fj][fi] = A;
            }
        }
    __syncthreads();
    const double invy = g.dy > 0.0 ? 1.0 / g.dy : 1.0;
    // faces with lo cell (j, i), rebuilt from LDS (beta is not needed past A)
    auto fx = [&](int j, int i) {
        const int fj = j - j0 + 1, fi = i - i0 + 2;
        const double* q = geo.gx + 3 * (((size_t)k * ny + j) * (nx - 1) + i);
        Face f;
        f.beta = 0.0;
        f.A = xfA[fj][fi];
        f.tr_lo = q[0];
        f.tr_hi = q[1];
        f.valid = q[2] != 0.0;
        f.inv = sinvx[fj];
        return f;
    };
    auto fy = [&](int j, int i) {
        const int fj = j - j0 + 2, fi = i - i0 + 1;
        const double* q = geo.gy + 3 * (((size_t)k * (ny - 1) + j) * nx + i);
        Face f;
        f.beta = 0.0;
        f.A = yfA[fj][fi];
        f.tr_lo = q[0];
        f.tr_hi = q[1];
        f.valid = q[2] != 0.0;
        f.inv = invy;
        return f;
    };
    if (limited) {
        for (int t = threadIdx.x; t < RR_H * RR_W; t += blockDim.x) {
            const int rj = t / RR_W, ri = t - rj * RR_W;
            const int j = j0 - 1 + rj, i = i0 - 1 + ri;
            if (j < 0 || j >= ny || i < 0 || i >= nx) continue;
            const double xc = xat(j, i);
            // beyond the edge: the ghost (or the cell itself); needed on the edge only
            const bool edge = i == 0 || i == nx - 1 || j == 0 || j == ny - 1;
            const double xe_ = has_g && edge ? gat(j, i) : xc;
            const double xw = i > 0 ? xat(j, i - 1) : xe_;
            const double xe = i < nx - 1 ? xat(j, i + 1) : xe_;
            const double xs_ = j > 0 ? xat(j - 1, i) : xe_;
            const double xn = j < ny - 1 ? xat(j + 1, i) : xe_;
            const double xmax = fmax(fmax(fmax(xc, xw), fmax(xe, xs_)), xn);
            const double xmin = fmin(fmin(fmin(xc, xw), fmin(xe, xs_)), xn);
            double pp = 0.0, pm = 0.0;
            auto add = [&](double a) {
                pp += fmax(a, 0.0);
                pm += fmax(-a, 0.0);
            };
            if (nx > 1) {
                if (i > 0) {
                    const Face f = fx(j, i - 1);
                    add(f.valid ? f.tr_hi * f.A * f.inv : 0.0);
                }
                if (i < nx - 1) {
                    const Face f = fx(j, i);
                    add(f.valid ? -f.tr_lo * f.A * f.inv : 0.0);
                }
            }
            if (ny > 1) {
                if (j > 0) {
                    const Face f = fy(j - 1, i);
                    add(f.valid ? f.tr_hi * f.A * f.inv : 0.0);
                }
                if (j < ny - 1) {
                    const Face f = fy(j, i);
                    add(f.valid ? -f.tr_lo * f.A * f.inv : 0.0);
                }
            }
            sRp[rj][ri] = pp > 0.0 ? fmin(1.0, (xmax - xc) / (dt * pp)) : 0.0;
            sRm[rj][ri] = pm > 0.0 ? fmin(1.0, (xc - xmin) / (dt * pm)) : 0.0;
        }
        __syncthreads();
    }
    auto Rp = [&](int j, int i) { return sRp[j - j0 + 1][i - i0 + 1]; };
    auto Rm = [&](int j, int i) { return sRm[j - j0 + 1][i - i0 + 1]; };
    for (int t = threadIdx.x; t < FT_J * FT_I; t += blockDim.x) {
        const int tj = t / FT_I, ti = t - tj * FT_I;
        const int j = j0 + tj, i = i0 + ti;
        if (j >= ny || i >= nx) continue;
        double acc = 0.0;
        if (nx > 1) {
            if (i > 0) {
                const Face f = fx(j, i - 1);
                const double C = limited ? face_C(f, Rp(j, i), Rm(j, i), Rp(j, i - 1), Rm(j, i - 1)) : 1.0;
                if (f.valid) acc += f.tr_hi * C * f.A * f.inv;
            }
            if (i < nx - 1) {
                const Face f = fx(j, i);
                const double C = limited ? face_C(f, Rp(j, i + 1), Rm(j, i + 1), Rp(j, i), Rm(j, i)) : 1.0;
                if (f.valid) acc -= f.tr_lo * C * f.A * f.inv;
            }
        }
        if (ny > 1) {
            if (j > 0) {
                const Face f = fy(j - 1, i);
                const double C = limited ? face_C(f, Rp(j, i), Rm(j, i), Rp(j - 1, i), Rm(j - 1, i)) : 1.0;
                if (f.valid) acc += f.tr_hi * C * f.A * f.inv;
            }
            if (j < ny - 1) {
                const Face f = fy(j, i);
                const double C = limited ? face_C(f, Rp(j + 1, i), Rm(j + 1, i), Rp(j, i), Rm(j, i)) : 1.0;
                if (f.valid) acc -= f.tr_lo * C * f.A * f.inv;
            }
        }
        const size_t n = lvl + (size_t)j * nx + i;
        const double val = acc * geo.rho[n];
        d[(size_t)v * ds + n] = base ? base[(size_t)v * bs + n] + val : val;
    }
}

// One bin: blockIdx.z = v * nz + k.
__global__ __launch_bounds__(256) void k_fct_tile(FctGrid g, FctGeo geo, const double* __restrict__ c, long long cs,
                                                  const double* __restrict__ cg, long long gs,
                                                  const int* __restrict__ gmask, double dt, int limited,
                                                  double* __restrict__ d, long long ds,
                                                  const double* __restrict__ base, long long bs) {
    fct_tile(g, geo, blockIdx.z % g.nz, blockIdx.z / g.nz, c, cs, cg, gs, gmask, dt, limited, d, ds, base, bs);
}

// Every bin of a pass in one launch (the bins' grids share the shape):
// blockIdx.z = (bin * nvec + v) * nz + k; the bin's grid and geometry come from
// bins[bin], and c, cg, d and base advance by their per-bin offsets.
struct FctBin {
    FctGrid g;
    FctGeo geo;
};
__global__ __launch_bounds__(256) void k_fct_tile_bins(const FctBin* __restrict__ bins, int nvec, int nz,
                                                       const double* __restrict__ c, long long cs, long long cb,
                                                       const double* __restrict__ cg, long long gs, long long gb,
                                                       const int* __restrict__ gmask, double dt, int limited,
                                                       double* __restrict__ d, long long ds, long long db,
                                                       const double* __restrict__ base, long long bs, long long bb) {
    const int k = blockIdx.z % nz, vb = blockIdx.z / nz;
    const int bin = vb / nvec, v = vb - bin * nvec;
    const FctGrid g = bins[bin].g;
    const FctGeo geo = bins[bin].geo;
    fct_tile(g, geo, k, v, c + (size_t)bin * cb, cs, cg ? cg + (size_t)bin * gb : nullptr, gs, gmask, dt, limited,
             d + (size_t)bin * db, ds, base ? base + (size_t)bin * bb : nullptr, bs);
}

// Frozen gains of L_AD: per x face (nz,ny,nx-1) and y face (nz,ny-1,nx), w_lo and w_hi.
__global__ void k_fct_gains(FctGrid g, const double* __restrict__ x, const double* __restrict__ Rp,
                            const double* __restrict__ Rm, double* __restrict__ wx, double* __restrict__ wy) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const size_t nfx = nx > 1 ? (size_t)nz * ny * (nx - 1) : 0;
    const size_t nfy = ny > 1 ? (size_t)nz * (ny - 1) * nx : 0;
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nfx) {
        const int i = (int)(t % (nx - 1)), j = (int)((t / (nx - 1)) % ny), k = (int)(t / ((size_t)(nx - 1) * ny));
        Face f = face_x(g, x, nullptr, k, j, i);
        const size_t lo = ((size_t)k * ny + j) * nx + i;
        const double C = face_C(f, Rp[lo + 1], Rm[lo + 1], Rp[lo], Rm[lo]);
        const double w = f.valid ? C * f.beta * f.inv : 0.0;
        wx[2 * t] = f.tr_lo * w;
        wx[2 * t + 1] = f.tr_hi * w;
    } else if (t < nfx + nfy) {
        const size_t u = t - nfx;
        const int i = (int)(u % nx), j = (int)((u / nx) % (ny - 1)), k = (int)(u / ((size_t)nx * (ny - 1)));
        Face f = face_y(g, x, nullptr, k, j, i);
        const size_t lo = ((size_t)k * ny + j) * nx + i;
        const double C = face_C(f, Rp[lo + nx], Rm[lo + nx], Rp[lo], Rm[lo]);
        const double w = f.valid ? C * f.beta * f.inv : 0.0;
        wy[2 * u] = f.tr_lo * w;
        wy[2 * u + 1] = f.tr_hi * w;
    }
}

// y = L_AD v (trans = 0) or L_AD^T v (trans = 1) from the frozen gains.
// L_x rows: hi cell of a face gets -w_hi (x_hi - x_lo), lo cell +w_lo (x_hi - x_lo);
// L_AD = diag(rho) L_x diag(1/rho).
__global__ void k_fct_linop(FctGrid g, const double* __restrict__ wx, const double* __restrict__ wy,
                            int trans, const double* __restrict__ v, double* __restrict__ y) {
    const size_t N = (size_t)g.nz * g.ny * g.nx;
    size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int nx = g.nx, ny = g.ny;
    const int i = (int)(n % nx), j = (int)((n / nx) % ny), k = (int)(n / ((size_t)nx * ny));
    const double rn = rho_at(g, n);
    double acc = 0.0;
    // u = v/rho (forward) or rho*v (transpose)
    auto u = [&](size_t m) { return trans ? v[m] * rho_at(g, m) : v[m] / rho_at(g, m); };
    const double un = u(n);
    if (nx > 1) {
        if (i > 0) {  // face (i-1, i): this cell is hi
            const size_t f = ((size_t)k * ny + j) * (nx - 1) + (i - 1);
            const double wlo = wx[2 * f], whi = wx[2 * f + 1];
            acc += trans ? (-whi * un + wlo * u(n - 1)) : (-whi * (un - u(n - 1)));
        }
        if (i < nx - 1) {  // face (i, i+1): this cell is lo
            const size_t f = ((size_t)k * ny + j) * (nx - 1) + i;
            const double wlo = wx[2 * f], whi = wx[2 * f + 1];
            acc += trans ? (-wlo * un + whi * u(n + 1)) : (wlo * (u(n + 1) - un));
        }
    }
    if (ny > 1) {
        if (j > 0) {
            const size_t f = ((size_t)k * (ny - 1) + (j - 1)) * nx + i;
            const double wlo = wy[2 * f], whi = wy[2 * f + 1];
            acc += trans ? (-whi * un + wlo * u(n - nx)) : (-whi * (un - u(n - nx)));
        }
        if (j < ny - 1) {
            const size_t f = ((size_t)k * (ny - 1) + j) * nx + i;
            const double wlo = wy[2 * f], whi = wy[2 * f + 1];
            acc += trans ? (-wlo * un + whi * u(n + nx)) : (wlo * (u(n + nx) - un));
        }
    }
    y[n] = trans ? acc / rn : acc * rn;
}

// (T rho)/rho * 86400 per cell from the opset's structured transport block
// (dtau*T stored), in the boundary k-major layout.
__global__ void k_rowsum_day(const double* __restrict__ V, const double* __restrict__ H,
                             const double* __restrict__ tdiag, const double* __restrict__ rho16,
                             double* __restrict__ out, int ncol, int nz, int nx, int ny, int px,
                             double inv_dtau) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ncol * nz) return;
    const int k = g / ncol, col = g - k * ncol;
    const int j = col / nx, i = col - j * nx;
    const double* Vr = V + (size_t)col * 256 + k * 16;
    double acc = tdiag[(size_t)col * 16 + k] * rho16[(size_t)col * 16 + k];
    for (int m = 0; m < 16; ++m)
        if (m != k) acc += Vr[m] * rho16[(size_t)col * 16 + m];
    const double* Hc = H + (size_t)col * 64 + k;
    const int cw = west_col(col, i, nx, px), ce = east_col(col, i, nx, px);
    if (cw >= 0) acc += Hc[0] * rho16[(size_t)cw * 16 + k];
    if (ce >= 0) acc += Hc[16] * rho16[(size_t)ce * 16 + k];
    if (j > 0) acc += Hc[32] * rho16[(size_t)(col - nx) * 16 + k];
    if (j < ny - 1) acc += Hc[48] * rho16[(size_t)(col + nx) * 16 + k];
    out[g] = acc * inv_dtau / rho16[(size_t)col * 16 + k] * 86400.0;
}

}  // namespace orbit

using namespace orbit;

struct orbit_fct {
    orbit_ctx* ctx = nullptr;
    int nz = 0, ny = 0, nx = 0, conc = 1;
    double* wx = nullptr;   // 2 per x face
    double* wy = nullptr;   // 2 per y face
    double* dP = nullptr;
    double* Dz = nullptr;
    double* tmp = nullptr;  // 2*N for host-pointer applies
};

namespace {

int check_metrics(const orbit_grid_metrics* g) {
    if (!g) return fail(ORBIT_ERR_INVALID, "fct: metrics is NULL");
    if (g->nz < 1 || g->ny < 1 || g->nx < 1) return fail(ORBIT_ERR_INVALID, "fct: bad grid shape");
    if (!g->u_plus || !g->v_plus || !g->dx || !g->dP || (g->concentration_state && !g->Dz))
        return fail(ORBIT_ERR_INVALID, "fct: missing metric array");
    if (g->split_fluxes && (!g->u_minus || !g->v_minus))
        return fail(ORBIT_ERR_INVALID, "fct: split fluxes need u_minus and v_minus");
    return ORBIT_OK;
}

// Device copies of the metric arrays (host pointers in g).
struct DevMetrics {
    std::vector<void*> bufs;
    FctGrid fg{};
    ~DevMetrics() {
        for (void* p : bufs) hipFree(p);
    }
    int make(hipStream_t st, const orbit_grid_metrics* g, bool host) {
        const size_t N = (size_t)g->nz * g->ny * g->nx;
        fg.nz = g->nz; fg.ny = g->ny; fg.nx = g->nx;
        fg.split = g->split_fluxes; fg.conc = g->concentration_state; fg.dy = g->dy;
        auto cp = [&](const double* src, size_t n, const double** dst) -> int {
            if (!src) {
                *dst = nullptr;
                return ORBIT_OK;
            }
            if (!host) {
                *dst = src;
                return ORBIT_OK;
            }
            void* p = nullptr;
            ORBIT_HIP_CHECK(hipMalloc(&p, n * 8));
            bufs.push_back(p);
            ORBIT_HIP_CHECK(hipMemcpyAsync(p, src, n * 8, hipMemcpyHostToDevice, st));
            *dst = (const double*)p;
            return ORBIT_OK;
        };
        int rc = cp(g->u_plus, N, &fg.up);
        if (!rc) rc = cp(g->split_fluxes ? g->u_minus : nullptr, N, &fg.um);
        if (!rc) rc = cp(g->v_plus, N, &fg.vp);
        if (!rc) rc = cp(g->split_fluxes ? g->v_minus : nullptr, N, &fg.vm);
        if (!rc) rc = cp(g->dx, (size_t)g->ny, &fg.dx);
        if (!rc) rc = cp(g->dP, N, &fg.dP);
        if (!rc) rc = cp(g->concentration_state ? g->Dz : nullptr, N, &fg.Dz);
        return rc;
    }
};

inline dim3 g1(size_t n) { return dim3((unsigned)((n + 255) / 256)); }
// launch shape of the cell kernels (cell_index): 64 x 4 cells per block
inline dim3 cell_grid(const FctGrid& g, int nvec) {
    return dim3((unsigned)((g.nx + 63) / 64), (unsigned)((g.ny + 3) / 4), (unsigned)(g.nz * nvec));
}
// launch shape of k_fct_tile: one FT_I x FT_J tile of a level of a vector per block
inline dim3 tile_grid(const FctGrid& g, int nvec) {
    return dim3((unsigned)((g.nx + FT_I - 1) / FT_I), (unsigned)((g.ny + FT_J - 1) / FT_J), (unsigned)(g.nz * nvec));
}
// [a, a + na) and [b, b + nb) (in doubles) share memory
inline bool overlaps(const double* a, size_t na, const double* b, size_t nb) {
    return a < b + nb && b < a + na;
}

int fct_source_impl(orbit_ctx* ctx, const orbit_grid_metrics* gm, bool host, const double* c,
                    double dt, int limited, const double* ghost, double* d) {
    int rc = check_metrics(gm);
    if (rc) return rc;
    if (!ctx || !c || !d) return fail(ORBIT_ERR_INVALID, "fct: NULL argument");
    if (!(dt > 0.0)) return fail(ORBIT_ERR_INVALID, "fct: dt must be positive");
    hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    DevMetrics M;
    rc = M.make(st, gm, host);
    if (rc) return rc;
    const size_t N = (size_t)gm->nz * gm->ny * gm->nx;
    std::vector<void*> tmp;
    auto dalloc = [&](size_t n) -> double* {
        void* p = nullptr;
        if (hipMalloc(&p, n * 8) != hipSuccess) return nullptr;
        tmp.push_back(p);
        return (double*)p;
    };
    const double *dc = c, *dg = ghost;
    double* dd = d;
    if (!host && (overlaps(d, N, c, N) || (ghost && overlaps(d, N, ghost, N))))
        rc = fail(ORBIT_ERR_INVALID, "fct: d must not overlap c or ghost");
    if (!rc && host) {
        double* a = dalloc(N);
        double* b = ghost ? dalloc(N) : nullptr;
        double* o = dalloc(N);
        if (!a || !o || (ghost && !b)) rc = fail(ORBIT_ERR_NOMEM, "fct: device allocation failed");
        if (!rc && hipMemcpyAsync(a, c, N * 8, hipMemcpyHostToDevice, st) != hipSuccess) rc = fail(ORBIT_ERR_HIP, "fct: copy");
        if (!rc && ghost && hipMemcpyAsync(b, ghost, N * 8, hipMemcpyHostToDevice, st) != hipSuccess)
            rc = fail(ORBIT_ERR_HIP, "fct: copy");
        dc = a;
        dg = b;
        dd = o;
    }
    double* geo = rc ? nullptr : dalloc(geo_doubles(gm->nz, gm->ny, gm->nx));
    if (!rc && !geo) rc = fail(ORBIT_ERR_NOMEM, "fct: device allocation failed");
    if (!rc) {
        const FctGeo gv = geo_view(geo, gm->nz, gm->ny, gm->nx);
        hipLaunchKernelGGL(k_fct_geometry, g1(N), dim3(256), 0, st, M.fg, (double*)gv.gx, (double*)gv.gy,
                           (double*)gv.invx, (double*)gv.rho);
        hipLaunchKernelGGL(k_fct_tile, tile_grid(M.fg, 1), dim3(256), 0, st, M.fg, gv, dc, 0LL, dg, 0LL,
                           (const int*)nullptr, dt, limited, dd, 0LL, (const double*)nullptr, 0LL);
        if (hipGetLastError() != hipSuccess) rc = fail(ORBIT_ERR_HIP, "fct: kernel launch failed");
    }
    if (!rc && host && hipMemcpyAsync(d, dd, N * 8, hipMemcpyDeviceToHost, st) != hipSuccess)
        rc = fail(ORBIT_ERR_HIP, "fct: copy back");
    hipStreamSynchronize(st);
    for (void* p : tmp) hipFree(p);
    return rc;
}

}  // namespace

extern "C" int orbit_fct_source(orbit_ctx* ctx, const orbit_grid_metrics* g, const double* c, double dt,
                                int limited, const double* ghost, double* d_out) {
    return fct_source_impl(ctx, g, true, c, dt, limited, ghost, d_out);
}

extern "C" int orbit_fct_source_dev(orbit_ctx* ctx, const orbit_grid_metrics* g, const double* c_dev, double dt,
                                    int limited, const double* ghost_dev, double* d_dev) {
    return fct_source_impl(ctx, g, false, c_dev, dt, limited, ghost_dev, d_dev);
}

extern "C" int orbit_fct_linop_create(orbit_ctx* ctx, const orbit_grid_metrics* gm, const double* c_base,
                                      double dt, orbit_fct** out) {
    int rc = check_metrics(gm);
    if (rc) return rc;
    if (!ctx || !c_base || !out) return fail(ORBIT_ERR_INVALID, "fct_linop: NULL argument");
    if (!(dt > 0.0)) return fail(ORBIT_ERR_INVALID, "fct_linop: dt must be positive");
    hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    DevMetrics M;
    rc = M.make(st, gm, true);
    if (rc) return rc;
    const int nz = gm->nz, ny = gm->ny, nx = gm->nx;
    const size_t N = (size_t)nz * ny * nx;
    const size_t nfx = nx > 1 ? (size_t)nz * ny * (nx - 1) : 0, nfy = ny > 1 ? (size_t)nz * (ny - 1) * nx : 0;
    auto* L = new orbit_fct();
    L->ctx = ctx;
    L->nz = nz; L->ny = ny; L->nx = nx; L->conc = gm->concentration_state;
    bool ok = hipMalloc((void**)&L->wx, std::max<size_t>(2 * nfx, 1) * 8) == hipSuccess &&
              hipMalloc((void**)&L->wy, std::max<size_t>(2 * nfy, 1) * 8) == hipSuccess &&
              hipMalloc((void**)&L->dP, N * 8) == hipSuccess && hipMalloc((void**)&L->Dz, N * 8) == hipSuccess &&
              hipMalloc((void**)&L->tmp, 2 * N * 8) == hipSuccess;
    double* cb = nullptr;
    double* R = nullptr;
    ok = ok && hipMalloc((void**)&cb, 2 * N * 8) == hipSuccess && hipMalloc((void**)&R, 2 * N * 8) == hipSuccess;
    if (!ok) {
        if (cb) hipFree(cb);
        if (R) hipFree(R);
        orbit_fct_linop_destroy(L);
        return fail(ORBIT_ERR_NOMEM, "fct_linop: device allocation failed");
    }
    hipMemcpyAsync(L->dP, gm->dP, N * 8, hipMemcpyHostToDevice, st);
    if (gm->concentration_state)
        hipMemcpyAsync(L->Dz, gm->Dz, N * 8, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(cb, c_base, N * 8, hipMemcpyHostToDevice, st);
    double* xb = cb + N;
    hipLaunchKernelGGL(k_fct_x, g1(N), dim3(256), 0, st, M.fg, (const double*)cb, 0LL, xb, 0LL, (int)N);
    hipLaunchKernelGGL(k_fct_R, cell_grid(M.fg, 1), dim3(256), 0, st, M.fg, (const double*)xb,
                       (const double*)nullptr, dt, R, R + N, 0LL, 0LL);
    if (nfx + nfy > 0)
        hipLaunchKernelGGL(k_fct_gains, g1(nfx + nfy), dim3(256), 0, st, M.fg, (const double*)xb, R, R + N, L->wx,
                           L->wy);
    hipError_t e = hipGetLastError();
    hipStreamSynchronize(st);
    hipFree(cb);
    hipFree(R);
    if (e != hipSuccess) {
        orbit_fct_linop_destroy(L);
        return fail(ORBIT_ERR_HIP, std::string("fct_linop: ") + hipGetErrorString(e));
    }
    *out = L;
    return ORBIT_OK;
}

static FctGrid linop_grid(const orbit_fct* L) {
    FctGrid g{};
    g.nz = L->nz; g.ny = L->ny; g.nx = L->nx; g.conc = L->conc;
    g.dP = L->dP; g.Dz = L->Dz;
    return g;
}

extern "C" int orbit_fct_linop_apply_dev(orbit_ctx* ctx, orbit_fct* L, int trans, const double* x_dev,
                                         double* y_dev) {
    if (!ctx || !L || !x_dev || !y_dev) return fail(ORBIT_ERR_INVALID, "fct_linop_apply: NULL argument");
    hipSetDevice(ctx->device);
    const size_t N = (size_t)L->nz * L->ny * L->nx;
    hipLaunchKernelGGL(k_fct_linop, g1(N), dim3(256), 0, ctx->stream, linop_grid(L), L->wx, L->wy, trans ? 1 : 0,
                       x_dev, y_dev);
    ORBIT_HIP_CHECK(hipGetLastError());
    return ORBIT_OK;
}

extern "C" int orbit_fct_linop_apply(orbit_ctx* ctx, orbit_fct* L, int trans, const double* x, double* y) {
    if (!ctx || !L || !x || !y) return fail(ORBIT_ERR_INVALID, "fct_linop_apply: NULL argument");
    hipSetDevice(ctx->device);
    const size_t N = (size_t)L->nz * L->ny * L->nx;
    hipStream_t st = ctx->stream;
    ORBIT_HIP_CHECK(hipMemcpyAsync(L->tmp, x, N * 8, hipMemcpyHostToDevice, st));
    int rc = orbit_fct_linop_apply_dev(ctx, L, trans, L->tmp, L->tmp + N);
    if (rc) return rc;
    ORBIT_HIP_CHECK(hipMemcpyAsync(y, L->tmp + N, N * 8, hipMemcpyDeviceToHost, st));
    ORBIT_HIP_CHECK(hipStreamSynchronize(st));
    return ORBIT_OK;
}

extern "C" int orbit_fct_linop_csr(orbit_ctx* ctx, orbit_fct* L, int32_t* rowptr, int32_t* colidx, double* vals,
                                   int* nnz) {
    if (!ctx || !L || !rowptr || !colidx || !vals || !nnz) return fail(ORBIT_ERR_INVALID, "fct_linop_csr: NULL argument");
    hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    const int nz = L->nz, ny = L->ny, nx = L->nx;
    const size_t N = (size_t)nz * ny * nx;
    const size_t nfx = nx > 1 ? (size_t)nz * ny * (nx - 1) : 0, nfy = ny > 1 ? (size_t)nz * (ny - 1) * nx : 0;
    std::vector<double> wx(2 * nfx), wy(2 * nfy), dP(N), Dz(N);
    if (nfx) ORBIT_HIP_CHECK(hipMemcpyAsync(wx.data(), L->wx, 2 * nfx * 8, hipMemcpyDeviceToHost, st));
    if (nfy) ORBIT_HIP_CHECK(hipMemcpyAsync(wy.data(), L->wy, 2 * nfy * 8, hipMemcpyDeviceToHost, st));
    ORBIT_HIP_CHECK(hipMemcpyAsync(dP.data(), L->dP, N * 8, hipMemcpyDeviceToHost, st));
    ORBIT_HIP_CHECK(hipMemcpyAsync(Dz.data(), L->Dz, N * 8, hipMemcpyDeviceToHost, st));
    ORBIT_HIP_CHECK(hipStreamSynchronize(st));
    auto rho = [&](size_t n) {
        if (!L->conc) return 1.0;
        return (dP[n] > 0.0 && Dz[n] > 0.0) ? dP[n] / (kGrav * Dz[n]) : 1.0;
    };
    // each row: up to 5 entries, columns sorted: S, W, self, E, N
    int32_t cnt = 0;
    for (size_t n = 0; n < N; ++n) {
        rowptr[n] = cnt;
        const int i = (int)(n % nx), j = (int)((n / nx) % ny), k = (int)(n / ((size_t)nx * ny));
        double cS = 0, cW = 0, cC = 0, cE = 0, cN = 0;
        bool hS = false, hW = false, hE = false, hN = false;
        if (nx > 1 && i > 0) {
            const size_t f = ((size_t)k * ny + j) * (nx - 1) + (i - 1);
            cC += -wx[2 * f + 1];
            cW += wx[2 * f + 1];
            hW = true;
        }
        if (nx > 1 && i < nx - 1) {
            const size_t f = ((size_t)k * ny + j) * (nx - 1) + i;
            cC += -wx[2 * f];
            cE += wx[2 * f];
            hE = true;
        }
        if (ny > 1 && j > 0) {
            const size_t f = ((size_t)k * (ny - 1) + (j - 1)) * nx + i;
            cC += -wy[2 * f + 1];
            cS += wy[2 * f + 1];
            hS = true;
        }
        if (ny > 1 && j < ny - 1) {
            const size_t f = ((size_t)k * (ny - 1) + j) * nx + i;
            cC += -wy[2 * f];
            cN += wy[2 * f];
            hN = true;
        }
        const double rn = rho(n);
        auto put = [&](size_t m, double v) {
            colidx[cnt] = (int32_t)m;
            vals[cnt] = v * rn / rho(m);
            ++cnt;
        };
        if (hS) put(n - nx, cS);
        if (hW) put(n - 1, cW);
        put(n, cC);
        if (hE) put(n + 1, cE);
        if (hN) put(n + nx, cN);
    }
    rowptr[N] = cnt;
    *nnz = cnt;
    return ORBIT_OK;
}

extern "C" int orbit_fct_linop_destroy(orbit_fct* L) {
    if (!L) return ORBIT_OK;
    if (L->ctx) {
        hipSetDevice(L->ctx->device);
        hipStreamSynchronize(L->ctx->stream);
    }
    if (L->wx) hipFree(L->wx);
    if (L->wy) hipFree(L->wy);
    if (L->dP) hipFree(L->dP);
    if (L->Dz) hipFree(L->Dz);
    if (L->tmp) hipFree(L->tmp);
    delete L;
    return ORBIT_OK;
}

// ------------------------------------------------------------------------- //
// air-mass-balance pre-flight
// ------------------------------------------------------------------------- //
namespace {
// numpy's default (linear interpolation) percentile of an unsorted sample
double percentile(std::vector<double> v, double q) {
    if (v.empty()) return std::nan("");
    std::sort(v.begin(), v.end());
    const double pos = q / 100.0 * (double)(v.size() - 1);
    const size_t lo = (size_t)std::floor(pos);
    const size_t hi = std::min(lo + 1, v.size() - 1);
    const double f = pos - (double)lo;
    return v[lo] + f * (v[hi] - v[lo]);
}
}  // namespace

extern "C" int orbit_mass_balance(orbit_ctx* ctx, orbit_opset* op, const double* rho, orbit_mb_stats* st) {
    if (!ctx || !op || !rho || !st) return fail(ORBIT_ERR_INVALID, "mass_balance: NULL argument");
    if (!op->has_tdiag) return fail(ORBIT_ERR_INVALID, "mass_balance: opset was not made with the shared form");
    if (op->ngroups != 1) return fail(ORBIT_ERR_INVALID, "mass_balance: opset has per-tracer transport");
    if (op->nz > 32) return fail(ORBIT_ERR_INVALID, "mass_balance: nz > 32");
    hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    const int nz = op->nz, ny = op->ny, nx = op->nx, ncol = op->ncol, N = op->N;
    std::vector<double> rho16((size_t)op->n16, 1.0);
    for (int n = 0; n < N; ++n) {
        const int k = n / ncol, c = n - k * ncol;
        rho16[(size_t)c * 16 + k] = rho[n];
    }
    double *drho = nullptr, *dout = nullptr;
    ORBIT_HIP_CHECK(hipMalloc(&drho, (size_t)op->n16 * 8));
    if (hipMalloc(&dout, (size_t)N * 8) != hipSuccess) {
        hipFree(drho);
        return fail(ORBIT_ERR_NOMEM, "mass_balance: device allocation failed");
    }
    hipMemcpyAsync(drho, rho16.data(), (size_t)op->n16 * 8, hipMemcpyHostToDevice, s);
    std::vector<double> r((size_t)N);
    std::vector<double> ri, rt;
    std::vector<std::vector<double>> lay(nz), lays(nz);
    int rc = ORBIT_OK;
    for (int b = 0; b < op->nbins && !rc; ++b) {
        const BinOps& B = op->bins[b];
        hipLaunchKernelGGL(k_rowsum_day, g1(N), dim3(256), 0, s, B.V64, B.H64, B.tdiag, drho, dout, ncol, nz, nx, ny,
                           op->px, 1.0 / op->dtau);
        if (hipMemcpyAsync(r.data(), dout, (size_t)N * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            rc = fail(ORBIT_ERR_HIP, "mass_balance: copy back failed");
            break;
        }
        for (int k = 0; k < nz; ++k)
            for (int j = 1; j < ny - 1; ++j)
                for (int i = 1; i < nx - 1; ++i) {
                    const double v = r[(size_t)k * ncol + (size_t)j * nx + i];
                    if (!std::isfinite(v)) continue;
                    (k < nz - 1 ? ri : rt).push_back(std::fabs(v));
                    lay[k].push_back(std::fabs(v));
                    lays[k].push_back(v);
                }
    }
    hipFree(drho);
    hipFree(dout);
    if (rc) return rc;
    if (ri.empty() && rt.empty()) return fail(ORBIT_ERR_INVALID, "mass_balance: grid has no interior cells");
    st->interior_median = percentile(ri, 50.0);
    st->interior_p90 = percentile(ri, 90.0);
    st->interior_p99 = percentile(ri, 99.0);
    st->interior_max = ri.empty() ? std::nan("") : *std::max_element(ri.begin(), ri.end());
    size_t a1 = 0, a2 = 0;
    for (double v : ri) {
        a1 += v > 0.01;
        a2 += v > 0.1;
    }
    st->interior_frac_above_0_01 = ri.empty() ? 0.0 : (double)a1 / ri.size();
    st->interior_frac_above_0_1 = ri.empty() ? 0.0 : (double)a2 / ri.size();
    st->top_median = percentile(rt, 50.0);
    st->top_p90 = percentile(rt, 90.0);
    for (int k = 0; k < 32; ++k) {
        st->layer_p90[k] = 0.0;
        st->layer_mean_signed[k] = 0.0;
    }
    for (int k = 0; k < nz; ++k) {
        st->layer_p90[k] = percentile(lay[k], 90.0);
        double sum = 0.0;
        for (double v : lays[k]) sum += v;
        st->layer_mean_signed[k] = lays[k].empty() ? std::nan("") : sum / lays[k].size();
    }
    return ORBIT_OK;
}


// ------------------------------------------------------------------------- //
// device-resident grid metrics of one bin and the batched source
// ------------------------------------------------------------------------- //
struct orbit_fct_grid {
    orbit_ctx* ctx = nullptr;
    std::vector<void*> bufs;
    FctGrid fg{};
    FctGeo geo{};  // in bufs
};

extern "C" int orbit_fct_grid_create(orbit_ctx* ctx, const orbit_grid_metrics* gm, orbit_fct_grid** out) {
    int rc = check_metrics(gm);
    if (rc) return rc;
    if (!ctx || !out) return fail(ORBIT_ERR_INVALID, "fct_grid: NULL argument");
    hipSetDevice(ctx->device);
    auto* G = new orbit_fct_grid();
    G->ctx = ctx;
    DevMetrics M;
    rc = M.make(ctx->stream, gm, true);
    if (rc) {
        delete G;
        return rc;
    }
    G->fg = M.fg;
    G->bufs.swap(M.bufs);  // ownership moves to the grid object
    // the source's geometry, once for every tracer and pass
    void* geo = nullptr;
    const size_t N = (size_t)gm->nz * gm->ny * gm->nx;
    if (hipMalloc(&geo, geo_doubles(gm->nz, gm->ny, gm->nx) * 8) != hipSuccess) {
        orbit_fct_grid_destroy(G);
        return fail(ORBIT_ERR_NOMEM, "fct_grid: device allocation failed");
    }
    G->bufs.push_back(geo);
    G->geo = geo_view((double*)geo, gm->nz, gm->ny, gm->nx);
    hipLaunchKernelGGL(k_fct_geometry, g1(N), dim3(256), 0, ctx->stream, G->fg, (double*)G->geo.gx,
                       (double*)G->geo.gy, (double*)G->geo.invx, (double*)G->geo.rho);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess) {
        orbit_fct_grid_destroy(G);
        return fail(ORBIT_ERR_HIP, "fct_grid: geometry kernel failed");
    }
    *out = G;
    return ORBIT_OK;
}

extern "C" int orbit_fct_grid_destroy(orbit_fct_grid* G) {
    if (!G) return ORBIT_OK;
    if (G->ctx) {
        hipSetDevice(G->ctx->device);
        hipStreamSynchronize(G->ctx->stream);
    }
    for (void* p : G->bufs) hipFree(p);
    delete G;
    return ORBIT_OK;
}

namespace {
// The batched source with the ghost rows on a stride of their own.
// gmask_dev (nullable, device, nvec ints): vector v reads the ghost only where gmask_dev[v] != 0.
int fct_source_batch(orbit_ctx* ctx, orbit_fct_grid* G, int nvec, const double* c_dev, long long c_stride,
                     const double* ghost_dev, long long ghost_stride, double dt, int limited, const double* base_dev,
                     long long base_stride, double* out_dev, long long out_stride, const int* gmask_dev = nullptr) {
    if (!ctx || !G || !c_dev || !out_dev || nvec < 1) return fail(ORBIT_ERR_INVALID, "fct_source_batch: bad argument");
    if (!(dt > 0.0)) return fail(ORBIT_ERR_INVALID, "fct_source_batch: dt must be positive");
    if (nvec > 65535 || (long long)nvec * G->fg.nz > 65535)
        return fail(ORBIT_ERR_INVALID, "fct_source_batch: nvec * nz > 65535");
    hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    const size_t N = (size_t)G->fg.nz * G->fg.ny * G->fg.nx;
    // tiles read c (and the ghost) over a halo that other tiles' outputs cover:
    // the output must not share memory with them (base may be the output itself)
    auto span = [&](long long stride) { return (size_t)(nvec - 1) * (size_t)stride + N; };
    if (c_stride < 0 || out_stride < 0 || base_stride < 0 || ghost_stride < 0)
        return fail(ORBIT_ERR_INVALID, "fct_source_batch: negative stride");
    if (overlaps(out_dev, span(out_stride), c_dev, span(c_stride)) ||
        (ghost_dev && overlaps(out_dev, span(out_stride), ghost_dev, span(ghost_stride))))
        return fail(ORBIT_ERR_INVALID, "fct_source_batch: out must not overlap c or ghost");
    if (base_dev && overlaps(out_dev, span(out_stride), base_dev, span(base_stride)) &&
        !(base_dev == out_dev && base_stride == out_stride))
        return fail(ORBIT_ERR_INVALID, "fct_source_batch: base overlaps out other than as out itself");
    hipLaunchKernelGGL(k_fct_tile, tile_grid(G->fg, nvec), dim3(256), 0, st, G->fg, G->geo, c_dev, c_stride, ghost_dev,
                       ghost_stride, ghost_dev ? gmask_dev : (const int*)nullptr, dt, limited, out_dev, out_stride,
                       base_dev, base_stride);
    ORBIT_HIP_CHECK(hipGetLastError());
    return ORBIT_OK;
}
}  // namespace

extern "C" int orbit_fct_source_batch_dev(orbit_ctx* ctx, orbit_fct_grid* G, int nvec, const double* c_dev,
                                          long long c_stride, const double* ghost_dev, double dt, int limited,
                                          const double* base_dev, long long base_stride, double* out_dev,
                                          long long out_stride) {
    return fct_source_batch(ctx, G, nvec, c_dev, c_stride, ghost_dev, c_stride, dt, limited, base_dev, base_stride,
                            out_dev, out_stride);
}

// Forward FCT deferred correction (reference orbit/cli.py:1212-1241) with
// every pass on the device: pass 0 solves with e; pass p adds the limited
// anti-diffusive source of the previous orbit's end-of-bin states
// d_AD(c[tau+1]) to e and solves again.  FP64: warm-started from the previous
// c0, the same launches as the host loop over orbit_fct_source_batch_dev and
// orbit_solve_forward_dev.  Mixed precision: each pass after the first starts
// its states from the previous orbit (solve_forward_pass), so its first
// residual is only that of the change of the forcing.
// The forward passes; pass_orbits (nullable) receives the orbits of passes
// 0 .. n_passes-2 in turn, (n_passes - 1, nt, n_bins + 1, N).
static int forward_fct_passes(orbit_ctx* ctx, orbit_opset* op, orbit_fct_grid* const* grids, const int* tracers,
                              int nt, const double* e_dev, const double* ghost_dev, const int* ghost_tracers,
                              int n_passes, const orbit_opts* opts, double* orbit_dev, orbit_info* info,
                              double* pass_orbits) {
    if (!ctx || !op || !grids || !tracers || nt < 1 || !e_dev || !orbit_dev || !info)
        return fail(ORBIT_ERR_INVALID, "forward_fct: NULL argument or empty tracer list");
    if (n_passes < 1) return fail(ORBIT_ERR_INVALID, "forward_fct: n_passes must be >= 1");
    if (opts && opts->rhs_is_s) return fail(ORBIT_ERR_INVALID, "forward_fct: the passes need e, not s");
    const int nbn = op->nbins;
    const size_t N = (size_t)op->N;
    for (int b = 0; b < nbn; ++b) {
        const orbit_fct_grid* G = grids[b];
        if (!G || (size_t)G->fg.nz * G->fg.ny * G->fg.nx != N || G->fg.nx != op->nx || G->fg.ny != op->ny)
            return fail(ORBIT_ERR_INVALID, "forward_fct: every bin needs an FCT grid of the operator set's shape");
        if (G->ctx != ctx) return fail(ORBIT_ERR_INVALID, "forward_fct: FCT grids and the solves must share ctx");
    }
    orbit_opts o;
    orbit_opts_default(&o);
    if (opts) o = *opts;
    o.warm_orbit = 0;
    hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    const bool inc = o.precision == ORBIT_PREC_MIXED && n_passes > 1;
    const size_t ne = (size_t)nt * nbn * N, n16 = (size_t)op->n16;
    // scratch: the pass's forcing, c0 (FP64 passes), s~ (mixed passes) and the
    // per-tracer ghost flags
    double *e_eff = nullptr, *c0 = nullptr, *s_state = nullptr;
    int* gmask = nullptr;
    FctBin* bins = nullptr;  // every bin's grid, for one FCT launch per pass
    const size_t nbin_bytes = ((size_t)nbn * sizeof(FctBin) + 255) / 256 * 256;
    const bool one_launch = (long long)op->nz * nt * nbn <= 65535;
    if (n_passes > 1) {
        int rc0 = ctx->ensure(8, nbin_bytes + (ne + (size_t)nt * N + (size_t)nt * n16) * 8 + (size_t)nt * 4);
        if (rc0) return rc0;
        bins = ctx->buf<FctBin>(8);
        e_eff = reinterpret_cast<double*>(reinterpret_cast<char*>(bins) + nbin_bytes);
        c0 = e_eff + ne;
        s_state = c0 + (size_t)nt * N;
        if (ghost_dev && ghost_tracers) {
            gmask = reinterpret_cast<int*>(s_state + (size_t)nt * n16);
            ORBIT_HIP_CHECK(hipMemcpyAsync(gmask, ghost_tracers, (size_t)nt * 4, hipMemcpyHostToDevice, st));
        }
        std::vector<FctBin> hb(nbn);
        for (int b = 0; b < nbn; ++b) hb[b] = FctBin{grids[b]->fg, grids[b]->geo};
        ORBIT_HIP_CHECK(hipMemcpyAsync(bins, hb.data(), (size_t)nbn * sizeof(FctBin), hipMemcpyHostToDevice, st));
        ORBIT_HIP_CHECK(hipStreamSynchronize(st));  // the host arrays may go away
    }
    PassDelta pd;
    pd.s_state = s_state;
    int rc = inc ? solve_forward_pass(ctx, op, tracers, nt, e_dev, nullptr, &pd, &o, orbit_dev, info)
                 : orbit_solve_forward_dev(ctx, op, tracers, nt, e_dev, nullptr, &o, orbit_dev, info);
    if (rc < 0) return rc;
    int worst = rc;
    if (n_passes == 1) return worst;
    const long long os = (long long)(nbn + 1) * N, es = (long long)nbn * N;
    for (int p = 1; p < n_passes; ++p) {
        if (pass_orbits)
            ORBIT_HIP_CHECK(hipMemcpyAsync(pass_orbits + (size_t)(p - 1) * nt * os, orbit_dev, (size_t)nt * os * 8,
                                           hipMemcpyDeviceToDevice, st));
        if (one_launch) {
            // every bin at once: the same per-(bin, tracer, level) tiles as a launch per bin
            const FctGrid& g0 = grids[0]->fg;
            const dim3 grid((unsigned)((g0.nx + FT_I - 1) / FT_I), (unsigned)((g0.ny + FT_J - 1) / FT_J),
                            (unsigned)(g0.nz * nt * nbn));
            hipLaunchKernelGGL(k_fct_tile_bins, grid, dim3(256), 0, st, (const FctBin*)bins, nt, g0.nz,
                               (const double*)(orbit_dev + N), os, (long long)N, ghost_dev, es, (long long)N,
                               (const int*)(ghost_dev ? gmask : nullptr), op->dtau, 1, e_eff, es, (long long)N,
                               e_dev, es, (long long)N);
            ORBIT_HIP_CHECK(hipGetLastError());
        } else {
            for (int tau = 0; tau < nbn; ++tau) {
                rc = fct_source_batch(ctx, grids[tau], nt, orbit_dev + (size_t)(tau + 1) * N, os,
                                      ghost_dev ? ghost_dev + (size_t)tau * N : nullptr, es, op->dtau, 1,
                                      e_dev + (size_t)tau * N, es, e_eff + (size_t)tau * N, es, gmask);
                if (rc) return rc;
            }
        }
        if (inc) {
            orbit_opts oi = o;
            oi.warm_orbit = 1;
            rc = solve_forward_pass(ctx, op, tracers, nt, e_eff, orbit_dev, &pd, &oi, orbit_dev,
                                    info + (size_t)p * nt);
        } else {
            ORBIT_HIP_CHECK(
                hipMemcpy2DAsync(c0, N * 8, orbit_dev, (size_t)os * 8, N * 8, nt, hipMemcpyDeviceToDevice, st));
            rc = orbit_solve_forward_dev(ctx, op, tracers, nt, e_eff, c0, &o, orbit_dev, info + (size_t)p * nt);
        }
        if (rc < 0) return rc;
        worst = std::max(worst, rc);
    }
    return worst;
}

extern "C" int orbit_forward_fct_dev(orbit_ctx* ctx, orbit_opset* op, orbit_fct_grid* const* grids,
                                     const int* tracers, int nt, const double* e_dev, const double* ghost_dev,
                                     const int* ghost_tracers, int n_passes, const orbit_opts* opts,
                                     double* orbit_dev, orbit_info* info) {
    return forward_fct_passes(ctx, op, grids, tracers, nt, e_dev, ghost_dev, ghost_tracers, n_passes, opts,
                              orbit_dev, info, nullptr);
}

extern "C" int orbit_forward_fct_passes_dev(orbit_ctx* ctx, orbit_opset* op, orbit_fct_grid* const* grids,
                                            const int* tracers, int nt, const double* e_dev, const double* ghost_dev,
                                            const int* ghost_tracers, int n_passes, const orbit_opts* opts,
                                            double* orbit_dev, orbit_info* info, double* pass_orbits_dev) {
    if (n_passes > 1 && !pass_orbits_dev) return fail(ORBIT_ERR_INVALID, "forward_fct_passes: NULL pass_orbits");
    return forward_fct_passes(ctx, op, grids, tracers, nt, e_dev, ghost_dev, ghost_tracers, n_passes, opts,
                              orbit_dev, info, pass_orbits_dev);
}

extern "C" int orbit_fct_linop_create_dev(orbit_ctx* ctx, orbit_fct_grid* G, const double* c_base_dev, double dt,
                                          orbit_fct** out) {
    if (!ctx || !G || !c_base_dev || !out) return fail(ORBIT_ERR_INVALID, "fct_linop_create_dev: NULL argument");
    if (!(dt > 0.0)) return fail(ORBIT_ERR_INVALID, "fct_linop_create_dev: dt must be positive");
    hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    const FctGrid& fg = G->fg;
    const int nz = fg.nz, ny = fg.ny, nx = fg.nx;
    const size_t N = (size_t)nz * ny * nx;
    const size_t nfx = nx > 1 ? (size_t)nz * ny * (nx - 1) : 0, nfy = ny > 1 ? (size_t)nz * (ny - 1) * nx : 0;
    auto* L = new orbit_fct();
    L->ctx = ctx;
    L->nz = nz; L->ny = ny; L->nx = nx; L->conc = fg.conc;
    double* R = nullptr;
    bool ok = hipMalloc((void**)&L->wx, std::max<size_t>(2 * nfx, 1) * 8) == hipSuccess &&
              hipMalloc((void**)&L->wy, std::max<size_t>(2 * nfy, 1) * 8) == hipSuccess &&
              hipMalloc((void**)&L->dP, N * 8) == hipSuccess && hipMalloc((void**)&L->Dz, N * 8) == hipSuccess &&
              hipMalloc((void**)&L->tmp, 2 * N * 8) == hipSuccess && hipMalloc((void**)&R, 3 * N * 8) == hipSuccess;
    if (!ok) {
        if (R) hipFree(R);
        orbit_fct_linop_destroy(L);
        return fail(ORBIT_ERR_NOMEM, "fct_linop_create_dev: device allocation failed");
    }
    hipMemcpyAsync(L->dP, fg.dP, N * 8, hipMemcpyDeviceToDevice, st);
    if (fg.conc && fg.Dz) hipMemcpyAsync(L->Dz, fg.Dz, N * 8, hipMemcpyDeviceToDevice, st);
    double* xb = R + 2 * N;
    hipLaunchKernelGGL(k_fct_x, g1(N), dim3(256), 0, st, fg, c_base_dev, 0LL, xb, 0LL, (int)N);
    hipLaunchKernelGGL(k_fct_R, cell_grid(fg, 1), dim3(256), 0, st, fg, (const double*)xb, (const double*)nullptr, dt,
                       R, R + N, 0LL, 0LL);
    if (nfx + nfy > 0)
        hipLaunchKernelGGL(k_fct_gains, g1(nfx + nfy), dim3(256), 0, st, fg, (const double*)xb, R, R + N, L->wx,
                           L->wy);
    hipError_t e = hipGetLastError();
    hipStreamSynchronize(st);
    hipFree(R);
    if (e != hipSuccess) {
        orbit_fct_linop_destroy(L);
        return fail(ORBIT_ERR_HIP, std::string("fct_linop_create_dev: ") + hipGetErrorString(e));
    }
    *out = L;
    return ORBIT_OK;
}
