// Device-side assembly of the shared transport block of one bin from its grid
// fields, straight into the device layout of an opset (or as a 7-point stencil
// for the caller).  Replaces the host CSR assembly of the reference:
//   assemble_transport_block      orbit/core/operator.py:281-353
//   assemble_horizontal_convdiff  orbit/core/convdiff.py:42-387 (Patankar exp.)
//   interface_omega, assemble_vertical_advection  orbit/core/advection.py:383-525
//   assemble_vertical_diffusion   orbit/core/mixing.py:14-117
//   assemble_cmfmc_transport      orbit/core/convection.py:25-140
//   density_weights               orbit/core/grid_data.py:543-553
//
// Every entry of a row of T depends only on the cell's own fields and those of
// its six neighbours, so one thread computes one row: the diagonal, the four
// same-level horizontal neighbours (W, E, S, N) and the two vertical ones
// (down = k-1, up = k+1), first in the pressure measure (face by face, in the
// order the reference accumulates them), then moved to the concentration form
// diag(rho) T diag(1/rho).  The opset path writes I + dtau*(in-column part) into
// the column blocks V, dtau*(horizontal part) into the stencil H and dtau*diag
// into the shared transport diagonal, all 16 padded levels per column, so no
// array needs clearing first; the per-tracer diagonals follow as
// d = dtau*T_nn + dtau*D_t,n.  Everything is FP64.
//
// A periodic longitude adds the wrap face between i = nx-1 (its low side) and
// i = 0 (convdiff.py:150-200) and drops the west / east outflow; its couplings
// sit in the W slot of i = 0 and the E slot of i = nx-1.

#include <cmath>

#include "orbit_internal.h"

// Setup code, not a hot path: no multiply-add contraction, so every entry is
// rounded as the host (and numpy) rounds it and d = dtau*T_nn + dtau*D matches
// the CSR builder's host arithmetic bit for bit.
#pragma clang fp contract(off)

using namespace orbit;

namespace {

constexpr double GRAV = 9.80665;             // grid_data.py:15
constexpr double EARTH_R = 6378137.0;        // grid_data.py:14
constexpr double DEG2RAD = 3.14159265358979323846 / 180.0;

struct AsmArgs {
    int nz, ny, nx;
    int split, conc;
    int periodic;             // periodic_lon: no west / east outflow
    int px;                   // the wrap face exists (periodic_lon and nx > 1)
    const double *Uw, *Upw, *Umw, *Kmw;  // wrap face wind, split winds, meander K; (nz, ny)
    double dy, trS, trN;
    const double* dx;
    const double *dP, *Dz, *U, *V, *Up, *Um, *Vp, *Vm, *Kh, *Kmu, *Kmv, *Kzz;
    const double *wpe, *wme;  // interface omega (nz+1 levels), or null
    const double *wpc, *wmc;  // cell-centred omega_plus / omega_minus, or null
    const double* cm;         // CMFMC (nz+1 levels), or null
    const double *Ue, *Upe, *Ume, *Vn, *Vpn, *Vmn;  // outer faces, or null
};

__host__ __device__ inline size_t cid(const AsmArgs& a, int k, int j, int i) {
    return ((size_t)k * a.ny + j) * a.nx + i;
}

__host__ __device__ inline double harm(double p, double q) { return 2.0 * p * q / (p + q + 1e-30); }

// Patankar rates of a face of width h: loss rate of the low-side cell (rL) and of
// the high-side cell (rR) (convdiff.py:15-39, 89-119).
__host__ __device__ inline void face_rates(double u, double Kf, double h, bool split, double up, double um,
                                           double& rL, double& rR) {
    const double sh = h > 0.0 ? h : 1.0;
    const double D = Kf / sh;
    const double sK = Kf > 1e-30 ? Kf : 1e-30;
    const double pe = fmin(fabs(u * sh / sK), 500.0);
    const double A = pe < 1e-6 ? 1.0 - 0.5 * pe : (pe > 500.0 ? 0.0 : pe / expm1(pe));
    const double diff = D * A / sh;
    if (split) {
        rL = diff + up / sh;
        rR = diff + um / sh;
    } else {
        rL = diff + fmax(u, 0.0) / sh;
        rR = diff + fmax(-u, 0.0) / sh;
    }
}

__host__ __device__ inline double wplus(const AsmArgs& a, int kk, int j, int i) {
    if (kk <= 0) return 0.0;  // the ground is closed
    if (a.wpe) return a.wpe[cid(a, kk, j, i)];
    return (a.wpc && kk < a.nz) ? a.wpc[cid(a, kk, j, i)] : 0.0;
}
__host__ __device__ inline double wminus(const AsmArgs& a, int kk, int j, int i) {
    if (kk <= 0) return 0.0;
    if (a.wme) return a.wme[cid(a, kk, j, i)];
    return (a.wmc && kk < a.nz) ? a.wmc[cid(a, kk, j, i)] : 0.0;
}

__host__ __device__ inline double rho_of(const AsmArgs& a, size_t n) {
    const double p = a.dP[n], z = a.Dz[n];
    return (p > 0.0 && z > 0.0) ? p / (GRAV * z) : 1.0;
}

// Row of T for cell (k, j, i): r = {diag, W, E, S, N, down, up}.
__host__ __device__ void transport_row(const AsmArgs& a, int k, int j, int i, double r[7]) {
    const int nz = a.nz, ny = a.ny, nx = a.nx;
    const size_t n = cid(a, k, j, i);
    const bool split = a.split != 0;
    const double dxj = a.dx[j];
    const bool vj = dxj > 0.0;
    const double sdx = vj ? dxj : 1.0;
    const double dPn = a.dP[n];
    double hd = 0.0, W = 0.0, E = 0.0, S = 0.0, Nn = 0.0;
    // ---- horizontal convection-diffusion, pressure measure ----
    if (nx > 1 && i > 0 && vj) {  // west face: this cell is the high side
        const size_t m = n - 1;
        double Kf = harm(a.Kh[m], a.Kh[n]);
        if (!split) Kf = Kf + a.Kmu[n];
        double rL, rR;
        face_rates(a.U[n], Kf, dxj, split, split ? a.Up[n] : 0.0, split ? a.Um[n] : 0.0, rL, rR);
        const double pL = a.dP[m];
        const bool ok = pL > 0.0 && dPn > 0.0;
        const double tr = ok ? (0.5 * (pL + dPn)) / dPn : 1.0;
        hd += rR * tr;
        W = -rL * tr;
    }
    if (nx > 1 && i < nx - 1 && vj) {  // east face: this cell is the low side
        const size_t m = n + 1;
        double Kf = harm(a.Kh[n], a.Kh[m]);
        if (!split) Kf = Kf + a.Kmu[m];
        double rL, rR;
        face_rates(a.U[m], Kf, dxj, split, split ? a.Up[m] : 0.0, split ? a.Um[m] : 0.0, rL, rR);
        const double pR = a.dP[m];
        const bool ok = dPn > 0.0 && pR > 0.0;
        const double tr = ok ? (0.5 * (dPn + pR)) / dPn : 1.0;
        hd += rL * tr;
        E = -rR * tr;
    }
    if (a.px && vj && (i == 0 || i == nx - 1)) {  // wrap face: low side i = nx-1, high side i = 0
        const size_t kj = (size_t)k * ny + j;
        const size_t nl = cid(a, k, j, nx - 1), nr = cid(a, k, j, 0);
        double Kf = harm(a.Kh[nl], a.Kh[nr]);
        if (!split) Kf = Kf + a.Kmw[kj];
        double rL, rR;
        face_rates(a.Uw[kj], Kf, dxj, split, split ? a.Upw[kj] : 0.0, split ? a.Umw[kj] : 0.0, rL, rR);
        const double pL = a.dP[nl], pR = a.dP[nr];
        const bool ok = pL > 0.0 && pR > 0.0;
        if (i == 0) {
            const double tr = ok ? (0.5 * (pL + pR)) / pR : 1.0;
            hd += rR * tr;
            W = -rL * tr;
        }
        if (i == nx - 1) {
            const double tr = ok ? (0.5 * (pL + pR)) / pL : 1.0;
            hd += rL * tr;
            E = -rR * tr;
        }
    }
    if (!a.periodic && i == 0 && vj) {  // west outflow
        if (split) {
            const double um = a.Um[n];
            if (um > 0.0) hd += um / sdx;
        } else {
            const double u = a.U[n];
            if (u < 0.0) hd += fabs(u) / sdx;
        }
    }
    if (!a.periodic && i == nx - 1 && vj) {  // east outflow through the true outer face
        const size_t kj = (size_t)k * ny + j;
        if (split) {
            const double up = a.Upe ? a.Upe[kj] : a.Up[n];
            if (up > 0.0) hd += up / sdx;
        } else {
            const double ue = nx > 1 ? (a.Ue ? a.Ue[kj] : a.U[n]) : a.U[cid(a, k, j, 0)];
            if (ue >= 0.0) hd += ue / sdx;
        }
    }
    const double sdy = a.dy > 0.0 ? a.dy : 1.0;
    if (ny > 1 && j > 0 && vj && a.dx[j - 1] > 0.0) {  // south face: this cell is the high side
        const size_t m = n - nx;
        double Kf = harm(a.Kh[m], a.Kh[n]);
        if (!split) Kf = Kf + a.Kmv[n];
        double rL, rR;
        face_rates(a.V[n], Kf, sdy, split, split ? a.Vp[n] : 0.0, split ? a.Vm[n] : 0.0, rL, rR);
        const double pS = a.dP[m];
        const bool ok = pS > 0.0 && dPn > 0.0;
        const double xf = 0.5 * (a.dx[j - 1] + dxj);
        const double tr = ok ? (xf / sdx) * ((0.5 * (pS + dPn)) / dPn) : 1.0;
        hd += rR * tr;
        S = -rL * tr;
    }
    if (ny > 1 && j < ny - 1 && vj && a.dx[j + 1] > 0.0) {  // north face: this cell is the low side
        const size_t m = n + nx;
        double Kf = harm(a.Kh[n], a.Kh[m]);
        if (!split) Kf = Kf + a.Kmv[m];
        double rL, rR;
        face_rates(a.V[m], Kf, sdy, split, split ? a.Vp[m] : 0.0, split ? a.Vm[m] : 0.0, rL, rR);
        const double pN = a.dP[m];
        const bool ok = dPn > 0.0 && pN > 0.0;
        const double xf = 0.5 * (dxj + a.dx[j + 1]);
        const double tr = ok ? (xf / sdx) * ((0.5 * (dPn + pN)) / dPn) : 1.0;
        hd += rL * tr;
        Nn = -rR * tr;
    }
    if (j == 0 && vj) {  // south outflow
        if (split) {
            const double vm = a.Vm[n];
            if (vm > 0.0) hd += vm / sdy * a.trS;
        } else {
            const double v = a.V[n];
            if (v < 0.0) hd += fabs(v) / sdy * a.trS;
        }
    }
    if (j == ny - 1 && vj) {  // north outflow through the true outer face
        const size_t ki = (size_t)k * nx + i;
        if (split) {
            const double vp = a.Vpn ? a.Vpn[ki] : a.Vp[n];
            if (vp > 0.0) hd += vp / sdy * a.trN;
        } else {
            const double vn = ny > 1 ? (a.Vn ? a.Vn[ki] : a.V[n]) : a.V[cid(a, k, 0, i)];
            if (vn >= 0.0) hd += vn / sdy * a.trN;
        }
    }
    // ---- vertical omega advection on the dP measure ----
    double vad = 0.0, vadn = 0.0, vaup = 0.0;
    if (dPn > 0.0) {
        if (k > 0) {
            const double wp = wplus(a, k, j, i), wm = wminus(a, k, j, i);
            if (wp > 0.0) vad = wp / dPn;
            if (wm > 0.0) vadn = -wm / dPn;
        }
        if (k < nz - 1) {
            const double wp = wplus(a, k + 1, j, i);
            if (wp > 0.0) vaup = -wp / dPn;
        }
        const double wm = wminus(a, k + 1, j, i);  // the domain top included
        if (wm > 0.0) vad = vad + wm / dPn;
    }
    // ---- vertical Kzz diffusion, divergence on the dP measure ----
    double vdd = 0.0, vddn = 0.0, vdup = 0.0;
    {
        const double z = a.Dz[n];
        const double rk = (z > 0.0 && dPn > 0.0) ? dPn / (GRAV * z) : 0.0;
        if (k > 0) {  // interface with the layer below: this cell is its upper side
            const size_t m = n - (size_t)ny * nx;
            const double zb = a.Dz[m], pb = a.dP[m];
            if (zb > 0.0 && z > 0.0) {
                const double Kf = 2.0 * a.Kzz[m] * a.Kzz[n] / (a.Kzz[m] + a.Kzz[n] + 1e-30);
                const double dist = 0.5 * (zb + z);
                const double rb = (zb > 0.0 && pb > 0.0) ? pb / (GRAV * zb) : 0.0;
                const double rf = 0.5 * (rb + rk);
                const double ratio = (rb > 0.0 && rk > 0.0) ? rf / rk : 1.0;
                const double c = Kf / (z * dist) * ratio;
                vddn = -c;
                vdd += c;
            }
        }
        if (k < nz - 1) {  // interface with the layer above: this cell is its lower side
            const size_t m = n + (size_t)ny * nx;
            const double za = a.Dz[m], pa = a.dP[m];
            if (z > 0.0 && za > 0.0) {
                const double Kf = 2.0 * a.Kzz[n] * a.Kzz[m] / (a.Kzz[n] + a.Kzz[m] + 1e-30);
                const double dist = 0.5 * (z + za);
                const double ra = (za > 0.0 && pa > 0.0) ? pa / (GRAV * za) : 0.0;
                const double rf = 0.5 * (rk + ra);
                const double ratio = (rk > 0.0 && ra > 0.0) ? rf / rk : 1.0;
                const double c = Kf / (z * dist) * ratio;
                vdup = -c;
                vdd += c;
            }
        }
    }
    // ---- CMFMC updraft and compensating subsidence ----
    double cvd = 0.0, cvdn = 0.0, cvup = 0.0;
    if (a.cm) {
        auto okc = [&](size_t q) { return a.Dz[q] > 0.0 && a.dP[q] > 0.0; };
        auto sDz = [&](size_t q) { return a.Dz[q] > 0.0 ? a.Dz[q] : 1.0; };
        auto srho = [&](size_t q) {
            const double r0 = a.dP[q] / (GRAV * sDz(q));
            return r0 > 0.0 ? r0 : 1.0;
        };
        if (k > 0) {  // updraft from below arrives here; subsidence leaves here
            const size_t m = n - (size_t)ny * nx;
            const double F = fmax(a.cm[n], 0.0);  // interface k = top of layer k-1
            if (okc(m) && F > 0.0 && okc(n)) {
                const double rcv = F / (srho(n) * sDz(n));
                cvdn = -rcv;
                cvd += rcv;
            }
        }
        const double F = fmax(a.cm[n + (size_t)ny * nx], 0.0);  // interface k+1
        if (okc(n) && F > 0.0) {
            const double src = F / (srho(n) * sDz(n));
            cvd += src;
            if (k < nz - 1 && okc(n + (size_t)ny * nx)) cvup = -src;
        }
    }
    r[0] = (hd + (vad + vdd)) + cvd;
    r[1] = W;
    r[2] = E;
    r[3] = S;
    r[4] = Nn;
    r[5] = (vadn + vddn) + cvdn;
    r[6] = (vaup + vdup) + cvup;
    if (a.conc) {  // diag(rho) T diag(1/rho), entry by entry
        const double rn = rho_of(a, n);
        r[0] = (rn * r[0]) * (1.0 / rn);
        if (i > 0 || a.px) r[1] = (rn * r[1]) * (1.0 / rho_of(a, i > 0 ? n - 1 : n + nx - 1));
        if (i < nx - 1 || a.px) r[2] = (rn * r[2]) * (1.0 / rho_of(a, i < nx - 1 ? n + 1 : n - (nx - 1)));
        if (j > 0) r[3] = (rn * r[3]) * (1.0 / rho_of(a, n - nx));
        if (j < ny - 1) r[4] = (rn * r[4]) * (1.0 / rho_of(a, n + nx));
        if (k > 0) r[5] = (rn * r[5]) * (1.0 / rho_of(a, n - (size_t)ny * nx));
        if (k < nz - 1) r[6] = (rn * r[6]) * (1.0 / rho_of(a, n + (size_t)ny * nx));
    }
}

// Stencil output, k-major: out[q*N + n], q = diag, W, E, S, N, down, up.
__global__ void k_assemble_stencil(AsmArgs a, double* __restrict__ out) {
    const size_t N = (size_t)a.nz * a.ny * a.nx;
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int i = (int)(n % a.nx), j = (int)((n / a.nx) % a.ny), k = (int)(n / ((size_t)a.nx * a.ny));
    double r[7];
    transport_row(a, k, j, i, r);
#pragma unroll
    for (int q = 0; q < 7; ++q) out[(size_t)q * N + n] = r[q];
}

// Opset layout: one thread per (column, padded level); writes the full row of
// V (16 values), the 4 stencil values of H and the transport diagonal.
__global__ void k_assemble_opset(AsmArgs a, double dtau, double* __restrict__ V, double* __restrict__ H,
                                 double* __restrict__ tdiag) {
    const int ncol = a.ny * a.nx;
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (size_t)ncol * 16) return;
    const int col = (int)(g >> 4), k = (int)(g & 15);
    double r[7] = {0, 0, 0, 0, 0, 0, 0};
    if (k < a.nz) transport_row(a, k, col / a.nx, col % a.nx, r);
    double* Vr = V + (size_t)col * 256 + k * 16;
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        double v = m == k ? 1.0 : 0.0;
        if (m == k - 1) v += dtau * r[5];
        if (m == k + 1) v += dtau * r[6];
        Vr[m] = v;
    }
    double* Hc = H + (size_t)col * 64 + k;
    Hc[0] = dtau * r[1];
    Hc[16] = dtau * r[2];
    Hc[32] = dtau * r[3];
    Hc[48] = dtau * r[4];
    tdiag[(size_t)col * 16 + k] = dtau * r[0];
}

// Lateral boundary inflow rate r_in >= 0 (1/s) per cell (reference
// compute_lateral_boundary_inflow, orbit/core/lateral_bc.py:58-92): the
// first-order upwind rate of the air entering through the cell's boundary
// faces -- the west face of i = 0 and the true east outer face of i = nx-1
// (none on a periodic longitude), the south face of j = 0 and the true north
// outer face of j = ny-1 times the outer faces' zonal width over the edge
// rows' -- the transport block's open-boundary outflow with the wind reversed,
// and no terrain ratio.  Accumulated west, east, south, north.
__global__ void k_lateral_inflow(AsmArgs a, double* __restrict__ out) {
    const size_t N = (size_t)a.nz * a.ny * a.nx;
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int i = (int)(n % a.nx), j = (int)((n / a.nx) % a.ny), k = (int)(n / ((size_t)a.nx * a.ny));
    const bool split = a.split != 0;
    const double dxj = a.dx[j];
    const bool vj = dxj > 0.0;
    const double sdx = vj ? dxj : 1.0;
    double r = 0.0;
    if (!a.periodic) {
        if (i == 0) {
            const size_t nw = cid(a, k, j, 0);
            const double u = split ? a.Up[nw] : fmax(a.U[nw], 0.0);
            if (vj && u > 0.0) r += u / sdx;
        }
        if (i == a.nx - 1) {
            const size_t kj = (size_t)k * a.ny + j, ne = cid(a, k, j, a.nx - 1);
            const double u = split ? (a.Ume ? a.Ume[kj] : a.Um[ne]) : fmax(-(a.Ue ? a.Ue[kj] : a.U[ne]), 0.0);
            if (vj && u > 0.0) r += u / sdx;
        }
    }
    const double sdy = a.dy > 0.0 ? a.dy : 1.0;
    if (j == 0) {
        const size_t ns = cid(a, k, 0, i);
        const double v = split ? a.Vp[ns] : fmax(a.V[ns], 0.0);
        if (vj && v > 0.0) r += v / sdy * a.trS;
    }
    if (j == a.ny - 1) {
        const size_t ki = (size_t)k * a.nx + i, nn = cid(a, k, a.ny - 1, i);
        const double v = split ? (a.Vmn ? a.Vmn[ki] : a.Vm[nn]) : fmax(-(a.Vn ? a.Vn[ki] : a.V[nn]), 0.0);
        if (vj && v > 0.0) r += v / sdy * a.trN;
    }
    out[n] = r;
}

// d[t] = tdiag + dtau * D[t] (D k-major, N per tracer), zero on padded levels.
__global__ void k_tracer_diag(const double* __restrict__ tdiag, const double* __restrict__ D, double dtau,
                              double* __restrict__ d, int ncol, int nz, int ntr) {
    const size_t n16 = (size_t)ncol * 16;
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n16 * ntr) return;
    const int t = (int)(g / n16);
    const size_t r = g - (size_t)t * n16;
    const int col = (int)(r >> 4), k = (int)(r & 15);
    d[g] = k < nz ? tdiag[r] + dtau * D[(size_t)t * nz * ncol + (size_t)k * ncol + col] : 0.0;
}

inline dim3 blocks_for(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

// Device copies of one bin's fields (one allocation).  The holder frees it
// after the stream has drained, so no copy or kernel still reads it.
struct FieldUpload {
    hipStream_t st = nullptr;
    void* buf = nullptr;
    AsmArgs a{};
    ~FieldUpload() {
        if (buf) {
            hipStreamSynchronize(st);
            hipFree(buf);
        }
    }
};

int check_fields(const orbit_grid_fields* g) {
    if (!g) return fail(ORBIT_ERR_INVALID, "assemble: NULL grid fields");
    if (g->nz < 1 || g->nz > NZP || g->ny < 1 || g->nx < 1)
        return fail(ORBIT_ERR_INVALID, "assemble: nz must be in [1, 16], ny and nx positive");
    if (g->periodic_lon && g->nx > 1) {
        if (!g->UAvg_wrap) return fail(ORBIT_ERR_INVALID, "assemble: a periodic longitude needs UAvg_wrap");
        if (g->split_fluxes && (!g->UAvg_plus_wrap || !g->UAvg_minus_wrap))
            return fail(ORBIT_ERR_INVALID, "assemble: split fluxes need UAvg_plus_wrap and UAvg_minus_wrap");
        if (!g->split_fluxes && !g->K_meander_u_wrap)
            return fail(ORBIT_ERR_INVALID, "assemble: net fluxes need K_meander_u_wrap");
    }
    if (!g->lat || !g->dx || !g->dP || !g->Dz || !g->UAvg || !g->VAvg || !g->Kxxyy || !g->Kzz)
        return fail(ORBIT_ERR_INVALID, "assemble: lat, dx, dP, Dz, UAvg, VAvg, Kxxyy and Kzz are required");
    if (g->split_fluxes && (!g->UAvg_plus || !g->UAvg_minus || !g->VAvg_plus || !g->VAvg_minus))
        return fail(ORBIT_ERR_INVALID, "assemble: split fluxes need UAvg_plus/minus and VAvg_plus/minus");
    if (!g->split_fluxes && (!g->K_meander_u || !g->K_meander_v))
        return fail(ORBIT_ERR_INVALID, "assemble: net fluxes need K_meander_u and K_meander_v");
    if ((g->omega_edge_plus == nullptr) != (g->omega_edge_minus == nullptr) ||
        (g->omega_plus == nullptr) != (g->omega_minus == nullptr))
        return fail(ORBIT_ERR_INVALID, "assemble: omega plus/minus must come in pairs");
    return ORBIT_OK;
}

int upload_fields(hipStream_t st, const orbit_grid_fields* g, FieldUpload& fu) {
    const size_t nz = g->nz, ny = g->ny, nx = g->nx, N = nz * ny * nx, N1 = (nz + 1) * ny * nx;
    struct Item {
        const double* src;
        size_t n;
        const double** dst;
    };
    AsmArgs& a = fu.a;
    fu.st = st;
    const bool wrap = g->periodic_lon && nx > 1;
    const Item items[] = {
        {g->dx, ny, &a.dx}, {g->dP, N, &a.dP}, {g->Dz, N, &a.Dz}, {g->UAvg, N, &a.U}, {g->VAvg, N, &a.V},
        {g->split_fluxes ? g->UAvg_plus : nullptr, N, &a.Up}, {g->split_fluxes ? g->UAvg_minus : nullptr, N, &a.Um},
        {g->split_fluxes ? g->VAvg_plus : nullptr, N, &a.Vp}, {g->split_fluxes ? g->VAvg_minus : nullptr, N, &a.Vm},
        {g->Kxxyy, N, &a.Kh}, {g->split_fluxes ? nullptr : g->K_meander_u, N, &a.Kmu},
        {g->split_fluxes ? nullptr : g->K_meander_v, N, &a.Kmv}, {g->Kzz, N, &a.Kzz},
        {g->omega_edge_plus, N1, &a.wpe}, {g->omega_edge_minus, N1, &a.wme},
        {g->omega_edge_plus ? nullptr : g->omega_plus, N, &a.wpc},
        {g->omega_edge_plus ? nullptr : g->omega_minus, N, &a.wmc},
        {g->CMFMC, N1, &a.cm},
        {g->U_face_e, nz * ny, &a.Ue}, {g->U_plus_face_e, nz * ny, &a.Upe}, {g->U_minus_face_e, nz * ny, &a.Ume},
        {g->V_face_n, nz * nx, &a.Vn}, {g->V_plus_face_n, nz * nx, &a.Vpn}, {g->V_minus_face_n, nz * nx, &a.Vmn},
        {wrap ? g->UAvg_wrap : nullptr, nz * ny, &a.Uw},
        {wrap && g->split_fluxes ? g->UAvg_plus_wrap : nullptr, nz * ny, &a.Upw},
        {wrap && g->split_fluxes ? g->UAvg_minus_wrap : nullptr, nz * ny, &a.Umw},
        {wrap && !g->split_fluxes ? g->K_meander_u_wrap : nullptr, nz * ny, &a.Kmw},
    };
    size_t total = 0;
    for (const Item& it : items)
        if (it.src) total += (it.n + 31) / 32 * 32;
    if (hipMalloc(&fu.buf, total * 8) != hipSuccess) return fail(ORBIT_ERR_NOMEM, "assemble: device allocation failed");
    double* p = (double*)fu.buf;
    for (const Item& it : items) {
        *it.dst = nullptr;
        if (!it.src) continue;
        ORBIT_HIP_CHECK(hipMemcpyAsync(p, it.src, it.n * 8, hipMemcpyHostToDevice, st));
        *it.dst = p;
        p += (it.n + 31) / 32 * 32;
    }
    a.nz = g->nz;
    a.ny = g->ny;
    a.nx = g->nx;
    a.split = g->split_fluxes ? 1 : 0;
    a.conc = g->concentration_state ? 1 : 0;
    a.periodic = g->periodic_lon ? 1 : 0;
    a.px = wrap ? 1 : 0;
    a.dy = g->dy;
    // zonal width of the outer south / north faces over the edge rows' width
    // (convdiff.py:335-341, 360-365)
    const double dx0 = g->dx[0] > 0.0 ? g->dx[0] : 1.0, dxn = g->dx[ny - 1] > 0.0 ? g->dx[ny - 1] : 1.0;
    a.trS = EARTH_R * std::cos((g->lat[0] - g->dlat / 2.0) * DEG2RAD) * g->dlon * DEG2RAD / dx0;
    a.trN = EARTH_R * std::cos((g->lat[ny - 1] + g->dlat / 2.0) * DEG2RAD) * g->dlon * DEG2RAD / dxn;
    return ORBIT_OK;
}

}  // namespace

extern "C" int orbit_transport_assemble(orbit_ctx* ctx, const orbit_grid_fields* g, double* stencil) {
    if (!ctx || !stencil) return fail(ORBIT_ERR_INVALID, "assemble: NULL argument");
    int rc = check_fields(g);
    if (rc) return rc;
    hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    FieldUpload fu;
    rc = upload_fields(st, g, fu);
    if (rc) return rc;
    const size_t N = (size_t)g->nz * g->ny * g->nx;
    double* out = nullptr;
    if (hipMalloc(&out, 7 * N * 8) != hipSuccess) return fail(ORBIT_ERR_NOMEM, "assemble: device allocation failed");
    hipLaunchKernelGGL(k_assemble_stencil, blocks_for(N), dim3(256), 0, st, fu.a, out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(stencil, out, 7 * N * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    hipFree(out);
    if (e != hipSuccess) return fail(ORBIT_ERR_HIP, std::string("assemble: ") + hipGetErrorString(e));
    return ORBIT_OK;
}

extern "C" int orbit_lateral_inflow(orbit_ctx* ctx, const orbit_grid_fields* g, double* rate) {
    if (!ctx || !rate) return fail(ORBIT_ERR_INVALID, "lateral_inflow: NULL argument");
    int rc = check_fields(g);
    if (rc) return rc;
    hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    FieldUpload fu;
    rc = upload_fields(st, g, fu);
    if (rc) return rc;
    const size_t N = (size_t)g->nz * g->ny * g->nx;
    double* out = nullptr;
    if (hipMalloc(&out, N * 8) != hipSuccess) return fail(ORBIT_ERR_NOMEM, "lateral_inflow: device allocation failed");
    hipLaunchKernelGGL(k_lateral_inflow, blocks_for(N), dim3(256), 0, st, fu.a, out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(rate, out, N * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    hipFree(out);
    if (e != hipSuccess) return fail(ORBIT_ERR_HIP, std::string("lateral_inflow: ") + hipGetErrorString(e));
    return ORBIT_OK;
}

extern "C" int orbit_opset_create_grids(orbit_ctx* ctx, int n_bins, int n_tracers, double dtau,
                                        const orbit_grid_fields* grids, const double* D, orbit_opset** out) {
    if (!ctx || !grids || !D || !out) return fail(ORBIT_ERR_INVALID, "create_grids: NULL argument");
    *out = nullptr;
    if (n_bins < 1 || n_tracers < 1 || n_tracers > 4096)
        return fail(ORBIT_ERR_INVALID, "create_grids: n_bins >= 1 and n_tracers in [1, 4096]");
    for (int b = 0; b < n_bins; ++b) {
        int rc = check_fields(&grids[b]);
        if (rc) return rc;
        if (grids[b].nz != grids[0].nz || grids[b].ny != grids[0].ny || grids[b].nx != grids[0].nx)
            return fail(ORBIT_ERR_INVALID, "create_grids: every bin needs the same grid dimensions");
        if ((grids[b].periodic_lon != 0) != (grids[0].periodic_lon != 0))
            return fail(ORBIT_ERR_INVALID, "create_grids: every bin needs the same periodic_lon");
    }
    const int wrap = grids[0].periodic_lon && grids[0].nx > 1;
    hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    const int nz = grids[0].nz, ny = grids[0].ny, nx = grids[0].nx, ncol = nx * ny, N = nz * ncol;
    auto* op = new orbit_opset();
    op->ctx = ctx;
    op->nz = nz; op->ny = ny; op->nx = nx; op->nbins = n_bins; op->ntr = n_tracers;
    op->ngroups = 1; op->ncol = ncol; op->n16 = ncol * 16; op->N = N; op->dtau = dtau;
    op->has_tdiag = true;
    op->px = wrap;
    int rc = alloc_bins(op);
    double* dD = nullptr;
    if (!rc && hipMalloc(&dD, (size_t)n_tracers * N * 8) != hipSuccess)
        rc = fail(ORBIT_ERR_NOMEM, "create_grids: device allocation failed");
    for (int b = 0; b < n_bins && !rc; ++b) {
        BinOps& B = op->bins[b];
        FieldUpload fu;
        rc = upload_fields(st, &grids[b], fu);
        if (rc) break;
        hipLaunchKernelGGL(k_assemble_opset, blocks_for((size_t)ncol * 16), dim3(256), 0, st, fu.a, dtau, B.V64,
                           B.H64, B.tdiag);
        if (hipMemcpyAsync(dD, D + (size_t)b * n_tracers * N, (size_t)n_tracers * N * 8, hipMemcpyHostToDevice,
                           st) != hipSuccess) {
            rc = fail(ORBIT_ERR_HIP, "create_grids: copy of D failed");
            break;
        }
        hipLaunchKernelGGL(k_tracer_diag, blocks_for((size_t)n_tracers * op->n16), dim3(256), 0, st, B.tdiag, dD,
                           dtau, B.d64, ncol, nz, n_tracers);
        rc = derive_bin(ctx, op, b);
        // the field upload and D are reused by the next bin: finish this one first
        if (hipStreamSynchronize(st) != hipSuccess && !rc) rc = fail(ORBIT_ERR_HIP, "create_grids: assembly failed");
    }
    hipStreamSynchronize(st);
    if (dD) hipFree(dD);
    if (!rc) rc = finalize_opset(ctx, op);
    if (rc) {
        orbit_opset_destroy(op);
        return rc;
    }
    *out = op;
    return ORBIT_OK;
}

// ------------------------------------------------------------------------- //
// per-species diagonals: deposition and chemistry losses
// ------------------------------------------------------------------------- //
// One thread per cell computes the diagonal of every species' operator
// L_s = T + D_s (reference assemble_species_operators, operator.py:430-583,
// with assemble_deposition, deposition.py:101-362): the wet rate on every
// level, the dry velocity over the surface layer thickness, the
// partition-weighted gas/particle blends of TotalNH, TotalNO3 and the VBS bins,
// the SOA photolytic sink, then the SO2/NOx/O3/CO chemistry losses and the VBS
// aging loss; and the diagonal coupling blocks (pSO4 <- SO2, TotalNO3 <- NOx,
// the four aging steps).  Species indices are the reference's (deposition.py:55-73).
namespace {

constexpr int NSP = 14;
constexpr int S_C100 = 0, S_PM25 = 1, S_NH = 2, S_SO2 = 3, S_NOX = 4, S_PSO4 = 5, S_NO3 = 6, S_O3 = 7, S_CO = 8,
              S_C10 = 9, S_C1 = 10, S_C01 = 11, S_C1000 = 12, S_POA = 13;
// VBS bins by position (low to high C*), and the aging steps source -> destination
__device__ __constant__ int VBS_BIN[5] = {S_C01, S_C1, S_C10, S_C100, S_C1000};
constexpr int N_COUPLE = 6;  // pSO4<-SO2, NO3<-NOx, C100<-C1000, C10<-C100, C1<-C10, C01<-C1

__device__ __forceinline__ int vbs_pos(int s) {
    return s == S_C01 ? 0 : s == S_C1 ? 1 : s == S_C10 ? 2 : s == S_C100 ? 3 : s == S_C1000 ? 4 : -1;
}

__global__ void k_species_diag(orbit_species_fields f, int nz, int ny, int nx, const double* __restrict__ Dz,
                               double* __restrict__ d, double* __restrict__ K) {
    const size_t N = (size_t)nz * ny * nx, plane = (size_t)ny * nx;
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int k = (int)(n / plane);
    const size_t c = n - (size_t)k * plane;  // surface index of the column
    const double pwd = f.particle_wet_dep[n], ogw = f.other_gas_wet_dep[n];
    const double pdd0 = f.particle_dry_dep[c];
    const double dz0 = Dz[c];
    const double p_nh = f.nh_part[n], p_nh0 = f.nh_part[c];
    const double p_no3 = f.no3_part ? f.no3_part[n] : 0.0, p_no30 = f.no3_part ? f.no3_part[c] : 0.0;
    const double hno3_vd0 = f.HNO3_dry_dep ? f.HNO3_dry_dep[c] : 0.03;
    double o3_vd0;
    if (f.O3_dry_dep) o3_vd0 = f.O3_dry_dep[c];
    else o3_vd0 = (f.is_land ? f.is_land[c] : 1.0) > 0.0 ? 0.004 : 0.0005;
    const double voc0 = f.VOC_dry_dep ? f.VOC_dry_dep[c] : 1.0e-3;
    double age[5] = {0, 0, 0, 0, 0};
    if (f.archive_OH) {
        const double roh = f.k_age * f.archive_OH[n];
        for (int b = 1; b < 5; ++b) {  // sources C1 .. C1000 (position 0, C01, does not age)
            if (f.F_p_vbs) {
                double F = f.F_p_vbs[(size_t)b * N + n];
                F = F < 0.0 ? 0.0 : (F > 1.0 ? 1.0 : F);
                age[b] = roh * (1.0 - F);
            } else {
                age[b] = roh;
            }
        }
    }
    for (int s = 0; s < NSP; ++s) {
        const int b = vbs_pos(s);
        double wet = 0.0, vd = 0.0;
        if (b >= 0) {
            wet = f.F_p_vbs ? f.F_p_vbs[(size_t)b * N + n] * pwd : pwd;
            if (f.F_p_vbs_surface) {
                const double F = f.F_p_vbs_surface[(size_t)b * plane + c];
                vd = F * pdd0 + (1.0 - F) * voc0;
            } else {
                vd = pdd0;
            }
        } else if (s == S_PM25 || s == S_POA || s == S_PSO4) {
            wet = pwd;
            vd = pdd0;
        } else if (s == S_NH) {
            wet = (1.0 - p_nh) * ogw + p_nh * pwd;
            vd = (1.0 - p_nh0) * f.NH3_dry_dep[c] + p_nh0 * pdd0;
        } else if (s == S_SO2) {
            wet = f.SO2_wet_dep[n];
            vd = f.SO2_dry_dep[c];
        } else if (s == S_NOX) {
            vd = f.NOx_dry_dep[c];
        } else if (s == S_NO3) {
            wet = (1.0 - p_no3) * ogw + p_no3 * pwd;
            vd = (1.0 - p_no30) * hno3_vd0 + p_no30 * pdd0;
        } else if (s == S_O3) {
            vd = o3_vd0;
        }
        double v = wet;
        if (k == 0) v += dz0 > 0.0 ? vd / dz0 : 0.0;
        if (b >= 0 && f.j_no2_soa && f.a_photo > 0.0)
            v = v + f.a_photo * f.j_no2_soa[n] * (f.F_p_vbs ? f.F_p_vbs[(size_t)b * N + n] : 1.0);
        if (s == S_SO2) v = v + f.so2_ox[n];
        if (s == S_NOX && f.nox_to_no3) v = v + f.nox_to_no3[n];
        if (s == S_O3 && f.o3_loss) v = v + f.o3_loss[n];
        if (s == S_CO && f.co_loss) v = v + f.co_loss[n];
        if (b >= 1 && f.archive_OH) v = v + age[b];
        d[(size_t)s * N + n] = v;
    }
    K[n] = -f.so2_ox[n];
    K[N + n] = f.nox_to_no3 ? -f.nox_to_no3[n] : 0.0;
    for (int q = 0; q < 4; ++q)  // C1000 -> C100, C100 -> C10, C10 -> C1, C1 -> C01
        K[(size_t)(2 + q) * N + n] = f.archive_OH ? -f.aging_gain * age[4 - q] : 0.0;
}

}  // namespace

namespace {
// ISORROPIA cross-coupling diagonals of the iso-coupled marginal mode
// (reference orbit/modes/iso_coupling.py:58-172): per cell the particle-minus-
// gas deposition contrast of TotalNH (wet: particle - other gas; dry at the
// surface: (particle - NH3) / Dz) and of TotalNO3 (the same wet contrast; dry:
// (particle - HNO3) / Dz, HNO3 at 3 cm/s when absent), times the cross
// partials of the partitioning: out[0] = dk_NH f_nh_dno3 (TotalNH <- TotalNO3),
// out[1] = dk_NH f_nh_dso4 (TotalNH <- pSO4), out[2] = dk_NO3 f_no3_dnh
// (TotalNO3 <- TotalNH), out[3] = dk_NO3 f_no3_dso4 (TotalNO3 <- pSO4).
struct IsoCrossArgs {
    const double* Dz;
    const double* particle_dry_dep;
    const double* NH3_dry_dep;
    const double* HNO3_dry_dep;  // nullable
    const double* particle_wet_dep;
    const double* other_gas_wet_dep;
    const double* f[4];  // f_nh_dno3, f_no3_dnh, f_nh_dso4, f_no3_dso4
    double* out[4];
    long long stride;    // between bins, for every array
    unsigned mask;       // bit b: block b on
};

__global__ void k_iso_cross(IsoCrossArgs a, int nz, int ny, int nx) {
    const size_t N = (size_t)nz * ny * nx, plane = (size_t)ny * nx;
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const size_t o = (size_t)blockIdx.y * a.stride;
    const int k = (int)(n / plane);
    const size_t c = n - (size_t)k * plane;
    const double dk_wet = a.particle_wet_dep[o + n] - a.other_gas_wet_dep[o + n];
    double dk_nh = dk_wet, dk_no3 = dk_wet;
    if (k == 0) {
        const double dz = a.Dz[o + c];
        const double pdd = a.particle_dry_dep[o + c];
        const double hno3 = a.HNO3_dry_dep ? a.HNO3_dry_dep[o + c] : 0.03;
        dk_nh = dk_nh + (dz > 0.0 ? (pdd - a.NH3_dry_dep[o + c]) / dz : 0.0);
        dk_no3 = dk_no3 + (dz > 0.0 ? (pdd - hno3) / dz : 0.0);
    }
    const double v0 = dk_nh * a.f[0][o + n], v1 = dk_nh * a.f[2][o + n];
    const double v2 = dk_no3 * a.f[1][o + n], v3 = dk_no3 * a.f[3][o + n];
    a.out[0][o + n] = (a.mask & 1u) ? v0 : 0.0;
    a.out[1][o + n] = (a.mask & 2u) ? v1 : 0.0;
    a.out[2][o + n] = (a.mask & 4u) ? v2 : 0.0;
    a.out[3][o + n] = (a.mask & 8u) ? v3 : 0.0;
}
}  // namespace

extern "C" int orbit_iso_cross_blocks_dev(orbit_ctx* ctx, int n_bins, int nz, int ny, int nx, long long stride,
                                          const double* Dz, const double* particle_dry_dep,
                                          const double* NH3_dry_dep, const double* HNO3_dry_dep,
                                          const double* particle_wet_dep, const double* other_gas_wet_dep,
                                          const double* const* partials, unsigned block_mask, double* const* out) {
    if (!ctx || !Dz || !particle_dry_dep || !NH3_dry_dep || !particle_wet_dep || !other_gas_wet_dep || !partials ||
        !out)
        return fail(ORBIT_ERR_INVALID, "iso_cross_blocks: NULL argument");
    if (n_bins < 1 || nz < 1 || ny < 1 || nx < 1)
        return fail(ORBIT_ERR_INVALID, "iso_cross_blocks: bad dimensions");
    const size_t N = (size_t)nz * ny * nx;
    if (n_bins > 1 && stride < (long long)N)
        return fail(ORBIT_ERR_INVALID, "iso_cross_blocks: the bin stride is shorter than a field");
    IsoCrossArgs a{};
    a.Dz = Dz;
    a.particle_dry_dep = particle_dry_dep;
    a.NH3_dry_dep = NH3_dry_dep;
    a.HNO3_dry_dep = HNO3_dry_dep;
    a.particle_wet_dep = particle_wet_dep;
    a.other_gas_wet_dep = other_gas_wet_dep;
    for (int q = 0; q < 4; ++q) {
        if (!partials[q] || !out[q]) return fail(ORBIT_ERR_INVALID, "iso_cross_blocks: NULL partial or output");
        a.f[q] = partials[q];
        a.out[q] = out[q];
    }
    a.stride = stride;
    a.mask = block_mask;
    hipSetDevice(ctx->device);
    hipLaunchKernelGGL(k_iso_cross, dim3((unsigned)((N + 255) / 256), n_bins), dim3(256), 0, ctx->stream, a, nz, ny,
                       nx);
    ORBIT_HIP_CHECK(hipGetLastError());
    return ORBIT_OK;
}

extern "C" int orbit_species_diagonals(orbit_ctx* ctx, int nz, int ny, int nx, const double* Dz,
                                       const orbit_species_fields* f, double* d_out, double* k_out) {
    if (!ctx || !Dz || !f || !d_out || !k_out) return fail(ORBIT_ERR_INVALID, "species_diagonals: NULL argument");
    if (nz < 1 || ny < 1 || nx < 1) return fail(ORBIT_ERR_INVALID, "species_diagonals: bad dimensions");
    if (!f->particle_dry_dep || !f->SO2_dry_dep || !f->NOx_dry_dep || !f->NH3_dry_dep || !f->particle_wet_dep ||
        !f->SO2_wet_dep || !f->other_gas_wet_dep || !f->nh_part || !f->so2_ox)
        return fail(ORBIT_ERR_INVALID, "species_diagonals: the dry/wet deposition fields, the TotalNH "
                                       "partitioning and the SO2 oxidation rate are required");
    hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    const size_t N = (size_t)nz * ny * nx, plane = (size_t)ny * nx;
    // device copies: (nz,ny,nx) fields, surface (ny,nx) fields, the 5-bin VBS fields
    struct Item {
        const double* src;
        size_t n;
        const double** dst;
    };
    orbit_species_fields g = *f;
    const Item items[] = {
        {f->particle_dry_dep, plane, &g.particle_dry_dep}, {f->SO2_dry_dep, plane, &g.SO2_dry_dep},
        {f->NOx_dry_dep, plane, &g.NOx_dry_dep}, {f->NH3_dry_dep, plane, &g.NH3_dry_dep},
        {f->VOC_dry_dep, plane, &g.VOC_dry_dep}, {f->HNO3_dry_dep, plane, &g.HNO3_dry_dep},
        {f->O3_dry_dep, plane, &g.O3_dry_dep}, {f->is_land, plane, &g.is_land},
        {f->particle_wet_dep, N, &g.particle_wet_dep}, {f->SO2_wet_dep, N, &g.SO2_wet_dep},
        {f->other_gas_wet_dep, N, &g.other_gas_wet_dep}, {f->nh_part, N, &g.nh_part},
        {f->no3_part, N, &g.no3_part}, {f->F_p_vbs, 5 * N, &g.F_p_vbs},
        {f->F_p_vbs_surface, 5 * plane, &g.F_p_vbs_surface}, {f->j_no2_soa, N, &g.j_no2_soa},
        {f->so2_ox, N, &g.so2_ox}, {f->nox_to_no3, N, &g.nox_to_no3}, {f->o3_loss, N, &g.o3_loss},
        {f->co_loss, N, &g.co_loss}, {f->archive_OH, N, &g.archive_OH}, {Dz, plane, nullptr},
    };
    size_t total = 0;
    for (const Item& it : items)
        if (it.src) total += (it.n + 31) / 32 * 32;
    const size_t out_n = (size_t)(NSP + N_COUPLE) * N;
    double* buf = nullptr;
    if (hipMalloc(&buf, (total + out_n) * 8) != hipSuccess)
        return fail(ORBIT_ERR_NOMEM, "species_diagonals: device allocation failed");
    double* p = buf;
    const double* dDz = nullptr;
    hipError_t e = hipSuccess;
    for (const Item& it : items) {
        if (!it.src) {
            if (it.dst) *it.dst = nullptr;
            continue;
        }
        if (e == hipSuccess) e = hipMemcpyAsync(p, it.src, it.n * 8, hipMemcpyHostToDevice, st);
        if (it.dst) *it.dst = p;
        else dDz = p;
        p += (it.n + 31) / 32 * 32;
    }
    double* dd = p;
    double* dk = p + (size_t)NSP * N;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_species_diag, blocks_for(N), dim3(256), 0, st, g, nz, ny, nx, dDz, dd, dk);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(d_out, dd, (size_t)NSP * N * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(k_out, dk, (size_t)N_COUPLE * N * 8, hipMemcpyDeviceToHost, st);
    hipError_t e2 = hipStreamSynchronize(st);
    hipFree(buf);
    if (e == hipSuccess) e = e2;
    if (e != hipSuccess) return fail(ORBIT_ERR_HIP, std::string("species_diagonals: ") + hipGetErrorString(e));
    return ORBIT_OK;
}
