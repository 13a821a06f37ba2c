/*
 * orbit_hip.h -- C ABI of the MI355X periodic-orbit transport solve (liborbit_hip.so).
 *
 * Plain C types only: pointers, sizes, status codes.  Every entry point
 * returns ORBIT_OK (0), ORBIT_WARN_NOT_CONVERGED (>0: finished but an iteration
 * cap was hit, results are returned) or a negative error; the message of the
 * last error on the calling thread is orbit_last_error().
 *
 * Vectors at the boundary are FP64, flat (N,) in k-major order
 * n = k*ny*nx + j*nx + i  (reference orbit/core/indexing.py:23-25).
 * Matrices are CSR with int32 indices (reference orbit/core/solve.py:136-137
 * passes int32 Ap/Ai to UMFPACK).
 *
 * Threading: an orbit_ctx owns one HIP stream and is not thread-safe; use one
 * context per host thread.  Opsets are owned by the context that made them.
 */
#ifndef ORBIT_HIP_H
#define ORBIT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORBIT_OK 0
#define ORBIT_WARN_NOT_CONVERGED 1
#define ORBIT_ERR_INVALID (-1)
#define ORBIT_ERR_HIP (-2)
#define ORBIT_ERR_STRUCTURE (-3)
#define ORBIT_ERR_NOMEM (-4)

#define ORBIT_PREC_FP64 0   /* plain: FP64 inner solves, FP64 GMRES */
#define ORBIT_PREC_MIXED 1  /* FP32 inner solves, BF16 column blocks, FP64 GMRES + refinement of
                               every bin state against its exact FP64 residual */

typedef struct orbit_ctx orbit_ctx;
typedef struct orbit_opset orbit_opset;
typedef struct orbit_fct orbit_fct;

typedef struct {
    double tol;          /* ||s - (I-M)c0|| <= tol*||s||  (reference atol = tol*||s||, orbit.py:177-182) */
    int maxiter;         /* GMRES restart cycles (reference maxiter, orbit.py:180) */
    int restart;         /* Krylov dimension per cycle (reference 30) */
    int precision;       /* ORBIT_PREC_FP64 or ORBIT_PREC_MIXED */
    int inner_sweeps;    /* red-black column sweeps per inexact P apply; 0 = opset default */
    int acc_sweeps;      /* sweeps per accurate FP64 P apply (plain FP64 path); 0 = opset default */
    int use_graph;       /* replay the inner sweeps of M from a captured HIP graph */
    int verbose;
    int max_rounds;      /* mixed: refinement rounds cap (0 = maxiter); a mixed round's GMRES
                            cycle takes at most min(restart, 16) steps */
    int rhs_is_s;        /* 1: the input is s itself ([nt][N]) for (I-M)c0 = s (or
                            (I-M^T)lam0 = s_adj); states 1..n_bins are then the
                            unforced propagation of c0 */
    int warm_orbit;      /* 1: c_warm is a whole orbit [nt][n_bins+1][N] (an output
                            of an earlier solve); mixed precision starts every bin
                            state there (FP64 takes its c0) */
    int check_sweeps;    /* mixed: sweeps per inner solve of the residual sweep (each
                            refinement round's check and right-hand side); 0 = opset default */
} orbit_opts;

typedef struct {
    int gmres_iters;     /* total Arnoldi steps (reference pr_norm callback count) */
    int rounds;          /* GMRES cycles (refinement rounds in mixed precision) */
    int converged;       /* 1 when ||r|| <= tol*||s|| */
    int zero_source;     /* 1 when ||s|| < 1e-30 (orbit is identically zero) */
    double residual;     /* ||(I-M)c0 - s|| / ||s||: FP64 sweeps to 1e-12 (plain FP64), or the
                            exact FP64 bin residuals propagated by the FP32 sweeps (mixed) */
    double periodicity;  /* ||c8 - c0|| / (||c0|| + 1e-30)  (orbit.py:199-200) */
    double neg_frac;     /* fraction of negative entries over the 9 states (orbit.py:202-203) */
    double s_norm;
} orbit_info;

typedef struct {
    int nz, ny, nx, n_bins, n_tracers, n_groups;
    int default_inner_sweeps, default_acc_sweeps;
    double dtau;
    double device_bytes;  /* HBM held by the opset (blocks, inverses, stencils) */
    double sweep_contraction;  /* measured error contraction of one red-black sweep (with margin) */
    int periodic_x;  /* 1 when the operators couple i = 0 and i = nx-1 (periodic longitude) */
    double relaxation;  /* omega of the sweeps' correction x += omega P (b - A x) */
    int mfma;           /* 1: BF16-inverse sweeps multiply the shared column block on the matrix cores */
    int default_check_sweeps;  /* sweeps per inner solve of the mixed refinement's residual sweep */
    int patch_shape;    /* orbit_opset_set_patch's shape of the first sweep (0: two half-sweep launches) */
} orbit_opset_stats;

/* Air-mass-balance statistics (reference orbit/core/operator.py:598-652). */
typedef struct {
    double interior_median, interior_p90, interior_p99, interior_max;
    double interior_frac_above_0_01, interior_frac_above_0_1;
    double top_median, top_p90;
    double layer_p90[32];
    double layer_mean_signed[32];
} orbit_mb_stats;

/* Grid metrics of one bin for the FCT source (reference GridData fields used
 * by orbit/core/fct.py:256-357).  All (nz,ny,nx) arrays are k-major; dx is (ny,).
 * With split_fluxes=0, u_plus/v_plus hold UAvg/VAvg and u_minus/v_minus are NULL. */
typedef struct {
    int nz, ny, nx;
    int split_fluxes;
    int concentration_state;  /* 1: transported variable is c/rho with rho = dP/(g Dz) */
    const double* u_plus;
    const double* u_minus;
    const double* v_plus;
    const double* v_minus;
    const double* dx;
    double dy;
    const double* dP;
    const double* Dz;
} orbit_grid_metrics;

const char* orbit_last_error(void);
void orbit_opts_default(orbit_opts* o);

int orbit_ctx_create(int device, orbit_ctx** out);
int orbit_ctx_destroy(orbit_ctx* ctx);
int orbit_ctx_sync(orbit_ctx* ctx);
/* The HIP stream the context launches on (hipStream_t as void*), for callers
 * that pass device pointers produced on another stream. */
void* orbit_ctx_stream(orbit_ctx* ctx);

/* Factor backend: replaces _factor_one_bin x 8 x species
 * (reference orbit/core/orbit.py:79-96, orbit/core/solve.py:210-244).
 * One CSR of L_{bin,t} per (bin, tracer), index [bin*n_tracers + t].  A is
 * built as I + dtau*L.  The pattern must couple each cell only to cells of its
 * own column (any levels) and to its 4 horizontal neighbours on its own level;
 * anything else is ORBIT_ERR_STRUCTURE; couplings across a periodic x wrap face
 * (i = 0 with i = nx - 1, any nx) are accepted.  nz <= 16. */
int orbit_opset_create_csr(orbit_ctx* ctx, int nz, int ny, int nx, int n_bins, int n_tracers,
                           double dtau, const int32_t* const* rowptr,
                           const int32_t* const* colidx, const double* const* vals,
                           orbit_opset** out);
/* Same, for the shared-transport form L_{bin,t} = T_bin + diag(D_{bin,t})
 * (reference orbit/core/operator.py:572-585): one CSR per bin and a dense
 * diagonal D laid out [bin][t][N]. */
int orbit_opset_create_shared(orbit_ctx* ctx, int nz, int ny, int nx, int n_bins, int n_tracers,
                              double dtau, const int32_t* const* rowptr,
                              const int32_t* const* colidx, const double* const* T_vals,
                              const double* D, orbit_opset** out);
int orbit_opset_destroy(orbit_opset* op);
int orbit_opset_stats_get(orbit_opset* op, orbit_opset_stats* st);
/* Relaxation of the red-black sweeps (the correction x += omega P (b - A x);
 * A x = b stays the fixed point, so omega changes the contraction and the
 * sweep counts, not the answer).  Opsets choose omega at creation by measuring
 * the contraction; omega = 0 chooses again, 1 is plain block Gauss-Seidel, any
 * value in [0.5, 1.5] is measured and used (ORBIT_ERR_INVALID, and the opset
 * unchanged, if its sweeps then do not contract).  No counterpart in the
 * reference (its direct factorisation has no iteration).  Not to be called
 * while a solve on the opset runs. */
int orbit_opset_set_relaxation(orbit_ctx* ctx, orbit_opset* op, double omega);
/* Block product of the reduced-precision sweeps (FP32 vectors and operator,
 * BF16 inverse blocks) of a shared-transport set: on = 1 multiplies the shared
 * 16x16 column block by the four slots of a wave on the matrix cores (v_mfma
 * 16x16x4 FP32); 0, the default, uses the VALU loop (measured 1-3 % faster per
 * headline step at 64x70x15 with 12 slots).  Either way a tracer's result does not depend on its batch; the
 * two settings differ in the last bits.  No counterpart in the reference.  Not
 * to be called while a solve on the opset runs. */
int orbit_opset_set_mfma(orbit_opset* op, int on);
/* The first sweep (from x = 0) of every FP32 inner solve of a set with one
 * shared transport and two colours: one launch over column patches of the
 * given shape (1: 4 x 4 columns, 8 waves a workgroup; 2: 4 x 2, 4 waves, the
 * default; 3: 8 x 2, 8 waves; 4: 8 x 4, 16 waves), each workgroup computing
 * the red columns of its patch and of the ring just outside it that its black
 * columns touch, then its black columns; or 0, two half-sweep launches.  The
 * same bits either way.  No counterpart in the reference.  Not to be called
 * while a solve on the opset runs. */
int orbit_opset_set_patch(orbit_opset* op, int shape);
/* Bin hand-off (on by default): in a chain of FP32 inner solves over the bins
 * (a monodromy apply, the residual sweep), each bin's last black half-sweep
 * runs inside the next bin's first-sweep launch, whose black columns finish
 * the previous bin's sweep at their own column before starting their own
 * (every red of the previous bin is final by then): one launch fewer per bin
 * boundary, the same bits.  Applies where orbit_opset_set_patch's one-launch
 * first sweep does (group solves included, when every set has it on), with
 * two or more sweeps per inner solve and at most 48 slots a launch (larger
 * launches are bound by bandwidth and measured slower with it).  No
 * counterpart in the reference.  Not to be called while a solve on the opset
 * runs. */
int orbit_opset_set_handoff(orbit_opset* op, int on);
/* Replace the per-tracer diagonals (D laid out [bin][t][N]) and refactor the
 * column blocks in place: the outer chemistry loop changes only loss rates. */
int orbit_opset_update_diag(orbit_ctx* ctx, orbit_opset* op, const double* D);

/* LU-protocol shim: x = A_{bin,t}^{-1} b (trans=0) or A^{-T} b (trans=1), FP64,
 * host buffers of N doubles.  Replaces _UmfpackLU.solve(b, trans)
 * (reference orbit/core/solve.py:246-277). */
int orbit_apply_P(orbit_ctx* ctx, orbit_opset* op, int bin, int tracer, int trans,
                  const double* b, double* x);

/* Periodic forward solve for nt tracers (indices tracers[0..nt)):
 * e[nt][n_bins][N] emission rates, c_warm[nt][N] or NULL, orbit_out[nt][n_bins+1][N],
 * info[nt].  Replaces solve_orbit_one_species (reference orbit/core/orbit.py:123-210)
 * for a batch of species sharing a month's operator set. */
int orbit_solve_forward(orbit_ctx* ctx, orbit_opset* op, const int* tracers, int nt,
                        const double* e, const double* c_warm, const orbit_opts* opts,
                        double* orbit_out, orbit_info* info);
/* Periodic adjoint solve (I - M^T) lam0 = s_adj: R[nt][n_bins][N] receptors,
 * lam_out[nt][n_bins+1][N].  Replaces solve_orbit_adjoint_one_species
 * (reference orbit/core/orbit_adjoint.py:70-144). */
int orbit_solve_adjoint(orbit_ctx* ctx, orbit_opset* op, const int* tracers, int nt,
                        const double* R, const double* lam_warm, const orbit_opts* opts,
                        double* lam_out, orbit_info* info);
/* Same two solves with device pointers on the context's stream (inputs already
 * resident in HBM).  info is a host array. */
int orbit_solve_forward_dev(orbit_ctx* ctx, orbit_opset* op, const int* tracers, int nt,
                            const double* e_dev, const double* c_warm_dev,
                            const orbit_opts* opts, double* orbit_out_dev, orbit_info* info);
int orbit_solve_adjoint_dev(orbit_ctx* ctx, orbit_opset* op, const int* tracers, int nt,
                            const double* R_dev, const double* lam_warm_dev,
                            const orbit_opts* opts, double* lam_out_dev, orbit_info* info);

/* The forward solve of orbit_solve_forward_dev with no device allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Its scratch (forcing, orbit, Krylov basis, work vectors,
 * GMRES state) is carved from a fixed device array of
 * ORBIT_FIXED_SCRATCH_BYTES per GPU, held from the library's load until the
 * process ends; a call that needs more (at 64x70x15 about 63 MB per tracer in
 * mixed precision, 32 MB in FP64 with restart 30: 12 tracers fit either way)
 * is refused with ORBIT_ERR_INVALID.  The
 * M~ sweeps are launched directly, never from a captured graph; results are
 * bit-identical to orbit_solve_forward_dev.  Every call on one GPU shares
 * that GPU's array (whatever the last call left in it; every buffer is
 * written before it is read): calls that overlap on one GPU race. */
#define ORBIT_FIXED_SCRATCH_BYTES (1024ull << 20)
int orbit_solve_forward_dev_noalloc(orbit_ctx* ctx, orbit_opset* op, const int* tracers, int nt,
                                    const double* e_dev, const double* c_warm_dev,
                                    const orbit_opts* opts, double* orbit_out_dev, orbit_info* info);

/* Periodic solves of slots from several operator sets (of any grid sizes) as
 * one batch: every sweep, vector and reduction launch covers all slots (the
 * uneven-orbit batch of many months and sub-domains; no counterpart in the
 * reference, whose species and months are solved one LU at a time).  Slot s
 * solves tracer tracers[s] of sets[slot_set[s]]: forward (adjoint = 0) for the
 * emissions in_dev[s] or adjoint (1) for the receptors in_dev[s], each a
 * device array (n_bins, N) of its set, into out_dev[s] (n_bins + 1, N), info[s]
 * as orbit_solve_forward_dev.  Cold starts only (rhs_is_s and warm_orbit
 * refused).  Every set needs one shared transport per bin and two sweep
 * colours (not a periodic longitude with an odd nx), the same n_bins and
 * dtau.  A slot's result is bit-identical to its set solved alone (with the
 * matrix-core block product when every set has it on, else the VALU loop). */
int orbit_solve_group_dev(orbit_ctx* ctx, int nsets, orbit_opset* const* sets, int nt, const int* slot_set,
                          const int* tracers, const double* const* in_dev, int adjoint, const orbit_opts* opts,
                          double* const* out_dev, orbit_info* info);

/* Grid fields of one bin for the device-side transport assembly (the GridData
 * fields read by orbit/core/operator.py:281-353 and the assemblers it calls).
 * (nz,ny,nx) arrays are k-major; NULL marks an absent optional field. */
typedef struct {
    int nz, ny, nx;
    int split_fluxes;          /* has_split_fluxes: UAvg/VAvg_plus/minus drive advection */
    int concentration_state;   /* 1: return diag(rho) T diag(1/rho), rho = dP/(g Dz) */
    int periodic_lon;          /* 1: the x direction wraps (face between i = nx-1 and i = 0) */
    double dy, dlon, dlat;     /* m, degrees, degrees */
    const double* lat;         /* (ny,) degrees, cell centres */
    const double* dx;          /* (ny,) m */
    const double* dP;          /* Pa */
    const double* Dz;          /* m */
    const double* UAvg;        /* west-face wind, m/s */
    const double* VAvg;        /* south-face wind, m/s */
    const double* UAvg_plus;   /* split fluxes only */
    const double* UAvg_minus;
    const double* VAvg_plus;
    const double* VAvg_minus;
    const double* Kxxyy;       /* m^2/s */
    const double* K_meander_u; /* net fluxes only (face-centred) */
    const double* K_meander_v;
    const double* Kzz;         /* m^2/s */
    const double* omega_edge_plus;   /* (nz+1,ny,nx) interface flux, downward, or NULL */
    const double* omega_edge_minus;  /* (nz+1,ny,nx) upward */
    const double* omega_plus;        /* (nz,ny,nx) cell-centred fallback, or NULL */
    const double* omega_minus;
    const double* CMFMC;             /* (nz+1,ny,nx) kg/m^2/s, or NULL */
    const double* U_face_e;          /* (nz,ny) true east outer face, or NULL */
    const double* U_plus_face_e;
    const double* U_minus_face_e;
    const double* V_face_n;          /* (nz,nx) true north outer face, or NULL */
    const double* V_plus_face_n;
    const double* V_minus_face_n;
    const double* UAvg_wrap;         /* (nz,ny) wind on the wrap face (periodic_lon, nx > 1) */
    const double* UAvg_plus_wrap;    /* split fluxes only */
    const double* UAvg_minus_wrap;
    const double* K_meander_u_wrap;  /* (nz,ny) net fluxes only */
} orbit_grid_fields;

/* Transport block of one bin assembled on the device from its grid fields
 * (reference assemble_transport_block, orbit/core/operator.py:281-347: the
 * Patankar horizontal convection-diffusion, omega advection, Kzz diffusion and
 * CMFMC convection, then the concentration form).  stencil (host, 7*N doubles)
 * receives, k-major, the diagonal and the W, E, S, N, down (k-1), up (k+1)
 * entries of every row (0 where the neighbour does not exist); with a periodic
 * longitude W of i = 0 and E of i = nx-1 are the couplings across the wrap
 * face (convdiff.py:150-200). */
int orbit_transport_assemble(orbit_ctx* ctx, const orbit_grid_fields* g, double* stencil);

/* Lateral boundary inflow rate r_in >= 0 (1/s) of every cell of one bin
 * (reference compute_lateral_boundary_inflow, orbit/core/lateral_bc.py:58-92):
 * the open boundary's outflow rule with the wind reversed (west / true east
 * outer face unless periodic_lon, south / true north outer face with the outer
 * faces' zonal width over the edge rows'), no terrain ratio.  rate: host, N
 * doubles, k-major.  The orbit source of a boundary concentration c_bc is
 * r_in * c_bc (lateral_bc.py:95-137). */
int orbit_lateral_inflow(orbit_ctx* ctx, const orbit_grid_fields* g, double* rate);

/* Opset of n_bins bins whose shared transport is assembled on the device from
 * grids[n_bins] and whose per-tracer diagonal D[n_bins][n_tracers][N] (host,
 * k-major) is added to it: L_{bin,t} = T_bin + diag(D_{bin,t})
 * (reference assemble_single_species_L, orbit/core/operator.py:356-368), with
 * no host CSR.  The result is the same opset orbit_opset_create_shared builds
 * from the CSR of T.  Any nx is accepted with a periodic longitude (with an
 * odd one the sweeps give the last column of each row two colours of its own). */
int orbit_opset_create_grids(orbit_ctx* ctx, int n_bins, int n_tracers, double dtau,
                             const orbit_grid_fields* grids, const double* D, orbit_opset** out);

/* Deposition and chemistry fields of one bin for the per-species diagonals
 * (the GridData fields read by orbit/core/deposition.py:101-362 and the rates
 * passed to assemble_species_operators, orbit/core/operator.py:371-595).
 * (nz,ny,nx) arrays k-major; the dry velocities are read on level 0 only;
 * is_land and F_p_vbs_surface are (ny,nx) / (5,ny,nx); F_p_vbs is (5,nz,ny,nx)
 * with the VBS bins ordered C01, C1, C10, C100, C1000.  nh_part / no3_part are
 * the partition fractions already chosen (marginal or equilibrium, with the
 * reference's fallbacks); NULL marks an absent optional field. */
typedef struct {
    const double* particle_dry_dep;
    const double* SO2_dry_dep;
    const double* NOx_dry_dep;
    const double* NH3_dry_dep;
    const double* VOC_dry_dep;    /* NULL: 0.1 cm/s */
    const double* HNO3_dry_dep;   /* NULL: 3 cm/s at the surface */
    const double* O3_dry_dep;     /* NULL: 0.4 cm/s over land, 0.05 over water */
    const double* is_land;        /* NULL: all land */
    const double* particle_wet_dep;
    const double* SO2_wet_dep;
    const double* other_gas_wet_dep;
    const double* nh_part;
    const double* no3_part;       /* NULL: 0 (TotalNO3 deposits as HNO3) */
    const double* F_p_vbs;        /* NULL: pure particle */
    const double* F_p_vbs_surface;
    const double* j_no2_soa;      /* NULL: no SOA photolytic sink */
    double a_photo;
    const double* so2_ox;         /* SO2 -> pSO4 rate (1/s) */
    const double* nox_to_no3;     /* NULL: no NOx -> TotalNO3 coupling */
    const double* o3_loss;
    const double* co_loss;
    const double* archive_OH;     /* NULL: no VBS aging */
    double k_age, aging_gain;
} orbit_species_fields;

/* The 14 species' diagonals d_out[14][N] of L_s = T + diag(d_s) and the
 * diagonal coupling blocks k_out[6][N] (pSO4 <- SO2, TotalNO3 <- NOx,
 * C100 <- C1000, C10 <- C100, C1 <- C10, C01 <- C1; zero when inactive),
 * computed on the device (reference assemble_species_operators,
 * orbit/core/operator.py:430-583).  Dz is (nz,ny,nx); outputs are host buffers. */
int orbit_species_diagonals(orbit_ctx* ctx, int nz, int ny, int nx, const double* Dz,
                            const orbit_species_fields* fields, double* d_out, double* k_out);

/* ISORROPIA cross-coupling blocks of the iso-coupled marginal mode (replaces
 * assemble_iso_cross_blocks, orbit/modes/iso_coupling.py:101-172, with its
 * _delta_k_nh / _delta_k_no3 at :58-98): for n_bins bins (every array
 * advances by `stride` doubles from one bin to the next; device pointers on
 * the context's stream; (nz,ny,nx) k-major, the dry velocities and Dz read
 * on level 0) the four diagonal K blocks, out[0] TotalNH <- TotalNO3 =
 * dk_NH * f_nh_dno3, out[1] TotalNH <- pSO4 = dk_NH * f_nh_dso4, out[2]
 * TotalNO3 <- TotalNH = dk_NO3 * f_no3_dnh, out[3] TotalNO3 <- pSO4 = dk_NO3
 * * f_no3_dso4, with partials = {f_nh_dno3, f_no3_dnh, f_nh_dso4, f_no3_dso4}
 * and dk = (particle - gas) wet rate + (particle - gas) dry velocity / Dz at
 * the surface (gas: NH3 for TotalNH, HNO3 for TotalNO3, 3 cm/s when
 * HNO3_dry_dep is NULL).  A block whose bit in block_mask is clear is zero
 * (the reference's ORBIT_ISO_DISABLE_BLOCKS diagnostic). */
int orbit_iso_cross_blocks_dev(orbit_ctx* ctx, int n_bins, int nz, int ny, int nx, long long stride,
                               const double* Dz, const double* particle_dry_dep, const double* NH3_dry_dep,
                               const double* HNO3_dry_dep, const double* particle_wet_dep,
                               const double* other_gas_wet_dep, const double* const* partials,
                               unsigned block_mask, double* const* out);

/* Regular-grid multilinear lookup table on the device: the 7-D ISORROPIA
 * partitioning table of the closure (reference IsorropiaLUT.query,
 * orbit/core/isorropia_lut.py:123-147, i.e. scipy RegularGridInterpolator
 * with method "linear", bounds_error=False, fill_value=None: linear
 * extrapolation outside the axes).  Axes strictly ascending, >= 2 values
 * each, <= 1024 values in all, ndim <= 8; fields are FP32 tables in C order
 * (host), nfields <= 8. */
typedef struct orbit_lut orbit_lut;
int orbit_lut_create(orbit_ctx* ctx, int ndim, const int* axis_len, const double* const* axes, int nfields,
                     const float* const* fields, orbit_lut** out);
/* out_dev[f][p] = clamp(interp_f(coords_dev[0][p], ..., coords_dev[ndim-1][p]),
 * clamp_lo[f], clamp_hi[f]) for p < npts (device arrays of doubles on the
 * context's stream; clamp arrays are host, nullable = no clamp). */
int orbit_lut_query_dev(orbit_ctx* ctx, orbit_lut* lut, long long npts, const double* const* coords_dev,
                        const double* clamp_lo, const double* clamp_hi, double* const* out_dev);
int orbit_lut_destroy(orbit_lut* lut);
/* ISORROPIA partitioning update on the device (reference partitioning_per_bin
 * and the blend of update_grid_partitioning, orbit/core/isorropia_closure.py:
 * 75-478) over a 7-axis table whose fields field_nh4 / field_no3 are f_nh4 and
 * f_no3.  Per bin b < n_bins and cell n < ncell, from the end-of-bin states
 * state_dev[0..2] = pSO4, TotalNH, TotalNO3 (clipped at 0; bin b at
 * + b*state_stride) and met_dev[0..3] = T (K), RH (percent), dust_fine and
 * sea_salt_fine (ug/m3, nullable: 0; bin b at + b*met_stride): out_dev[f]
 * (bin b at + b*ncell) receives f = f_nh_eq, f_nh_marg, f_no3_eq, f_no3_marg
 * and, with cross != 0, f_nh_dno3, f_no3_dnh, f_nh_dso4, f_no3_dso4; with
 * prev_dev (same layout) and 0 < alpha < 1 each output is
 * alpha*new + (1 - alpha)*prev.  Device arrays of doubles on the context's
 * stream; out must not overlap the inputs, except out[f] == prev[f]. */
int orbit_lut_partitioning_dev(orbit_ctx* ctx, orbit_lut* lut, int field_nh4, int field_no3, int n_bins,
                               long long ncell, const double* const* state_dev, long long state_stride,
                               const double* const* met_dev, long long met_stride, int cross, double alpha,
                               const double* const* prev_dev, double* const* out_dev);

/* Air-mass-balance pre-flight over the shared transport blocks of an opset
 * made with orbit_opset_create_shared (reference orbit/core/operator.py:598-652);
 * rho is (N,).  The caller applies the abort rule of orbit/cli.py:1604-1614. */
int orbit_mass_balance(orbit_ctx* ctx, orbit_opset* op, const double* rho, orbit_mb_stats* st);

/* Horizontal FCT anti-diffusive source d_AD = rho*(-div(C*A)) for one species
 * and bin (reference compute_horizontal_fct_source, orbit/core/fct.py:459-492).
 * c, ghost (nullable) and d_out are host buffers of N doubles (device buffers
 * for the _dev variant, where d_out must not overlap c or ghost). */
int orbit_fct_source(orbit_ctx* ctx, const orbit_grid_metrics* g, const double* c, double dt,
                     int limited, const double* ghost, double* d_out);
int orbit_fct_source_dev(orbit_ctx* ctx, const orbit_grid_metrics* g, const double* c_dev,
                         double dt, int limited, const double* ghost_dev, double* d_dev);

/* Frozen-coefficient FCT tangent-linear operator L_AD at a baseline state
 * (reference assemble_fct_linear_operator, orbit/core/fct.py:495-572), kept
 * as per-face gains on the device; apply y = L_AD x (trans=0) or L_AD^T x. */
int orbit_fct_linop_create(orbit_ctx* ctx, const orbit_grid_metrics* g, const double* c_base,
                           double dt, orbit_fct** out);
int orbit_fct_linop_apply(orbit_ctx* ctx, orbit_fct* L, int trans, const double* x, double* y);
int orbit_fct_linop_apply_dev(orbit_ctx* ctx, orbit_fct* L, int trans, const double* x_dev,
                              double* y_dev);
/* Dense export of the 5-point operator as CSR arrays sized by the caller
 * (N+1 row pointers, up to 5*N entries); returns the entry count in *nnz. */
int orbit_fct_linop_csr(orbit_ctx* ctx, orbit_fct* L, int32_t* rowptr, int32_t* colidx,
                        double* vals, int* nnz);
int orbit_fct_linop_destroy(orbit_fct* L);

/* Device-resident metrics of one bin for repeated FCT sources (the forward
 * deferred-correction passes of orbit/cli.py:1212-1241 evaluate the source for
 * every species and bin on every outer pass). */
typedef struct orbit_fct_grid orbit_fct_grid;
int orbit_fct_grid_create(orbit_ctx* ctx, const orbit_grid_metrics* g, orbit_fct_grid** out);
int orbit_fct_grid_destroy(orbit_fct_grid* G);
/* Batched compute_horizontal_fct_source (reference orbit/core/fct.py:459-492) on
 * nvec device vectors: vector v is c_dev + v*c_stride (ghost likewise, nullable);
 * out + v*out_stride receives base + v*base_stride + d_AD (base nullable: d_AD).
 * All pointers are device pointers on the context's stream; strides are >= 0.
 * The span of out ([out, out + (nvec-1)*out_stride + N)) must not overlap the
 * spans of c or ghost (ORBIT_ERR_INVALID), and base may overlap out only as out
 * itself with the same stride. */
int orbit_fct_source_batch_dev(orbit_ctx* ctx, orbit_fct_grid* G, int nvec, const double* c_dev,
                               long long c_stride, const double* ghost_dev, double dt, int limited,
                               const double* base_dev, long long base_stride, double* out_dev,
                               long long out_stride);

/* Forward FCT deferred correction on the device (reference orbit/cli.py:1212-1241,
 * the passes of forward mode): pass 0 solves the periodic orbits of e; pass p
 * = 1..n_passes-1 adds d_AD of the previous orbit's end-of-bin states (limited,
 * grids[bin] of the operator set's shape and context) to e and solves again:
 * FP64 from the previous c0; mixed precision starts every bin state from the
 * previous orbit, so its first residual is that of the forcing's change.  The
 * passes converge to the same tolerance either way.  e_dev (nt, n_bins, N) and
 * ghost_dev (nullable, the same layout: the concentration beyond the lateral
 * edge per bin, the lateral boundary values of orbit/core/lateral_bc.py) are
 * device arrays; ghost_tracers (nullable, host, nt ints): tracer s reads its
 * ghost rows only where ghost_tracers[s] != 0, the others keep closed edges
 * (the reference passes a ghost only for species with boundary values); null
 * means every tracer has one.  orbit_dev (nt, n_bins+1, N) receives the last
 * pass's orbits and info[p*nt + s] the solve information of pass p.  opts as
 * for orbit_solve_forward (rhs_is_s refused).  Returns the worst status of the
 * passes; the caller's stream is not synchronised with torch's. */
int orbit_forward_fct_dev(orbit_ctx* ctx, orbit_opset* op, orbit_fct_grid* const* grids, const int* tracers,
                          int nt, const double* e_dev, const double* ghost_dev, const int* ghost_tracers,
                          int n_passes, const orbit_opts* opts, double* orbit_dev, orbit_info* info);

/* The same passes, also writing the orbit every pass before the last ends
 * with into pass_orbits_dev (n_passes - 1, nt, n_bins + 1, N), device: the
 * baseline states of each pass's FCT source, which the backward pass of the
 * FCT orbit through the passes freezes its operators at.  Extends
 * orbit_forward_fct_dev (the reference keeps only the last orbit). */
int orbit_forward_fct_passes_dev(orbit_ctx* ctx, orbit_opset* op, orbit_fct_grid* const* grids,
                                 const int* tracers, int nt, const double* e_dev, const double* ghost_dev,
                                 const int* ghost_tracers, int n_passes, const orbit_opts* opts,
                                 double* orbit_dev, orbit_info* info, double* pass_orbits_dev);

/* Frozen L_AD from a device-resident grid and a device baseline state. */
int orbit_fct_linop_create_dev(orbit_ctx* ctx, orbit_fct_grid* G, const double* c_base_dev, double dt,
                               orbit_fct** out);

/* Species coupling block K (N x N CSR over k-major indices) on the device, for
 * the forward RHS term e_s -= K_(s,src) c_src[tau] (reference _solve_one_species,
 * orbit/core/orbit.py:263-307) and the adjoint receptor term
 * R_src -= dtau K_(r,src)^T (lam_r - R_r) (orbit/modes/adjoint.py:650-676). */
typedef struct orbit_spmat orbit_spmat;
int orbit_spmat_create(orbit_ctx* ctx, int n, const int32_t* rowptr, const int32_t* colidx,
                       const double* vals, orbit_spmat** out);
/* y += alpha * K x (trans = 0) or alpha * K^T x (trans = 1); x and y are device
 * vectors of n doubles on the context's stream and must not alias. */
int orbit_spmat_apply_dev(orbit_ctx* ctx, orbit_spmat* K, int trans, double alpha,
                          const double* x_dev, double* y_dev);
int orbit_spmat_destroy(orbit_spmat* K);

/* Health check of an opset's stored arrays: counts[14] receives the number of
 * non-finite entries, summed over bins, of V64, V32, VT64, VT32, H64, H32,
 * HT64, HT32, d64, d32, P64, P16, PT64, PT16 (no counterpart in the reference;
 * the reference's factor backend reports failures through UMFPACK status). */
int orbit_opset_check(orbit_ctx* ctx, orbit_opset* op, long long* counts);

/* Diagnostic with no counterpart in the reference: average device time (ms) of
 * one red-black half-sweep launch of `bin` on nslots tracer slots (tracers
 * 0..nslots-1), precision ORBIT_PREC_FP64 (FP64 blocks), ORBIT_PREC_MIXED (the
 * inner sweep of the mixed solve: FP32 vectors, BF16 blocks) or
 * ORBIT_SWEEP_ACC_MIXED (FP64 vectors and operator with BF16 blocks; no solve
 * runs this variant since the mixed path refines its states against exact FP64
 * residuals), measured with HIP events over `reps` back-to-back launches on the
 * context's stream.  Also returns the algorithmic bytes one launch moves (each
 * operand read once, the written colour once). */
#define ORBIT_SWEEP_ACC_MIXED 2
int orbit_bench_halfsweep(orbit_ctx* ctx, orbit_opset* op, int bin, int precision, int nslots,
                          int reps, double* ms_per_launch, double* bytes_per_launch);

/* Kernel timing of the half-sweeps (diagnostic, no counterpart in the
 * reference): between orbit_timing_start and orbit_timing_stop every
 * half-sweep the context launches records its own start and stop events
 * (hipExtLaunchKernelGGL; events made by orbit_timing_start for up to
 * max_launches launches, later ones run untimed), and solves launch their
 * inexact sweeps directly instead of replaying captured graphs.  Stop returns
 * per kernel variant (sorted by total time) the launch count, the sum of the
 * launches' own durations and of their algorithmic bytes (as
 * orbit_bench_halfsweep counts them, less what the variant does not read). */
typedef struct {
    int precision;       /* ORBIT_PREC_FP64, ORBIT_SWEEP_ACC_MIXED or ORBIT_PREC_MIXED (FP32 inner) */
    int slots_per_wave;  /* 1 or 4 */
    int mfma;            /* 1: the shared column block's product on the matrix cores */
    int own_zero;        /* 1: x of the colour known zero (not read) */
    int nbr_zero;        /* 1: x of the other colour known zero (not read) */
    long long launches;
    double total_ms;
    double total_bytes;
    int first_sweep;     /* 1: the first sweep from x = 0, both colours in one launch; 2: that launch
                            carrying the previous bin's last black half-sweep too (bin hand-off) */
} orbit_sweep_timing;
int orbit_timing_start(orbit_ctx* ctx, int max_launches);
int orbit_timing_stop(orbit_ctx* ctx, orbit_sweep_timing* out, int max_out, int* n_out, long long* untimed);

/* Launch shape of the red-black half-sweep (the column SpMV + block apply of
 * every inner solve): slots_per_wave 1 (a wave is 4 columns of one tracer slot)
 * or 4 (one column x 4 slots sharing the column's block loads), threads per
 * workgroup 64, 128 or 256 (the column tile), xcd_map 1 (each of the 8 XCDs
 * takes a contiguous range of column tiles) or 0 (linear block order).
 * Every shape computes the same per-lane arithmetic: results are bit-identical
 * whatever the shape.  Tables are [3][5]: rows FP64 sweeps, accurate mixed
 * sweeps (FP64 vectors, BF16 blocks), FP32 inner sweeps; columns slot buckets
 * 1, 2-3, 4-7, 8-15, 16+.  No counterpart in the reference (its factor backend
 * has no launch shape); this is the autotuner of the batched solves, run once
 * per sub-domain size class.  Not to be called while a solve on the opset runs. */
typedef struct {
    int slots_per_wave;
    int threads;
    int xcd_map;
} orbit_sweep_shape;
#define ORBIT_N_SWEEP_PRECS 3
#define ORBIT_N_SLOT_BUCKETS 5
/* Time every shape on bin 0 (slot counts 1, 2, 4, 8, 16; slot s runs tracer
 * s mod n_tracers; `reps` launches each, best of two rounds) and keep the
 * fastest per precision and bucket; the shape in place before the call stays
 * unless another is more than 3 % faster.  table / us_table (nullable) receive
 * the chosen shapes and their microseconds per launch. */
int orbit_opset_autotune(orbit_ctx* ctx, orbit_opset* op, int reps, orbit_sweep_shape* table,
                         double* us_table);
int orbit_opset_get_shapes(orbit_opset* op, orbit_sweep_shape* table);
/* Install a table (e.g. one tuned on another opset of the same size class). */
int orbit_opset_set_shapes(orbit_opset* op, const orbit_sweep_shape* table);

#ifdef __cplusplus
}
#endif

#endif /* ORBIT_HIP_H */
