/*
 * cmdg.h -- C ABI of the MI355X DG right-hand side and LSRK54 stepper for the
 * dry AtmosModel balance law (spectral-element DG, GLL nodes, hexahedra).
 *
 * Every entry point returns an int status (CMDG_OK == 0) and never throws.
 * Arrays use the reference's column-major storage:
 *   state       (Np, nstate, nelem)    node n = i + Nq*(j + Nq*k), 0-based i,j,k
 *   vgeo        (Np, nvgeo, nelem)     field order xi1x1,xi2x1,xi3x1,xi1x2,...,
 *                                      xi3x3, M, MI, MH, x1, x2, x3, JcV[, x_xi]
 *   sgeo        (5, Nfp, 6, nelem)     n1, n2, n3, sM, vMI
 *   vmap-/vmap+ (Nfp, 6, nelem) int64  1-based linear volume dof indices
 *   elemtobndy  (6, nelem) int64       0 = connected face, else boundary tag; every
 *                                      non-zero tag is the one boundary condition the
 *                                      library implements, the reference's default
 *                                      AtmosBC(Impenetrable(FreeSlip()), Insulating())
 *                                      (src/Atmos/Model/boundaryconditions.jl:43-110):
 *                                      a caller whose boundary-condition tuple holds
 *                                      other conditions cannot use this library
 *   D           (Nq, Nq)               D[j + Nq*k] = l_k'(xi_j)
 * Reference layout sources:
 *   src/Numerics/Mesh/Grids.jl:76-130 (vgeo/sgeo field ids),
 *   src/Numerics/Mesh/Grids.jl:559-637 (vmap-/vmap+ construction),
 *   src/Numerics/Mesh/Elements.jl:60-82 (D),
 *   src/Arrays/MPIStateArrays.jl:46-172 ((Np, nstate, nelem) storage),
 *   src/Atmos/Model/AtmosModel.jl:387-497 (variable order).
 *
 * Geometry, connectivity and the initial auxiliary state are host pointers read
 * once by cmdg_create (the context derives its device copy in FP64).  State,
 * tendency and dQ pointers passed to the stepping entry points are device
 * pointers (hipMalloc'd, caller-owned); work is enqueued on the context stream.
 */
#ifndef CMDG_H
#define CMDG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define CMDG_OK 0
#define CMDG_EINVAL 1
#define CMDG_EHIP 2
#define CMDG_ENOMEM 3
#define CMDG_EUNSUPPORTED 4
#define CMDG_ERANGE 5 /* a size beyond an entry point's fixed maximum */

/* orientation: src/Common/Orientations/Orientations.jl:138-146, 246-257 */
#define CMDG_NO_ORIENTATION 0
#define CMDG_FLAT_ORIENTATION 1
/* reference state: src/Atmos/Model/ref_state.jl:24, 36-60 */
#define CMDG_NO_REFERENCE_STATE 0
#define CMDG_HYDROSTATIC_STATE 1
/* turbulence closure: src/Common/TurbulenceClosures/TurbulenceClosures.jl */
#define CMDG_CONSTANT_DYNAMIC_VISCOSITY 0 /* :312-404 */
#define CMDG_CONSTANT_KINEMATIC_VISCOSITY 1
#define CMDG_SMAGORINSKY_LILLY 2 /* :426-499 */
#define CMDG_LEARNED_CLOSURE 3   /* per-node MLP eddy viscosity (no reference) */
/* first-order numerical flux: src/Numerics/DGMethods/NumericalFluxes.jl */
#define CMDG_RUSANOV 0 /* :219-285 */
#define CMDG_CENTRAL 1 /* :298-340 */
/* arithmetic of the tendency kernels */
#define CMDG_FP64 0
#define CMDG_MIXED 1 /* FP32 fluxes/derivatives, FP64 state/dQ/reference */

#define CMDG_NSTATE 5 /* rho, rhou1, rhou2, rhou3, rhoe */

typedef struct cmdg_desc {
    int32_t N;          /* polynomial order (same in all three directions) */
    int32_t nvgeo;      /* fields per node in vgeo (16 or 25) */
    int64_t nelem;
    int32_t naux;       /* auxiliary variables per node in the caller layout */
    /* positions of the auxiliary fields inside one node's aux vector (-1 = absent)
     * src/Atmos/Model/AtmosModel.jl:481-497 */
    int32_t aux_coord;      /* 3 consecutive: x1, x2, x3 */
    int32_t aux_Phi;        /* orientation.Phi */
    int32_t aux_gradPhi;    /* 3 consecutive: orientation.gradPhi */
    int32_t aux_ref_rho;    /* ref_state.rho */
    int32_t aux_ref_p;      /* ref_state.p */
    int32_t aux_Delta;      /* turbulence.Delta (Smagorinsky length scale) */
    int32_t aux_theta_v;    /* moisture.theta_v  (written by every RHS; by the last
                             * step of a cmdg_lsrk54_steps call) */
    int32_t aux_air_T;      /* moisture.air_T    (the same) */
    const double *vgeo;
    const double *sgeo;
    const int64_t *vmap_minus;
    const int64_t *vmap_plus;
    const int64_t *elemtobndy;
    const double *D;
    const double *omega;
    const double *aux_init; /* (Np, naux, nelem) initial auxiliary state */
    /* thermodynamic and model constants */
    double R_d, cp_d, cv_d, T_0, grav, MSLP;
    int32_t orientation;
    int32_t ref_state;
    int32_t subtract_off;   /* HydrostaticState(...; subtract_off) */
    int32_t gravity_source; /* source = (Gravity(),) */
    int32_t turbulence;
    int32_t with_divergence;/* ConstantViscosity WithDivergence() */
    double viscosity;       /* rho*nu (dynamic) or nu (kinematic) */
    double C_smag;
    double inv_Pr_turb;
    int32_t numerical_flux;
    int32_t precision;
} cmdg_desc;

typedef struct cmdg_ctx cmdg_ctx;

/* Build a DG context (the DGModel of src/Numerics/DGMethods/DGModel.jl:22-65):
 * copies geometry/connectivity to the device, takes ownership of a device copy of
 * the auxiliary state (DGModel.state_auxiliary) and allocates scratch (gradient
 * flux, ping-pong state).  `stream` is a hipStream_t (NULL = default stream). */
int cmdg_create(const cmdg_desc *desc, void *stream, cmdg_ctx **out);

/* An axis-aligned brick of ne[0] x ne[1] x ne[2] elements with break points x[a]
 * (ne[a] + 1 increasing values along axis a), stored with the vertical index fastest:
 * element e = k + ne[2] (i + ne[0] j) at lattice position (i, j, k), the order of the
 * reference's StackedBrickTopology (src/Numerics/Mesh/Topologies.jl:632-).  periodic[a]
 * joins the two ends of axis a; otherwise tag[a][0] / tag[a][1] is the boundary tag of
 * its lower / upper faces (non-zero: a wall, as any non-zero elemtobndy tag). */
typedef struct cmdg_brick {
    int64_t ne[3];
    const double *x[3];
    int32_t periodic[3];
    int64_t tag[3][2];
} cmdg_brick;

/* cmdg_create for a brick described by `brick` instead of per-node arrays (no
 * reference counterpart: a drop-in caller passes the grid's vgeo/sgeo/vmaps to
 * cmdg_create).  desc's vgeo, sgeo, vmap_minus, vmap_plus, elemtobndy and nvgeo are
 * ignored and may be null; D, omega and aux_init are read as by cmdg_create.  The
 * context is the one cmdg_create builds from the same brick's arrays (metric terms
 * 2/h, J = h1 h2 h3 / 8, the same element records, classes and traversal order), so
 * results are bitwise the same; it saves building and passing the 25 + 2 x 6 x 5 per-node
 * words of geometry and connectivity of a large mesh. */
int cmdg_create_brick(const cmdg_desc *desc, const cmdg_brick *brick, void *stream, cmdg_ctx **out);
int cmdg_destroy(cmdg_ctx *ctx);

/* tendency = alpha * dQ/dt(Q, t) + beta * tendency
 * replaces (dg::DGModel)(tendency, Q, _, t, alpha, beta),
 * src/Numerics/DGMethods/DGModel.jl:85-427; the auxiliary theta_v / air_T
 * are refreshed from Q as in DGModel.jl:110-116.  t is accepted for the
 * reference's signature; no flux, source or boundary condition the library
 * implements depends on time, so the tendency does not either. */
int cmdg_rhs(cmdg_ctx *ctx, double *tendency, const double *Q, double t,
             double alpha, double beta);

/* One LSRK54CarpenterKennedy step from t to t + dt:
 * replaces dostep!(Q, lsrk::LowStorageRungeKutta2N, p, t),
 * src/Numerics/ODESolvers/LowStorageRungeKuttaMethod.jl:102-144, with the
 * coefficients of :293-327.  dQ is the solver's low-storage register. */
int cmdg_lsrk54_step(cmdg_ctx *ctx, double *Q, double *dQ, double t, double dt);

/* nsteps consecutive LSRK54CarpenterKennedy steps from t with step dt: the loop of
 * solve!(Q, solver; numberofsteps) with no callbacks, src/Numerics/ODESolvers/
 * ODESolvers.jl:110-158 over dostep! (LowStorageRungeKuttaMethod.jl:102-144).
 * Identical to nsteps calls of cmdg_lsrk54_step (Q, dQ and the auxiliary theta_v and
 * air_T on return; a zero may differ in sign).  The stages alternate the state
 * between Q and a context buffer (each stage reads its neighbours' state while it
 * writes the next); the result is in Q on return, and a call with an odd nsteps
 * copies Q into the context buffer once at its start.  Between two steps of the call
 * the zeros the last stage leaves in dQ (the scheme's first coefficient, A_1 = 0,
 * scales it) are neither stored nor read, and only the last step writes theta_v and
 * air_T. */
int cmdg_lsrk54_steps(cmdg_ctx *ctx, double *Q, double *dQ, double t, double dt, int64_t nsteps);

/* One low-storage stage: dQ <- RHS(Q, t_stage) + dQ; Q += rkb*dt*dQ;
 * dQ *= rka_next.  (rhs! call + update! kernel,
 * LowStorageRungeKuttaMethod.jl:122-142, 146-158) */
int cmdg_lsrk_stage(cmdg_ctx *ctx, double *Q, double *dQ, double t_stage,
                    double dt, double rka_next, double rkb);

/* The bare update! kernel on n values (LowStorageRungeKuttaMethod.jl:146-158):
 * Q[i] += rkb*dt*dQ[i]; dQ[i] *= rka.  `stream` is a hipStream_t. */
int cmdg_lsrk_update(double *dQ, double *Q, int64_t n, double rka, double rkb,
                     double dt, void *stream);

/* Device pointer to the context-owned auxiliary state (Np, naux, nelem). */
int cmdg_aux(cmdg_ctx *ctx, double **aux_device);

/* Device pointer to the gradient flux of the last RHS (Np, ngradflux, nelem),
 * FP64; ngradflux is 9 (constant viscosity) or 10 (Smagorinsky). */
int cmdg_gradient_flux(cmdg_ctx *ctx, double **gradflux_device, int32_t *ngradflux);

/* Device pointer to the FP32 gradient flux of the last mixed-precision RHS
 * (Np, ngradflux, nelem), in stress form: the closure (constant viscosity,
 * Smagorinsky-Lilly or the learned nu_0) folded into what the second-order flux
 * needs.  Planes: tau, the viscous stress factor -2 nu_d S_d. (+ (2 nu / 3) tr S on
 * the diagonal with divergence); e_d = (nu_d / Pr_t) dh_tot/dx_d (3); the viscosity.
 * ngradflux 10 when the viscosity is isotropic (no buoyancy correction): tau as 6
 * symmetric planes xx xy xz yy yz zz, e at 6-8, nu at 9; 15 otherwise: tau as 9
 * row-major planes, e at 9-11, nu_d at 12-14.  Row d of the second-order flux is
 * (0, rho tau_d., tau_d . rho u - rho e_d) (tendencies_momentum.jl:37-42,
 * tendencies_energy.jl:27-56). */
int cmdg_gradient_flux_f32(cmdg_ctx *ctx, float **gradflux_device, int32_t *ngradflux);

/* ---------------------------------------------------------------------------
 * Learned closure (no reference counterpart; it replaces nu_0 of
 * SmagorinskyLilly, src/Common/TurbulenceClosures/TurbulenceClosures.jl:478-491).
 * A per-node MLP 10 -> 32 -> 32 -> 1 with SiLU hidden units.  Raw inputs: the
 * velocity gradient du_c/dx_d at index 3c + d and the filter width Delta at 9;
 * the network sees x = (raw - shift) * scale; its output y, a signed coefficient, gives
 *     nu_0 = max(nu_floor + clamp(y, -e^-1, e^-1) * Delta^2 * |S|, 0),   |S| = sqrt(2 S:S)
 * (the clamp keeps the network's part within 0.37 Delta^2 |S| either way, a
 * Smagorinsky coefficient of 0.6; a negative y takes viscosity away, backscatter, never
 * below zero in total); nu_floor is the descriptor's `viscosity` (the molecular
 * kinematic viscosity), 1e-5 when that is zero, so that a constant output C_s^2 with
 * viscosity 0 is Smagorinsky-Lilly.
 * Weights are BF16 (row-major [out][in]: w1 32x10, w2 32x32, w3 1x32), biases FP32
 * (b1 32, b2 32, b3 1), shift/scale 10 FP32 each; all host pointers.  Evaluated
 * with BF16 MFMA / FP32 accumulation inside the mixed-precision gradient kernel. */
int cmdg_mlp_set_weights(cmdg_ctx *ctx, const uint16_t *w1_bf16, const float *b1,
                         const uint16_t *w2_bf16, const float *b2,
                         const uint16_t *w3_bf16, const float *b3,
                         const float *in_shift, const float *in_scale);

/* The network alone on n raw input rows (device, row-major n x 10) -> y (device, n). */
int cmdg_mlp_forward(cmdg_ctx *ctx, const float *x_raw, float *y, int64_t n);

/* Velocity gradient du_c/dx_d of Q at (node, 3c + d, element), FP64 device
 * (Np, 9, nelem), formed as the gradient pass forms it (volume derivative plus
 * lifted central jump, wall rule of bc_momentum.jl:24-43).  Affine bricks. */
int cmdg_velocity_gradient(cmdg_ctx *ctx, const double *Q, double *grad_u);

/* 2:1 element-agglomeration filter (no reference counterpart): coarse element c is
 * the union of the 8 fine elements children[8c + hx + 2hy + 4hz] (h = 0: lower
 * half along that axis).  The coarse nodal values are the L2 projection of the
 * fine DG field, proj[(h*Nq + i)*Nq + a] the 1-D projection from fine node a of half
 * h to coarse node i (host pointer).  Filtered: rho, rho u (3), rho e and
 * rho u_a u_b (11,12,13,22,23,33) -> moments (Np, 11, ncoarse).  Device arrays. */
int cmdg_box_filter(const double *Q_fine, const int32_t *children, int64_t ncoarse, int32_t N,
                    const double *proj, double *moments, void *stream);

/* Training samples of the subgrid stress, one per coarse node, SoA with leading
 * dimension ld starting at column `offset`: rows 0-8 raw velocity gradient, row 9
 * Delta, rows 10-12 P, Q, R such that the squared deviatoric-stress error of the
 * closure is c^2 P + 2 c Q + R, c = clamp(y, -e^-1, e^-1) (the stress the network must supply is the
 * subgrid stress plus the molecular stress 2 (nu_mol - nu_floor) S^d that the
 * closure's floor nu_floor does not already give).  element_average != 0 replaces
 * P, Q, R by their mean over each element (locally averaged targets). */
int cmdg_closure_samples(const double *moments, const double *grad_u, int64_t nelem, int32_t N,
                         double Delta, double nu_mol, double nu_floor, int32_t element_average,
                         float *samples, int64_t ld, int64_t offset, void *stream);

/* MLP training, FP32 master parameters (CMDG_MLP_NPARAM, flat: W1[32][10], b1[32],
 * W2[32][32], b2[32], W3[32], b3).  Loss and gradient over a batch of samples
 * (rows idx[0..batch) of the sample table, or its first `batch` columns when idx is
 * NULL): grad[0..NPARAM) and grad[NPARAM] = loss = sum (c^2 P + 2 c Q + R) *
 * inv_norm.  work: device scratch of at least ceil(batch/64) * (NPARAM+1) floats. */
#define CMDG_MLP_NPARAM 1441
int cmdg_mlp_loss_grad(const float *params, const float *shift, const float *scale,
                       const float *samples, int64_t ld, const int64_t *idx, int64_t batch,
                       double inv_norm, float *grad, float *work, int64_t work_floats,
                       void *stream);

/* cmdg_mlp_loss_grad without the work argument: the same kernels and the same
 * result, bit for bit, with the per-workgroup partial gradients kept in a fixed
 * device array of the library (5.6 MiB per GPU, held until the process ends) instead
 * of caller-allocated scratch.  No reference counterpart: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc); it calls no hipMalloc, hipFree or allocating
 * library routine.  batch is at most CMDG_MLP_STATIC_MAX_BATCH (CMDG_ERANGE above
 * it).  Every call on one GPU shares that GPU's array: calls that overlap in time on
 * one GPU (other streams or threads) race and must be ordered by the caller.
 * Enqueues on `stream` (a hipStream_t of the device the array is taken from). */
#define CMDG_MLP_STATIC_MAX_BATCH 65536
int cmdg_mlp_loss_grad_static(const float *params, const float *shift, const float *scale,
                              const float *samples, int64_t ld, const int64_t *idx, int64_t batch,
                              double inv_norm, float *grad, void *stream);

/* One Adam step (bias-corrected, step >= 1) on n parameters. */
int cmdg_mlp_adam(float *params, const float *grad, float *m, float *v, int64_t n, float lr,
                  float beta1, float beta2, float eps, int64_t step, void *stream);

/* Velocity u = (rho u)/rho interpolated to p^3 equispaced points per element
 * (cell centres xi_m = -1 + (2m+1)/p; L[m*Nq + i] = l_i(xi_m), host) and written to
 * u[c][Z][Y][X] (c = 0..2; X = ix*p + mx, ...; dims nz, ny, nx points) from the
 * element's brick index eijk[3e..3e+2] = (ix, iy, iz): the uniform-grid input of
 * power_spectrum_3d (src/Common/Spectra/power_spectrum_les.jl:13-75), as
 * InterpolationBrick (src/Numerics/Mesh/Interpolation.jl:66-571) supplies it. */
int cmdg_interpolate_equispaced(const double *Q, const int32_t *eijk, int64_t nelem, int32_t N,
                                int32_t p, const double *L, int64_t nx, int64_t ny, int64_t nz,
                                double *u, void *stream);

/* Largest local Courant number of Q over all nodes (courant with the Atmos local
 * Courant functions: src/Numerics/DGMethods/SpaceDiscretization.jl:307-365,
 * DGModel_kernels.jl:3028-3095, src/Atmos/Model/courant.jl:11-73), with the
 * node spacing of kernel_min_neighbor_distance! (src/Numerics/Mesh/Grids.jl:1228-1334)
 * taken from the aux coordinates.  kind: 0 advective dt|u|/dx, 1 nondiffusive
 * dt(|u| + c)/dx, 2 diffusive dt nu/dx^2 (nu from the gradient flux of the last RHS,
 * formed as that RHS's kernels form it: FP64, or FP32 in mixed contexts, the learned nu_0
 * with the learned closure); direction: 0 EveryDirection, 1 HorizontalDirection,
 * 2 VerticalDirection.  Synchronises; *courant is a host double. */
int cmdg_courant(cmdg_ctx *ctx, const double *Q, double dt, int32_t kind, int32_t direction,
                 double *courant);

/* sum over all nodes of M * Q[:, state, :] (M the mass-matrix diagonal, vgeo field
 * M) accumulated in double-double: weightedsum(Q, state),
 * src/Arrays/MPIStateArrays.jl:655-674.  Synchronises; *sum is a host double. */
int cmdg_weightedsum(cmdg_ctx *ctx, const double *Q, int32_t state, double *sum);

/* Which state fields of Q (device, (Np, 5, nelem)) hold a NaN or Inf: bit s of *mask
 * is set when some Q[:, s, :] is not finite (0: all finite).  The map-reduce of
 * show_not_finite_fields (src/Arrays/MPIStateArrays.jl:884-903), done on the device;
 * a solver loop can call it every few steps to stop on a crash (the rhs and step
 * entry points themselves do not check).  Synchronises; *mask is a host int. */
int cmdg_not_finite_fields(cmdg_ctx *ctx, const double *Q, int32_t *mask);

/* Geometry representation the context chose (diagnostic; no reference counterpart),
 * a bit set: bit 0 = axis-aligned affine bricks (per-element metrics, detected bitwise
 * at create; clear = general per-node metrics); bit 1 = on those bricks, Phi and p_ref
 * agree bitwise on both sides of every interior face, so the face kernels take the
 * neighbour's values from their own side instead of its static face trace; bit 2 = on
 * those bricks, the static auxiliary fields (Phi, grad Phi, rho_ref, p_ref, Delta) of
 * many elements agree bitwise (every element layer of a brick whose reference state
 * depends on height alone), and the kernels read them from one shared block per class
 * instead of one per element; bit 3 = on those bricks, a pass walks the elements in
 * lattice tiles (z fastest, then x, then y, within and between tiles) so that a
 * block's face neighbours ran shortly before it on the same XCD: by default the
 * tendency pass in 8 x 8 x 8 tiles, and a mixed context whose face neighbours lie
 * farther apart in storage order than one XCD's L2 holds walks 12 x 12 element columns
 * spanning y in both passes.
 * Setting CMDG_GENERAL_GEOMETRY=1 in the environment before cmdg_create clears bit 0;
 * CMDG_NEIGHBOUR_STATIC_TRACE=1 clears bit 1; CMDG_STATIC_AUX_CLASSES=0 clears bit 2;
 * CMDG_ELEMENT_TILE=T sets cubic tiles of edge T (T = 0 clears bit 3) and
 * CMDG_ELEMENT_TILE=TX,TY,TZ the tile extents per axis (0: the whole axis);
 * CMDG_TILED_PASSES (tendency, gradient, both, none) the passes that walk the tiles. */
int cmdg_geometry_kind(cmdg_ctx *ctx, int32_t *kind);

/* Per-kernel device timing (diagnostic; no reference counterpart): when enabled,
 * every kernel the context launches is bracketed by HIP events on the context
 * stream.  cmdg_kernel_timing synchronises, returns the summed milliseconds and
 * launch counts per kind (0 gradient, 1 tendency, 2 update, 3 copy) and resets. */
int cmdg_set_kernel_timing(cmdg_ctx *ctx, int enable);
int cmdg_kernel_timing(cmdg_ctx *ctx, double *ms_total, int64_t *launches, int32_t nkinds);

int cmdg_sync(cmdg_ctx *ctx);
const char *cmdg_last_error(cmdg_ctx *ctx);
const char *cmdg_version(void);

#ifdef __cplusplus
}
#endif

#endif /* CMDG_H */
