/*
 * dsem_hip.h -- C ABI of the MI355X (gfx950) DSEM right-hand side and
 * low-storage RK4 step.
 *
 * The reference drives this path from Fortran with implicit interfaces and
 * derived types (`TYPE(domain) d(ngp)`, `TYPE(mortar) mrtr(nmp)`), neither of
 * which is C-interoperable, so the drop-in takes flat arrays:
 *
 *   reference                                            replaced by
 *   tstep(time,dt,d,ngrid,mrtr,nmort,resid)              dsem_rk_step
 *       src/tstep.f90:12 (called at src/3DDriver_parfmdf.f90:466)
 *   Fluxes(d,ngrid,mrtr,nmort,time) + update's           dsem_rhs
 *       divergence  src/Parabolic3D/stac3m/fluxesH_newa.f90:17,
 *                   src/stac3m/updatea.f90:100-113
 *   dtcalc(dt,d)  src/Parabolic3D/NS/NS/stac3m/dtcalc.f90:9 dsem_dtcalc
 *   set_points/set_matrices/set_geometry/set_mortar_nsign dsem_create
 *       src/Grid_Routines/stac3m_geomLib.f90:7,
 *       src/Spectral/stac3m/matrices.f90:20,
 *       src/Grid_Routines/set_domain_geometry.f90:8,
 *       src/Grid_Routines/set_mortars.f90:227
 *   d(id)%Q gather/scatter (Q stored as J*Q,             dsem_set_state / dsem_get_state
 *       src/3DDriver_parfmdf.f90:286-300)
 *   physics / rk_coefs module globals                    dsem_params
 *       src/paraFMDF/NS_data.f90:5-40, initializationsc2.f90:34-40
 *   Is_NaN/Is_INF check of resid                         DSEM_ERR_NONFINITE
 *       src/3DDriver_parfmdf.f90:489-499
 *
 * Conventions: every call returns 0 on success or a negative DSEM_ERR_*
 * code (message in dsem_last_error).  The context owns all device memory;
 * host buffers are only read or written during the call.  Calls are
 * synchronous with respect to the host unless named *_async.  One context
 * serves one device; a context is not thread-safe.
 *
 * State layout: J-weighted conserved variables (rho, rho u, rho v, rho w,
 * rho e) times J, FP64.
 *   layout 0  compact   [K][5][ncg^3], point n + ncg*(m + ncg*l)
 *   layout 1  reference [K][5][10][10][10], i.e. d(id)%Q(nx,ny,nz,neq) with
 *             nx=ny=nz=10 (caseFiles/tgv/size.f90), padding ignored on input
 *             and zero on output.
 */
#ifndef DSEM_HIP_H
#define DSEM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DSEM_OK               0
#define DSEM_ERR_ARG         -1
#define DSEM_ERR_HIP         -2
#define DSEM_ERR_NONFINITE   -3
#define DSEM_ERR_UNSUPPORTED -4
#define DSEM_ERR_NOMEM       -5

#define DSEM_PREC_FP64   0   /* plain FP64 drop-in                           */
#define DSEM_PREC_FP32   1   /* FP32 contractions, FP64 state/RK/reductions  */
#define DSEM_PREC_BF16X3 2   /* BF16 3-term split contractions, FP32 accum   */

/* Metric storage.  PER_POINT (the zero value, so the default) keeps the
 * reference's metrics at every point and is the parity mode.  PER_ELEMENT
 * keeps one value per element (parallelepiped elements only) and differs
 * from the reference by its point-to-point rounding noise (about 3e-12
 * relative L2 on the RHS at 22^3, p=8); AUTO picks PER_ELEMENT when every
 * element is affine.  Both are opt-in. */
#define DSEM_METRICS_PER_POINT   0  /* reference: metrics at every point (default) */
#define DSEM_METRICS_AUTO        1  /* per element when every element is affine    */
#define DSEM_METRICS_PER_ELEMENT 2  /* affine (parallelepiped) elements only       */

typedef struct dsem_ctx dsem_ctx;

/* Conforming hex8 mesh.  Mortar side 0 is the master face (the one whose
 * geometry defines the normal, as in set_domain_geometry.f90:254-345).
 * Faces use the reference numbering 1..6 (1: eta=0, 2: eta=1, 3: zeta=0,
 * 4: xi=1, 5: zeta=1, 6: xi=0).  mortar_orient is the orientation of the
 * slave face against the mortar as faces_match returns it
 * (src/Grid_Routines/face_operations.f90:404-460): DFLT, B1F2, B1B2, F1B2,
 * F2F1, B2F1, B2B1, F2B1 = 0..7 (element_library.f90:99-101); mortar point
 * (i,j) takes the slave face point that rotate_to_mortar
 * (src/Parabolic3D/mortar_operations.f90:461-605) maps to it. */
typedef struct {
    int32_t K;                     /* elements                               */
    int32_t p;                     /* order: ncg = p+1 Gauss, nc = p+2 Lobatto, 1..8 */
    int32_t nmortar;
    int32_t reserved;
    const int32_t *mortar_elem;    /* [nmortar][2] 0-based element ids       */
    const int32_t *mortar_face;    /* [nmortar][2] 1..6                       */
    const int32_t *mortar_sign;    /* [nmortar] +-1 (slave flux sign) or NULL */
    const int32_t *mortar_nsign;   /* [nmortar] +-1 (normal sign) or NULL     */
    const int32_t *mortar_orient;  /* [nmortar] 0..7, or NULL (all 0)         */
    const double  *corners;        /* [K][8][3] hex8 corners, FE order        */
} dsem_mesh_desc;

#define DSEM_EV_DUCROS    1   /* multiply mu_h by the Ducros sensor ('ducrosSens') */
#define DSEM_EV_SMOOTHING 2   /* smooth mu_h in each element ('Smoothing')         */
#define DSEM_EV_CMAX_CAP  4   /* cap mu_h by cmax*max(rho)*h*max(|u|+c) per element
                                 instead of mumax (the reference's commented-out
                                 variant, EViscosity_GL.f90:82)                  */

typedef struct {
    double  gamma, re, pr, mach;   /* physics module (mach = 1 with T0=100 gives Ma 0.1) */
    int32_t shock;                 /* entropy viscosity, once per step (input keyword 'shock') */
    int32_t precision;             /* DSEM_PREC_*                             */
    int32_t metrics;               /* DSEM_METRICS_*                          */
    int32_t ev_flags;              /* DSEM_EV_*                               */
    double  cmu, mumax, cmax;      /* entropy-viscosity coefficients (physics keywords Cmu, mumax) */
} dsem_params;

int  dsem_create(const dsem_mesh_desc *mesh, const dsem_params *params, int device, dsem_ctx **out);
int  dsem_destroy(dsem_ctx *ctx);
const char *dsem_last_error(const dsem_ctx *ctx);

/* Problem sizes: K, p, nmortar, degrees of freedom (5*K*ncg^3). */
int  dsem_sizes(const dsem_ctx *ctx, int32_t *K, int32_t *p, int32_t *nmortar, int64_t *dof);

/* Gauss-point coordinates [K][ncg^3][3] and jacob (=1/J) [K][ncg^3]
 * (set_domain_geometry.f90:86-97); either pointer may be NULL. */
int  dsem_gauss_points(dsem_ctx *ctx, double *xg, double *jacob);

/* Chebyshev-Lobatto cx [nc] and Chebyshev-Gauss cxg [ncg] points on [0,1]
 * (set_points, src/Grid_Routines/stac3m_geomLib.f90:7-47); either may be NULL. */
int  dsem_points(const dsem_ctx *ctx, double *cx, double *cxg);

/* The Gauss derivative and Gauss->Lobatto interpolation operators of
 * dermat / intrpmat (src/Spectral/stac3m/matrices.f90:144-207):
 * dmx [ncg][nc] with dmx[k][j] = dmx(k+1,j+1), bx [nc][ncg] with
 * bx[j][k] = bx(j+1,k+1); either may be NULL. */
int  dsem_operators(const dsem_ctx *ctx, double *dmx, double *bx);

/* Per-point metrics of the context's mesh as set_geometry leaves them in
 * d%gmetg, d%gmet and the mortars (src/Grid_Routines/set_domain_geometry.f90:8-348),
 * whatever the DSEM_METRICS mode; any pointer may be NULL.
 *   gmetg [K][ncg^3][9]: entry 3i+j = gmetg(i+1,j+1) at the Gauss points;
 *   gmet  [K][3][nc*ncg^2][3]: gmet(d+1,c+1) on staggered grid d;
 *   nhat  [nmortar][ncg^2][3], norm [nmortar][ncg^2] (mrtr%n_hat, mrtr%norm). */
int  dsem_get_metrics(dsem_ctx *ctx, double *gmetg, double *gmet, double *nhat, double *norm);

int  dsem_set_state(dsem_ctx *ctx, const double *JQ, int layout);
int  dsem_get_state(dsem_ctx *ctx, double *JQ, int layout);

/* Entropy viscosity mu_h on the three staggered grids, [K][3][nc*ncg^2]
 * (muhlgg/muhglg/muhggl of Vis_Fluxes_new.f90:126-190), used when shock=1.
 * dsem_rk_step recomputes it at the start of every step, as tstep calls
 * Entropy (src/tstep.f90:36-38); dsem_rhs uses the current field. */
int  dsem_set_viscosity(dsem_ctx *ctx, const double *muh);
int  dsem_get_viscosity(dsem_ctx *ctx, double *muhg, double *muh);

/* Entropy (src/EV_routines/Main/Driver.f90:21): mu_h from the current state
 * and the lagged staggered states (Qlgg/Qglg/Qggl after face averaging,
 * [K][3][5][nc*ncg^2]) that the last RK stage left behind.  Optional host
 * outputs at the Gauss points [K][ncg^3]: mu_h, the residual D and the
 * Ducros sensor.  Without lagged states (fresh start) mu_h is zero. */
int  dsem_ev_update(dsem_ctx *ctx, double *muhg, double *resid, double *ducros);
int  dsem_set_lagged_states(dsem_ctx *ctx, const double *stag);
int  dsem_get_lagged_states(dsem_ctx *ctx, double *stag);

/* Qt = d(JQ)/dt for the current state (Fluxes + update's divergence);
 * Qt_out [K][5][ncg^3] compact, or NULL to keep it on the device. */
int  dsem_rhs(dsem_ctx *ctx, double t, double *Qt_out);

/* One Carpenter-Kennedy RK4(5) 2N step (tstep).  resid_out (nullable)
 * receives max|Qt| over the last element in the last stage, as the
 * reference's resid; DSEM_ERR_NONFINITE is returned when any stage produced
 * a NaN or Inf anywhere (checked only when resid_out is non-NULL). */
int  dsem_rk_step(dsem_ctx *ctx, double t, double dt, double *resid_out);

/* Several steps back to back without host synchronisation in between. */
int  dsem_rk_steps(dsem_ctx *ctx, double t, double dt, int nsteps, double *resid_out);

/* Inviscid time-step bound before the CFL factor (dtcalc over all elements). */
int  dsem_dtcalc(dsem_ctx *ctx, double *dt_out);

/* FP64 flow diagnostics on the device.
 * Ek  (nullable): turbk of the TGV statistics (do_user_stuff,
 *     caseFiles/tgv/BackstepProblem-stoch.f90:629-679):
 *     1/2 sum (u^2+v^2+w^2) bmatu1b / jacob_1(1,1,1) / (8 pi^3), u = (JQ)_2/(JQ)_1;
 *     the driver's -dEk/dt is (Ek_previous - Ek)/dt (:692).
 * eps (nullable): energydissipation (src/Energydissipation/energydissipation.f90:2-89):
 *     100 * sum(|curl u|^2/jacob) / sum(1/jacob) / Re over all Gauss points,
 *     the velocity gradients taken from the lagged staggered states of the
 *     last RK stage (as the driver calls it right after tstep,
 *     src/3DDriver_parfmdf.f90:466-471).  Needs a step taken with shock = 1
 *     or after dsem_keep_lagged_states(ctx, 1).
 * Sums run per element, then over elements in a fixed order. */
int  dsem_reduce_ke(dsem_ctx *ctx, double *Ek, double *eps);

/* Adjoint (reverse mode) of the RHS and of the RK step, FP64, shock = 0,
 * Smagorinsky model off.
 * No counterpart in the reference, which has no derivative code: these
 * differentiate the same discrete operators as dsem_rhs / dsem_rk_step
 * (tstep, src/tstep.f90:12-120), with nu = 1/Re as a parameter (the RHS is
 * linear in it).
 * dsem_rhs_vjp: JQbar = (dR/dJQ)^T lam at the context's state, and
 *   nubar = <lam, dR/dnu>; host buffers [K][5][ncg^3].
 * dsem_rhs_vjp_device: the same at a device state dJQ, device buffers.
 * dsem_rk_step_vjp: adjoint of one dsem_rk_step from the context's state
 *   (the step's start state, left unchanged): Qbar holds the adjoint of the
 *   state after the step on entry and of the state before it on return;
 *   nubar receives this step's contribution.  The five stage states are
 *   recomputed from the start state.  _device: dQbar on the device. */
int  dsem_rhs_vjp(dsem_ctx *ctx, const double *lam, double *JQbar, double *nubar);
int  dsem_rhs_vjp_device(dsem_ctx *ctx, const double *dJQ, const double *dLam, double *dJQbar, double *nubar);
int  dsem_rk_step_vjp(dsem_ctx *ctx, double dt, double *Qbar, double *nubar);
int  dsem_rk_step_vjp_device(dsem_ctx *ctx, double dt, double *dQbar, double *nubar);

/* Step tape for checkpointed adjoints (no reference counterpart).
 * dsem_tape_reserve allocates nslots slots, each holding one step's five
 * stage input states and the forward face intermediates the adjoint reads
 * (5 * (K*5*ncg^3 + 2*K*30*ncg^2 + nmortar*5*ncg^2) doubles per slot).
 * dsem_rk_step_record takes one step of the context's state (advancing it,
 * as dsem_rk_step) and records it in `slot`; dsem_rk_step_vjp_tape is the
 * adjoint of the step recorded there (dQbar on the device, in/out). */
int  dsem_tape_reserve(dsem_ctx *ctx, int nslots);
int  dsem_rk_step_record(dsem_ctx *ctx, double dt, int slot);
int  dsem_rk_step_vjp_tape(dsem_ctx *ctx, double dt, int slot, double *dQbar, double *nubar);

/* Adjoint with entropy viscosity (shock = 1).  No counterpart in the
 * reference.  tstep computes mu_h once per step from the step's start state
 * and the lagged staggered states the previous step's last stage left
 * (Entropy, src/EV_routines/Main/Driver.f90:21-60; Entropy_Viscosity,
 * EViscosity_GL.f90:19-101; RHS_ent_trans, RHS.f90:21-190), so the state of
 * a shock run is the pair (J*Q, lagged states) and its adjoint the pair
 * (Qbar, Slagbar).
 * dsem_ev_vjp: at the context's state and lagged states, JQbar =
 *   (d mu_h / d JQ)^T muhbar and Slagbar = (d mu_h / d Slag)^T muhbar; host
 *   buffers muhbar [K][3][nc*ncg^2], JQbar [K][5][ncg^3], Slagbar
 *   [K][3][5][nc*ncg^2].  The clamps take the branch the forward took.
 * dsem_rk_step_vjp_tape_lagged: the adjoint of the step recorded in `slot`
 *   (dsem_rk_step_record keeps its mu_h and the lagged states it read);
 *   dQbar and dSlagbar (device) hold on entry the adjoints of the state and
 *   of the lagged states after the step, on return those before it.
 *   The same holds with the Smagorinsky model (dsem_set_les, sensor
 *   switches included), whose nu_t tstep also computes once per step from
 *   the start state and the lagged states (simpleSmag,
 *   staticSmag.f90:32-103): the slot keeps nu_t, and its adjoint goes back
 *   through k_smag's (clamp and prolongation zeroing take the forward's
 *   branch; the sensors are recomputed from the same states); nubar includes
 *   nu_t's own dependence on nu = 1/Re (nu_t = Re (C_s Delta)^2 |S|).  The
 *   tape's layout follows the Smagorinsky switch at dsem_tape_reserve: after
 *   switching, reserve again before recording.
 * dsem_get/set_lagged_states_device: device copies for checkpoints; valid
 *   receives whether the context holds lagged states (a fresh start does
 *   not); set with NULL makes the next step a fresh start (mu_h = 0). */
int  dsem_ev_vjp(dsem_ctx *ctx, const double *muhbar, double *JQbar, double *Slagbar);
int  dsem_rk_step_vjp_tape_lagged(dsem_ctx *ctx, double dt, int slot, double *dQbar, double *dSlagbar, double *nubar);
int  dsem_get_lagged_states_device(dsem_ctx *ctx, double *dstag, int32_t *valid);
int  dsem_set_lagged_states_device(dsem_ctx *ctx, const double *dstag);

/* modalfilter (src/Energydissipation/energydissipation.f90:720-959; the
 * driver calls it every 50 steps before tstep, src/3DDriver_parfmdf.f90:462-463)
 * on the context's state: per element the momentum J*Q times jacob(1,1,1)
 * through a 3-D DCT-II (FFTW REDFT10) with the reference's rescaling, modes
 * with any index >= cutoff set to zero, and back through a 3-D DCT-III
 * (REDFT01).  cutoff = p+1 keeps every mode, as the reference as shipped
 * does (its mode edit is commented out): the identity up to round-off. */
int  dsem_modal_filter(dsem_ctx *ctx, int cutoff);

/* Device-to-device state transfer (checkpointing): dJQ [K][5][ncg^3]. */
int  dsem_set_state_device(dsem_ctx *ctx, const double *dJQ);
int  dsem_get_state_device(dsem_ctx *ctx, double *dJQ);

/* Gauss-point quadrature weights on [0,1]^3 used by dsem_reduce_ke
 * (bmatu1b, caseFiles/tgv/BackstepProblem-stoch.f90:345-361), [ncg^3]. */
int  dsem_quadrature_weights(const dsem_ctx *ctx, double *bmat);

/* Static Smagorinsky eddy viscosity (input keyword 'smagorinsky';
 * simpleSmag + nut_prolong, src/turbulence/staticSmago/staticSmag.f90:32-103,
 * 206-261, called once per step by tstep, src/tstep.f90:40-43).
 * nu_t = Re (C_s Delta)^2 |S| from the lagged staggered states, capped at
 * numax, prolonged to the staggered grids; the viscous fluxes use
 * amu = 1 + nu_t rho (+ mu_h), Vis_Fluxes_new.f90:123-125.  Zero on the first
 * step after dsem_set_state (sensor switches: dsem_set_les_sensors).
 * dsem_les_update recomputes nu_t from the current lagged states and returns
 * it at the Gauss points [K][ncg^3] and on the staggered grids
 * [K][3][nc*ncg^2] (either nullable); dsem_set_les_viscosity prescribes the
 * staggered field for dsem_rhs. */
int  dsem_set_les(dsem_ctx *ctx, int enable, double cs, double numax);

/* Sensor switches of the Smagorinsky model (input keywords shockSensor and
 * rhoSensor, staticSmag.f90:68-69): nu_t is multiplied by 1 - ducros and/or
 * by the density-based sensor (RHS.f90:263-275) before the clamp.  Both
 * sensors are computed by the entropy-viscosity update of the same step
 * (shock = 1, which tstep runs first, src/tstep.f90:36-43); without it they
 * stay zero, as in the reference.  dsem_get_ev_sensors returns the Gauss-point
 * sensors [K][ncg^3] of the last entropy-viscosity update (either nullable). */
#define DSEM_LES_SHOCK_SENSOR 1
#define DSEM_LES_RHO_SENSOR   2
int  dsem_set_les_sensors(dsem_ctx *ctx, int sensors);
int  dsem_get_ev_sensors(dsem_ctx *ctx, double *ducros, double *rho_sensor);
int  dsem_les_update(dsem_ctx *ctx, double *nut_g, double *nut_s);
int  dsem_set_les_viscosity(dsem_ctx *ctx, const double *nut_s);

/* Keep the last stage's face-averaged staggered states after every step
 * (d%Qlgg/Qglg/Qggl survive tstep in the reference) when entropy viscosity
 * is off, for the enstrophy diagnostic; costs one extra write of
 * 3*5*nc*ncg^2 doubles per element per step. */
int  dsem_keep_lagged_states(dsem_ctx *ctx, int enable);

/* Device-pointer entry points (same HIP device, stream-ordered on the
 * context's stream): dJQ and dQt are [K][5][ncg^3] FP64 device buffers. */
int  dsem_rhs_device(dsem_ctx *ctx, const double *dJQ, double *dQt);

/* dsem_rhs_device without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels and result as dsem_rhs_device, but the RHS intermediates (face
 * traces, mortar averages and fluxes, viscous face traces, partial residual:
 * 8 * (12 K ncg^2 * 5 + 15 nmortar ncg^2 + 5 K ncg^3) bytes, each part rounded
 * up to 256 bytes) live in one static device array of
 * DSEM_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends, instead of
 * the context's buffers.  Meshes that need more return DSEM_ERR_ARG (at p = 8
 * the limit is about 690 elements).  Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU (contexts on different streams, or
 * threads) race; calls on one context are ordered by its stream. */
#define DSEM_NOALLOC_SCRATCH_BYTES (64ull << 20)
int  dsem_rhs_device_noalloc(dsem_ctx *ctx, const double *dJQ, double *dQt);

int  dsem_state_pointers(dsem_ctx *ctx, double **dJQ, double **dG);
int  dsem_stream(dsem_ctx *ctx, void **hip_stream);
int  dsem_synchronize(dsem_ctx *ctx);

/* Per-kernel timing with HIP events on the context's stream.  When enabled,
 * every launch is bracketed by events and synchronised, so totals are
 * kernel durations; use it for profiling runs only. */
int  dsem_timing_enable(dsem_ctx *ctx, int enable);
int  dsem_timing_reset(dsem_ctx *ctx);
int  dsem_timing_count(void);
int  dsem_timing_query(const dsem_ctx *ctx, int kernel_id, const char **name, double *total_ms, int64_t *launches);

#ifdef __cplusplus
}
#endif
#endif /* DSEM_HIP_H */
