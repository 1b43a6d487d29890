/*
 * arwen_dycore.h -- C-ABI of the MI355X (gfx950) dycore-step library
 * libarwen_dycore.so.
 *
 * The reference binds its device code through one seam,
 * gpuwm/core/kernels/__init__.py:140-143 get_kernel(module, func), which
 * returns a CuPy RawFunction called as kern(grid, block, args).  This
 * library replaces that seam: every entry point takes plain device
 * pointers (FP32 arrays in the reference layout (nz, ny, nx), x fastest,
 * C-grid staggering u (nz,ny,nx+1), v (nz,ny+1,nx), w/php (nz+1,ny,nx)),
 * integer extents, scalars and a hipStream_t passed as void*.  Nothing
 * allocates: all storage is owned by the caller (the Python DomainState).
 * Every entry point returns ARW_OK (0) or a negative ARW_E* code and never
 * throws across the ABI; arw_last_error() returns a message for the last
 * failure on the calling thread.
 */
#ifndef ARWEN_DYCORE_H
#define ARWEN_DYCORE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ARW_ABI_VERSION 9

/* transported species beyond qv/qc/qr a domain can carry (NSSL's set) */
#define ARW_MAX_EXTRA 13

/* flags of the adjoint entry points that accumulate */
#define ARW_ADJ_CLEAR 1
#define ARW_ADJ_ACCUMULATE 2

#define ARW_OK 0
#define ARW_E_ARG (-1)      /* invalid argument (null pointer, bad extent) */
#define ARW_E_LAUNCH (-2)   /* a kernel launch or runtime call failed */
#define ARW_E_UNSUPPORTED (-3)

/*
 * One domain's device arrays.  Field names follow the reference
 * DomainState (gpuwm/core/state.py:389-695); scratch arrays replace the
 * reference's named scratch slots (acoustic_c2a/_a/_alpha/_gamma,
 * rk_ww, state.py:805-830).
 */
typedef struct arw_domain {
    int32_t nz, ny, nx;
    int32_t base3d;            /* thb/pb/alb/phb are (nz[+1],ny,nx) */
    /* prognostics */
    float *u, *v, *w, *thp, *php, *mup;
    /* EOS diagnostics */
    float *p, *al, *alt;
    /* RK time-t copies */
    float *u0, *v0, *w0, *thp0, *php0, *mup0;
    /* coupled slow tendencies */
    float *ru_t, *rv_t, *rw_t, *rth_t, *rph_t, *rmu_t;
    /* acoustic perturbations */
    float *u_pp, *v_pp, *w_pp, *th_pp, *ph_pp, *mu_pp;
    float *p_pp, *p_pp_old, *ww_pp, *al_pp;
    /* base state, terrain, base column mass */
    const float *thb, *pb, *alb, *phb, *mub2d, *ht;
    /* vertical coordinate */
    const float *c1h, *c2h, *c1f, *c2f, *dnw, *rdnw, *dn, *rdn, *fnm, *fnp;
    float cf1, cf2, cf3, cfn, cfn1;
    /* scratch: ww (Omega, nz+1,ny,nx), gam (LU factor, nz+1,ny,nx),
     * dmdt_rf (reference mass divergence, ny,nx), mup_next (ny,nx) */
    float *ww, *gam, *dmdt_rf, *mup_next;
    /* constant-K diffusion inverse spacings (reference diffusion.py:34-61):
     * half-level fields rdzf_h (nz-1), rdzc_h (nz); w rdzf_w (nz),
     * rdzc_w (nz+1).  May be NULL when khdif = kvdif = 0. */
    const float *rdzf_h, *rdzc_h, *rdzf_w, *rdzc_w;
    /* moist scalars (reference state.py:408-427), all NULL for a dry
     * domain.  qv/qc/qr hold time t on entry to arw_step and the new
     * state on exit; qv0/qc0/qr0 and qsv/qsc/qsr are stage buffers. */
    float *qv, *qc, *qr, *qv0, *qc0, *qr0, *qsv, *qsc, *qsr;
    /* moist scratch: substep-averaged mass fluxes (WRF sumflux; ru_m
     * (nz,ny,nx+1), rv_m (nz,ny+1,nx), ww_m (nz+1,ny,nx)) and the
     * positive-definite limiter's face fluxes (fxl/fxc like u, fyl/fyc
     * like v, fzl/fzc like w) and cell scales (nz,ny,nx) */
    float *ru_m, *rv_m, *ww_m;
    float *pd_fxl, *pd_fxc, *pd_fyl, *pd_fyc, *pd_fzl, *pd_fzc, *pd_sc;
    /* precision form only (else NULL): FP64 tridiagonal intermediates,
     * the LU factor gamma then the forward-eliminated w'', each
     * (nz+1,ny,nx); and pp, the perturbation pressure p - pb of the last
     * EOS (nz,ny,nx), which the perturbation-form EOS keeps for the
     * pressure-gradient terms */
    double *tri64;
    float *pp;
    /* diff_6th_opt > 0 only (else NULL): the held 6th-order tendencies of
     * u, v, w, theta' (shapes of u, v, w, thp), formed once per step from
     * the time-t fields (reference dycore.py:1345-1420) */
    float *d6_u, *d6_v, *d6_w, *d6_t;
    /* ABI 7: a moist state's transported species beyond qv/qc/qr (the
     * mixed-phase schemes' ice masses and number moments, reference
     * gpuwm/core/moist.py:107-240): their count, how many of them (the
     * first, in order) add to the condensate loading of the buoyancy
     * (reference dycore.cu:213-230: 3 for qi, qs, qg, 4 with qh), and per
     * species the state array (time t on entry, new on exit), its time-t
     * copy and a third stage buffer, all (nz,ny,nx). */
    int32_t n_extra, n_load;
    float *qx[ARW_MAX_EXTRA], *qx0[ARW_MAX_EXTRA], *qxs[ARW_MAX_EXTRA];
    /* the retained microphysics heating rate (K/s, (nz,ny,nx)); NULL when
     * mp_physics = 0.  Its coupled form is added to every stage's theta
     * tendency and the final stage's finish removes what the substeps
     * integrated (reference dycore.py:572-592, dycore.cu:1186-1194). */
    float *h_diabatic;
    /* WRF hypsometric_opt of the EOS's specific volume (reference
     * diagnostics.cu:37-89): 1 (or 0) from the layer's geopotential
     * thickness, 2 from the log of its dry-pressure ratio, which reads the
     * hybrid coefficients c3h, c4h (nz) and c3f, c4f (nz+1) and the model
     * top pressure p_top.  Plain form only. */
    int32_t hypsometric_opt;
    const float *c3h, *c4h, *c3f, *c4f;
    double p_top;
    /* moist_cq's stage coefficients (u, v and w shapes; scratch) */
    float *cqu, *cqv, *cqw;
    /* km_opt = 4 (arw_params): scratch of 9 nz ny nx + 5 (nz+1) ny nx
     * elements for the metrics, deformation, coefficients and stresses */
    float *sm_buf;
    /* moist states with held mixing (km_opt = 4 or diff_6th_opt): the held
     * tendency of every transported species, nz ny nx elements each, in
     * the order qv, qc, qr, then the extras */
    float *held_q;
} arw_domain;

/*
 * The same domain in double precision: identical fields and order, FP64
 * arrays.  The arw64_* entry points run the identical discretisation in
 * FP64 -- the reference the float32 step's per-field error is stated
 * against, and the precision of the adjoint's dot-product test.  No
 * reference counterpart: ArWen's verification mirrors are CPU-only
 * (gpuwm/verify/npref.py).
 */
typedef struct arw64_domain {
    int32_t nz, ny, nx;
    int32_t base3d;
    double *u, *v, *w, *thp, *php, *mup;
    double *p, *al, *alt;
    double *u0, *v0, *w0, *thp0, *php0, *mup0;
    double *ru_t, *rv_t, *rw_t, *rth_t, *rph_t, *rmu_t;
    double *u_pp, *v_pp, *w_pp, *th_pp, *ph_pp, *mu_pp;
    double *p_pp, *p_pp_old, *ww_pp, *al_pp;
    const double *thb, *pb, *alb, *phb, *mub2d, *ht;
    const double *c1h, *c2h, *c1f, *c2f, *dnw, *rdnw, *dn, *rdn, *fnm, *fnp;
    double cf1, cf2, cf3, cfn, cfn1;
    double *ww, *gam, *dmdt_rf, *mup_next;
    const double *rdzf_h, *rdzc_h, *rdzf_w, *rdzc_w;
    double *qv, *qc, *qr, *qv0, *qc0, *qr0, *qsv, *qsc, *qsr;
    double *ru_m, *rv_m, *ww_m;
    double *pd_fxl, *pd_fxc, *pd_fyl, *pd_fyc, *pd_fzl, *pd_fzc, *pd_sc;
    double *tri64;
    double *pp;                /* unused by the FP64 build */
    double *d6_u, *d6_v, *d6_w, *d6_t;
    int32_t n_extra, n_load;
    double *qx[ARW_MAX_EXTRA], *qx0[ARW_MAX_EXTRA], *qxs[ARW_MAX_EXTRA];
    double *h_diabatic;
    int32_t hypsometric_opt;
    const double *c3h, *c4h, *c3f, *c4f;
    double p_top;
    double *cqu, *cqv, *cqw;
    double *sm_buf;
    double *held_q;
} arw64_domain;

/* Step configuration (the RunConfig fields the step reads,
 * gpuwm/config.py:28-190).  Scalars that the reference derives in
 * double precision on the host and hands to its kernels as float32
 * (rdx = float32(1/dx), dtau = float32(dt/ns), ...) are derived here the
 * same way from these doubles. */
typedef struct arw_params {
    double dx, dy, dt;
    int32_t time_step_sound;
    double epssm, smdiv;
    double khdif, kvdif;
    int32_t damp_opt;
    double zdamp, dampcoef;
    int32_t top_lid;
    int32_t moist_adv_opt;     /* 1: PD limiter on the final RK stage */
    /* 1: precision form of the float32 step -- the EOS, the pressure
     * gradient, buoyancy and rhs_ph terms and the theta'' init / theta
     * finish in perturbation-from-base form (needs d->pp), compensated
     * (Kahan) column sums of the mass divergence and Omega, the implicit
     * w/phi tridiagonal factored and solved in FP64 (needs d->tri64).  No
     * reference counterpart; ignored by the FP64 build. */
    int32_t precision_form;
    /* WRF sixth_order_diffusion held once per step (reference dycore.py:
     * 1345-1420, kernels/diff6.cu): 0 off, 1 plain, 2 monotonic (no
     * up-gradient flux); the per-step factor (the caller applies the
     * model-clock scaling of dycore.py:1767-1786); the terrain-slope taper
     * switch and threshold.  arw_step only; dry states, plain form. */
    int32_t diff_6th_opt;
    double diff_6th_factor;
    int32_t diff_6th_slopeopt;
    double diff_6th_thresh;
    /* WRF h_sca_adv_order of rhs_ph's horizontal geopotential advection
     * (reference dycore.cu:541-748): 2, or 5 (the face-flux sum times the
     * centred difference (45 d1 - 9 d2 + d3) / 60); 0 means 2. */
    int32_t h_sca_adv_order;
    /* WRF moist_cq (reference acoustic.py:61-110, acoustic.cu:84-158 and
     * the uses :287, :307, :700-744, :899-905, dycore.cu:195, :208): 0
     * off; otherwise the number of moist mass species summed into
     * cq = 1 / (1 + 0.5 qtot) at faces and w levels -- 1 (qv), 3 (qv, qc,
     * qr) or 3 + n_load (the ice and hail masses).  cq scales the slow and
     * acoustic horizontal pressure gradients and the implicit w solve's
     * pressure terms; needs d->cqu, cqv, cqw.  Moist plain-form steps. */
    int32_t moist_cq;
    /* WRF diff_opt = 2 eddy viscosity (reference dycore.py:1223-1343,
     * kernels/smag2d.cu:272-1340): km_opt = 4 holds the 2-D Smagorinsky
     * stress divergences of u, v, w and the theta flux divergence of the
     * time-t fields, once per step, and adds them to every RK stage's
     * tendencies (before a 6th-order filter's); needs d->sm_buf and the
     * d6_* held buffers.  c_s is the Smagorinsky constant.  bl_pbl_physics
     * 0: the step also mixes vertically, with the surface policy isfflx
     * (0: the constant drag coefficient tke_drag_coefficient and the
     * kinematic heat flux tke_heat_flux, 2: the heat flux, 1: neither --
     * no surface-layer scheme runs inside the step); otherwise a PBL
     * scheme outside the step owns the vertical.  Other km_opt values
     * leave this off (km_opt = 1's constant K is khdif / kvdif).
     * arw_step only; dry states, plain form, nz >= 3. */
    int32_t km_opt;
    double c_s;
    int32_t bl_pbl_physics;
    int32_t isfflx;
    double tke_heat_flux, tke_drag_coefficient;
    /* moist states under diff_6th_opt: bit e set filters species e (qv,
     * qc, qr, then the extras); WRF's moist_mix6_off clears the moist
     * array's bits (reference dycore.py:1200-1220).  The held tendencies
     * of the species (d->held_q) join the unlimited scalar updates of
     * every RK stage and are folded into the time-t scalar ahead of the
     * final stage's positive-definite limiter (moist.py:293-346). */
    uint32_t diff6_species;
} arw_params;

int32_t arw_abi_version(void);
const char *arw_last_error(void);

/* Whole RK3 split-explicit step, in place (replaces gpuwm/core/dycore.py:
 * 2161-2503 ``step`` for periodic, msf = 1 states, dry or carrying the
 * qv/qc/qr scalars and any mixed-phase scheme's extra species (n_extra,
 * with h_diabatic for mp_physics != 0): moist EOS and buoyancy, WRF
 * sumflux, the coupled scalar stage update of moist.py:434-698 and the
 * positive-definite limiter of pd_advection.cu on the final stage when
 * moist_adv_opt = 1).  Leaves p/al/alt consistent with the new state,
 * like the reference's closing update_diagnostics (dycore.py:2482). */
int32_t arw_step(const arw_domain *d, const arw_params *prm, void *stream);

/* The advection-only step, step(acoustic=False) (replaces
 * gpuwm/core/dycore.py:2318-2343 with advection.py:134-182 and the stage
 * update dycore.py:1602-1643): three stages of length dt/3, dt/2, dt of
 * coupled flux-form advection of theta, u, v, w with Omega = 0 and the
 * constant-K diffusion, each advanced from the time-t copies (u0..mup0,
 * written here) with w, phi' and mu' frozen; then the kinematic lower
 * boundary.  Dry, plain form, no 6th-order filter (ARW_E_UNSUPPORTED
 * otherwise); the caller guarantees w == 0 (the reference raises
 * ValueError otherwise).  u_pp, v_pp, ww_pp and th_pp serve as work
 * arrays; p/al/alt are left at the last stage's EOS, as in the
 * reference. */
int32_t arw_step_advective(const arw_domain *d, const arw_params *prm,
                           void *stream);

/* arw_step with no workspace and no allocation at all.  No counterpart
 * in the reference: it is for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The step's scratch (d->ww, d->gam, d->dmdt_rf,
 * d->mup_next are ignored and may be null) comes from one array of
 * 64 MiB per GPU that the library holds until the process ends; a domain
 * needing more (2*(nz+1)*ny*nx + 2*ny*nx values, each part rounded up to
 * 256 bytes) returns ARW_E_ARG.  Dry plain-form steps without the
 * 6th-order filter (ARW_E_UNSUPPORTED otherwise).  Calls that overlap on
 * one GPU share the array and race.  Bit-identical to arw_step. */
int32_t arw_step_noalloc(const arw_domain *d, const arw_params *prm,
                         void *stream);

/*
 * Stage checkpoints of one step's scalar transport: per RK stage the
 * substep-averaged mass fluxes (WRF sumflux: ru_m (nz,ny,nx+1), rv_m
 * (nz,ny+1,nx), ww_m (nz+1,ny,nx)), the post-acoustic mu' (ny,nx) and
 * the stage length; mup_t is the time-t mu'.  Filled by arw_step_ckpt,
 * read by arw_tracer_step / arw_tracer_adjoint.  No reference
 * counterpart: ArWen has no adjoint (the checkpoints follow WRF's
 * rk_scalar_tend inputs, reference moist.py:434-698).
 */
typedef struct arw_tracer_ckpt {
    float *ru_m[3], *rv_m[3], *ww_m[3], *mup_new[3];
    float *mup_t;
    double dt[3];
} arw_tracer_ckpt;

/* arw_step that also records the stage checkpoints (the sumflux
 * averages are accumulated even for a dry domain, which then needs the
 * ru_m/rv_m/ww_m scratch). */
int32_t arw_step_ckpt(const arw_domain *d, const arw_params *prm,
                      arw_tracer_ckpt *ck, void *stream);

/* One RK3 step of unlimited flux-form transport of a scalar (reference
 * moist.py:385-398 per stage, moist_adv_opt = 0) with the checkpointed
 * fluxes and masses: q_out from q_in; work holds 2*nz*ny*nx; the final
 * stage clamps at zero when clamp != 0.  Bit-identical to the scalar
 * update inside arw_step. */
int32_t arw_tracer_step(const arw_domain *d, const arw_params *prm,
                        const arw_tracer_ckpt *ck, const float *q_in,
                        float *q_out, float *work, int32_t clamp,
                        void *stream);

/* Hand-written adjoint (transpose) of arw_tracer_step without an active
 * clamp: lam_out = J^T lam_in; work holds 3*nz*ny*nx.  No reference
 * counterpart. */
int32_t arw_tracer_adjoint(const arw_domain *d, const arw_params *prm,
                           const arw_tracer_ckpt *ck, const float *lam_in,
                           float *lam_out, float *work, void *stream);

/*
 * Ensemble: independent periodic domains stepped together.  No reference
 * counterpart: ArWen runs ensemble members one after another
 * (gpuwm/ensemble/engine.py:1-14).  Every pass of the step is one launch
 * over all members of a group; a work table maps each workgroup to a
 * member and to a workgroup of that member.  Members may differ in nx,
 * ny and nz; they share arw_params and base3d and are dry.  Each member
 * advances exactly as arw_step would advance it (same kernels, same
 * arithmetic, bit-identical results).
 *
 * order:  member order in the work tables (a permutation of 0..n-1, or
 *         NULL for 0..n-1); members early in the order start first.
 * groups: split the ordered members into this many groups of about equal
 *         work; each group is its own launch sequence, and a captured
 *         graph runs the groups as parallel branches (1 <= groups <= n).
 * col_tile: columns per workgroup of the batched column solve (64, 128
 *         or 256; results do not depend on it).
 * The handle owns its device tables until arw_ensemble_destroy.
 */
typedef struct arw_ensemble arw_ensemble;
int32_t arw_ensemble_create(const arw_domain *members, int32_t n,
                            const arw_params *prm, const int32_t *order,
                            int32_t groups, int32_t col_tile,
                            arw_ensemble **out);
/* nsteps steps, launched on stream. */
int32_t arw_ensemble_step(arw_ensemble *e, int32_t nsteps, void *stream);
/* Capture nsteps steps into one HIP graph held by the handle (replacing
 * an earlier one); arw_ensemble_launch runs it on stream.  With several
 * groups the branches rejoin every join_every steps (0: only at the end),
 * part of the launch schedule the autotuner picks. */
int32_t arw_ensemble_capture(arw_ensemble *e, int32_t nsteps,
                             int32_t join_every);
int32_t arw_ensemble_launch(arw_ensemble *e, void *stream);
int32_t arw_ensemble_destroy(arw_ensemble *e);

/*
 * Adjoint of one acoustic substep (arw_acoustic_substep; reference
 * acoustic.cu:160-982 advance_uv -> advance_mu_th -> advance_w_phi, which
 * has no adjoint in the reference).  The substep's input X is the
 * domain's small-step state (u_pp, v_pp, w_pp, th_pp, ph_pp, mu_pp, p_pp,
 * p_pp_old, al_pp) with the stage coefficients prepared as for the
 * forward substep (arw_calc_coefs); the domain is not modified.  l_*
 * are the adjoints of the substep's outputs (l_ww_pp may be NULL), g_*
 * receive J^T l for the inputs and the slow tendencies (same shapes as the
 * arw_domain fields of the same names).  The Jacobian is taken at X (the
 * new column mass enters three denominators).  work holds
 * arw_acoustic_adjoint_work(...) elements.
 */
typedef struct arw_acoustic_adj {
    const float *l_u_pp, *l_v_pp, *l_w_pp, *l_th_pp, *l_ph_pp, *l_mu_pp,
        *l_p_pp, *l_p_pp_old, *l_al_pp, *l_ww_pp;
    float *g_u_pp, *g_v_pp, *g_w_pp, *g_th_pp, *g_ph_pp, *g_mu_pp, *g_p_pp,
        *g_p_pp_old, *g_al_pp;
    float *g_ru_t, *g_rv_t, *g_rw_t, *g_rth_t, *g_rph_t, *g_rmu_t;
} arw_acoustic_adj;
int32_t arw_acoustic_substep_adjoint(const arw_domain *d,
                                     const arw_params *prm, double dtau,
                                     int32_t first,
                                     const arw_acoustic_adj *adj,
                                     float *work, void *stream);
int32_t arw_acoustic_adjoint_work(int32_t nz, int32_t ny, int32_t nx,
                                  int64_t *elements);

/* The substep adjoint with the stage reference state's share (no
 * counterpart in the reference, which has no adjoint; with the
 * small-step init/finish and tendency-pass adjoints it composes the
 * adjoint of a whole dycore.py:2161-2503 step).  The stage fields enter a
 * substep through c2a = gamma p / alpha_t and alpha_t of the stage EOS,
 * the column mass mu_t, theta, the full phi, w (the damper) and u, v
 * (dmdt_rf): the call ADDS J^T lam with respect to those to the
 * accumulators of acc (mut, grf: (ny,nx); c2a, alt, thr: mass points; ph,
 * w: w levels), which the caller clears at the stage start (or lets the
 * stage's first call write, ARW_ADJ_CLEAR) and keeps over the stage's
 * substeps.  fwd (optional): the substep's
 * outputs from d's small-step fields (u''..al'' and Omega'' after
 * arw_acoustic_substep), which the call then does not recompute.  lu
 * (optional): the stage's implicit-w Thomas factors from
 * arw_acoustic_adjoint_factors (3 fields of (nz+1)*ny*nx; they depend on
 * the stage state and dtau only), read instead of formed.  flags:
 * ARW_ADJ_CLEAR -- this call writes the accumulators instead of adding
 * (no clearing needed before a stage's first call); ARW_ADJ_ACCUMULATE
 * -- the tendency adjoints g_r*_t are added to what they hold (a stage's
 * substeps summed in place).  Otherwise as arw_acoustic_substep_adjoint
 * (same work size).  Dry states only.  arw_acoustic_reference_gradients
 * maps the accumulators to the stage fields' gradients g_u .. g_mup:
 * dmdt_rf's share (grf times the stage u, v and face masses) to g_u,
 * g_v, acc->w to g_w, and c2a / alt through the stage EOS (d->p, d->alt
 * must hold it) to g_thp, g_php, g_mup; written, or with
 * ARW_ADJ_ACCUMULATE added to what they hold; work: 2 fields of
 * (nz+1)(ny+1)(nx+1), e.g. the substep adjoint's work.
 * arw_set_w_surface_adjoint ADDS the (u, v) gradients of the kinematic
 * lower boundary w(sfc) = u . grad(ht) (replaces nothing: the transpose of
 * dycore.py:1999-2026) for the adjoint lam_w0 (ny,nx) of w(sfc). */
typedef struct arw_acoustic_ref {
    float *mut, *c2a, *alt, *thr, *ph, *w, *grf;
} arw_acoustic_ref;
typedef struct arw_acoustic_fwd {
    const float *u_pp, *v_pp, *w_pp, *th_pp, *mu_pp, *al_pp, *ww_pp;
} arw_acoustic_fwd;
int32_t arw_acoustic_substep_adjoint_ref(const arw_domain *d,
                                         const arw_params *prm, double dtau,
                                         int32_t first,
                                         const arw_acoustic_adj *adj,
                                         const arw_acoustic_ref *acc,
                                         const arw_acoustic_fwd *fwd,
                                         const float *lu, int32_t flags,
                                         float *work, void *stream);
int32_t arw_acoustic_adjoint_factors(const arw_domain *d,
                                     const arw_params *prm, double dtau,
                                     float *lu, void *stream);
int32_t arw_acoustic_reference_gradients(const arw_domain *d,
                                         const arw_params *prm,
                                         const arw_acoustic_ref *acc,
                                         float *g_u, float *g_v, float *g_w,
                                         float *g_thp, float *g_php, float *g_mup,
                                         float *g_qv, int32_t flags, float *work,
                                         void *stream);
int32_t arw_set_w_surface_adjoint(const arw_domain *d, const arw_params *prm,
                                  const float *lam_w0, float *g_u, float *g_v,
                                  void *stream);

/* One stage of the scalars' unlimited update with explicit fluxes and
 * masses (no counterpart in the reference as an entry point; its
 * arithmetic is moist.py:385-398 _update_scalar_kernel with the
 * substep-averaged fluxes): q_new = (C(mu_t) q0 + dt flux_div(q; ru, rv,
 * ww)) / C(mu_new), clamped at zero when clamp.  mup_t, mup_new: mu' of
 * time t and after the stage's finish.  arw_scalar_stage_adjoint: its
 * transpose for lam = adjoint of q_new; g_q is written, g_q0, g_ru, g_rv,
 * g_ww, g_mup_t, g_mup_new are added to; the clamp is differentiated as 1
 * where q_new > 0, 0 elsewhere; work: 3*nz*ny*nx elements. */
int32_t arw_scalar_stage(const arw_domain *d, const arw_params *prm,
                         double dt_stage, int32_t clamp, const float *q,
                         const float *q0, const float *ru, const float *rv,
                         const float *ww, const float *mup_t,
                         const float *mup_new, float *q_new, void *stream);
int32_t arw_scalar_stage_adjoint(const arw_domain *d, const arw_params *prm,
                                 double dt_stage, int32_t clamp,
                                 const float *q, const float *q0, const float *ru,
                                 const float *rv, const float *ww,
                                 const float *mup_t, const float *mup_new,
                                 const float *lam, float *g_q, float *g_q0,
                                 float *g_ru, float *g_rv, float *g_ww,
                                 float *g_mup_t, float *g_mup_new, float *work,
                                 void *stream);

/* The final RK stage's update of one scalar with WRF's positive-definite
 * limiter (replaces the moist_adv_opt = 1 branch of moist.py:434-698:
 * pd_advection.cu pd_fluxes -> pd_scale -> pd_renorm_apply, then the
 * clamped coupled update moist.py:385-398), with explicit averaged fluxes
 * and masses; bit-identical to the step's own final-stage scalar.
 * arw_pd_stage_adjoint: its transpose for lam = adjoint of q_new; g_q is
 * written, g_q0, g_ru, g_rv, g_ww, g_mup_t, g_mup_new are added to; each
 * branch (the clamp, a correction's donor cell, the factor's [0, 1]
 * clamps, the upwind CFL caps) is differentiated as taken.  work (both):
 * ARW_PD_WORK(nz, ny, nx) elements. */
#define ARW_PD_WORK(nz, ny, nx) \
    (26 * (size_t)((nz) + 1) * (size_t)((ny) + 1) * (size_t)((nx) + 1))
int32_t arw_pd_stage(const arw_domain *d, const arw_params *prm,
                     double dt_stage, const float *q, const float *q0,
                     const float *ru, const float *rv, const float *ww,
                     const float *mup_t, const float *mup_new, float *q_new,
                     float *work, void *stream);
int32_t arw_pd_stage_adjoint(const arw_domain *d, const arw_params *prm,
                             double dt_stage, const float *q, const float *q0,
                             const float *ru, const float *rv, const float *ww,
                             const float *mup_t, const float *mup_new,
                             const float *lam, float *g_q, float *g_q0,
                             float *g_ru, float *g_rv, float *g_ww,
                             float *g_mup_t, float *g_mup_new, float *work,
                             void *stream);

/* EOS p, al, alt from thp, php, mup (replaces diagnostics.cu:37-89
 * calc_p_alpha, launched by diagnostics.py:14-57; hypsometric_opt = 1). */
int32_t arw_calc_p_alpha(const arw_domain *d, void *stream);

/* Coupled mass fluxes and Omega at the current state (replaces
 * dycore.py:162-206 stage_fluxes: couple_momentum + _omega_ref).  ru is
 * (nz,ny,nx+1), rv (nz,ny+1,nx); Omega goes to d->ww. */
int32_t arw_stage_fluxes(const arw_domain *d, const arw_params *prm,
                         float *ru, float *rv, void *stream);

/* Flux-form advection kernels, periodic, msf = 1 (replace advection.cu:
 * 206-416 flux_div_scalar, 424-611 flux_div_u, 616-794 flux_div_v,
 * 803-1051 flux_div_w, launched by advection.py:31-129).  They ADD into
 * tend.  rdz is rdnw (scalar/u/v) or rdn (w). */
int32_t arw_flux_div_scalar(const float *q, const float *ru, const float *rv,
                            const float *rw, float *tend, const float *rdnw,
                            const float *fnm, const float *fnp, float dx_inv,
                            float dy_inv, int32_t nz, int32_t ny, int32_t nx,
                            void *stream);
int32_t arw_flux_div_u(const float *u, const float *ru, const float *rv,
                       const float *rw, float *tend, const float *rdnw,
                       const float *fnm, const float *fnp, float dx_inv,
                       float dy_inv, int32_t nz, int32_t ny, int32_t nx,
                       void *stream);
int32_t arw_flux_div_v(const float *v, const float *ru, const float *rv,
                       const float *rw, float *tend, const float *rdnw,
                       const float *fnm, const float *fnp, float dx_inv,
                       float dy_inv, int32_t nz, int32_t ny, int32_t nx,
                       void *stream);
int32_t arw_flux_div_w(const float *w, const float *ru, const float *rv,
                       const float *rw, float *tend, const float *rdn,
                       const float *fnm, const float *fnp, float dx_inv,
                       float dy_inv, int32_t nz, int32_t ny, int32_t nx,
                       void *stream);

/* Adjoint of the large-step tendency pass of one RK stage (no
 * counterpart in the reference): the stage-start EOS, the coupled mass
 * fluxes and Omega, the advection of theta/u/v/w, the large-step PGF,
 * buoyancy and rhs_ph (diagnostics.cu:37-89, dycore.py:102-206,
 * advection.cu:206-1051, dycore.cu:98-308, :541-667) for a dry state
 * without constant-K diffusion, plain form.  For the adjoints l_* of
 * ru_t, rv_t, rw_t, rth_t, rph_t it WRITES the gradients g_* of the stage
 * state u, v, w, thp, php, mup (the transposed Jacobian at the state d).
 * The state is not modified: its EOS, Omega and fluxes are recomputed
 * into work (arw_stage_adjoint_work elements). */
typedef struct arw_stage_adj {
    const float *l_ru_t, *l_rv_t, *l_rw_t, *l_rth_t, *l_rph_t;
    float *g_u, *g_v, *g_w, *g_thp, *g_php, *g_mup;
    float *g_qv, *g_qc, *g_qr;     /* moist states: the stage scalars' */
    const float *l_ru, *l_rv, *l_ww; /* optional: adjoints of the stage fluxes */
    int32_t flags;   /* ARW_ADJ_ACCUMULATE: add to g_* instead of writing */
} arw_stage_adj;
int32_t arw_stage_tendencies_adjoint(const arw_domain *d,
                                     const arw_params *prm,
                                     const arw_stage_adj *adj, float *work,
                                     void *stream);
int32_t arw_stage_adjoint_work(int32_t nz, int32_t ny, int32_t nx,
                               int64_t *elements);

/* Adjoints of the small-step initialisation and finish (no counterpart in
 * the reference; dycore.cu:947-1085 and :1087-1198, plain form, dry).
 * The initialisation's transpose takes the adjoints of u_pp..p_pp_old and
 * WRITES the gradients of the time-t fields (g_*0) and of the stage fields
 * (g_*, including through the stage EOS's p and alt, recomputed into
 * work).  The finish's transpose takes the adjoints of the new stage
 * fields and WRITES the gradients of the stage fields and of the
 * small-step fields.  work: arw_small_step_adjoint_work elements. */
typedef struct arw_init_adj {
    const float *l_u_pp, *l_v_pp, *l_w_pp, *l_th_pp, *l_ph_pp, *l_mu_pp,
        *l_al_pp, *l_p_pp, *l_p_pp_old;
    float *g_u0, *g_v0, *g_w0, *g_thp0, *g_php0, *g_mup0;
    float *g_u, *g_v, *g_w, *g_thp, *g_php, *g_mup;
    float *g_qv;                   /* moist states: the stage qv's */
    int32_t flags;   /* ARW_ADJ_ACCUMULATE: add to g_* instead of writing */
} arw_init_adj;
typedef struct arw_finish_adj {
    const float *l_u, *l_v, *l_w, *l_thp, *l_php, *l_mup;
    float *g_u, *g_v, *g_w, *g_thp, *g_php, *g_mup;
    float *g_u_pp, *g_v_pp, *g_w_pp, *g_th_pp, *g_ph_pp, *g_mu_pp;
} arw_finish_adj;
int32_t arw_small_step_init_adjoint(const arw_domain *d,
                                    const arw_init_adj *adj, float *work,
                                    void *stream);
int32_t arw_small_step_finish_adjoint(const arw_domain *d,
                                      const arw_finish_adj *adj, float *work,
                                      void *stream);
int32_t arw_small_step_adjoint_work(int32_t nz, int32_t ny, int32_t nx,
                                    int64_t *elements);

/* Adjoints of the four advection kernels above (no counterpart in the
 * reference, which has no adjoint): for the output adjoint lam (shaped like
 * that kernel's tend) they ADD the transposed Jacobian of tend with respect
 * to the advected field, ru, rv and rw, at the given (field, ru, rv, rw),
 * into g_q, g_ru, g_rv, g_rw (any of which may be NULL to skip it).  The
 * duplicate column / row of ru, rv and of the u / v fields are separate
 * inputs, as the forward reads them.  |vel| is differentiated as
 * sign(vel), sign(0) = 0.  Outputs must not alias inputs. */
int32_t arw_flux_div_scalar_adjoint(const float *q, const float *ru,
                                    const float *rv, const float *rw,
                                    const float *lam, float *g_q, float *g_ru,
                                    float *g_rv, float *g_rw,
                                    const float *rdz, const float *fnm,
                                    const float *fnp, float dx_inv,
                                    float dy_inv, int32_t nz, int32_t ny,
                                    int32_t nx, void *stream);
int32_t arw_flux_div_u_adjoint(const float *q, const float *ru,
                               const float *rv, const float *rw,
                               const float *lam, float *g_q, float *g_ru,
                               float *g_rv, float *g_rw, const float *rdz,
                               const float *fnm, const float *fnp,
                               float dx_inv, float dy_inv, int32_t nz,
                               int32_t ny, int32_t nx, void *stream);
int32_t arw_flux_div_v_adjoint(const float *q, const float *ru,
                               const float *rv, const float *rw,
                               const float *lam, float *g_q, float *g_ru,
                               float *g_rv, float *g_rw, const float *rdz,
                               const float *fnm, const float *fnp,
                               float dx_inv, float dy_inv, int32_t nz,
                               int32_t ny, int32_t nx, void *stream);
int32_t arw_flux_div_w_adjoint(const float *q, const float *ru,
                               const float *rv, const float *rw,
                               const float *lam, float *g_q, float *g_ru,
                               float *g_rv, float *g_rw, const float *rdz,
                               const float *fnm, const float *fnp,
                               float dx_inv, float dy_inv, int32_t nz,
                               int32_t ny, int32_t nx, void *stream);

/* Stage preparation: EOS (when do_eos), Omega, the slow tendencies
 * ru_t..rph_t (advection + PGF + buoyancy + rhs_ph + constant-K
 * diffusion; replaces dycore.py:502-569 _add_slow_tendencies and
 * diffusion.py:100-130), small-step initialisation (dycore.cu:947-1085)
 * and the implicit-w coefficients (acoustic.cu:675-759 calc_coefs).
 * dtau is the stage's acoustic step. */
int32_t arw_stage_prepare(const arw_domain *d, const arw_params *prm,
                          double dtau, int32_t do_eos, void *stream);

/* One acoustic substep (replaces acoustic.py:330-479: advance_uv
 * acoustic.cu:244-314, advance_mu_th :329-456, advance_w_phi :774-982).
 * Uses the coefficients of the last arw_stage_prepare (or
 * arw_calc_coefs).  first != 0 drops divergence damping (smdiv = 0). */
int32_t arw_acoustic_substep(const arw_domain *d, const arw_params *prm,
                             double dtau, int32_t first, void *stream);

/* Per-stage acoustic preparation for the current reference state
 * (replaces acoustic.py:299-327 prepare_acoustic_coefficients): Omega and
 * the reference-momentum mass divergence.  The implicit-w LU factors
 * (acoustic.cu:675-759 calc_coefs) are formed inside every substep from
 * the stage reference p/alt, so none are stored. */
int32_t arw_calc_coefs(const arw_domain *d, const arw_params *prm,
                       double dtau, void *stream);

/* Small-step finish (dycore.cu:1087-1198) followed by the EOS of the new
 * state. */
int32_t arw_small_step_finish(const arw_domain *d, void *stream);

/* w(sfc) = u . grad(ht) (replaces dycore.py:1999-2026 set_w_surface). */
int32_t arw_set_w_surface(const arw_domain *d, const arw_params *prm,
                          void *stream);

/* Health maxima (replaces health.cu:23-222 health_partial/health_final as
 * used by dycore.py:2513-2582 stability_report): out[0..3] = max|u|,
 * max|w|, max|thp|, max |w[k>=1]| / dz(k) with dz from the live
 * geopotential.  NaNs propagate.  out is a device array of 4 floats. */
int32_t arw_health(const arw_domain *d, float *out, void *stream);

/*
 * Full-state health gate (replaces the reference's StateHealthValidator,
 * gpuwm/core/health.py:611-770, and its kernel validate_full_state,
 * kernels/health.cu:224-297).  One descriptor per checked field, in
 * DEVICE memory: the field (an array of the build's precision, float for
 * arw_, double for arw64_), an optional auxiliary added before the check
 * (same shape, or with ARW_HEALTH_AUX_LEVEL one value per `plane`
 * elements), its bounds and the status bit its failure sets.  A value
 * fails when it is not finite, below `lower` (at or below with
 * ARW_HEALTH_STRICT_LOWER) or above `upper`.  result (device, two
 * uint64): the OR of the failing fields' status bits, and the smallest
 * (field number << 48 | flat index) of a failing value (all ones when
 * none fails).  Launch: field x chunk pairs, nchunks chunks of `chunk`
 * elements per field.
 */
#define ARW_HEALTH_LOWER 1u
#define ARW_HEALTH_UPPER 2u
#define ARW_HEALTH_STRICT_LOWER 4u
#define ARW_HEALTH_ADD_AUX 8u
#define ARW_HEALTH_AUX_LEVEL 16u
typedef struct arw_health_field {
    const void *values;
    const void *aux;
    uint64_t size;
    double lower, upper;
    uint32_t flags;
    uint32_t plane;
    uint64_t status_bit;
} arw_health_field;
int32_t arw_health_gate(const arw_health_field *fields, int32_t nfields,
                        uint64_t chunk, uint32_t nchunks, uint64_t *result,
                        void *stream);

/* FP64 domain dry-mass sum(mub2d + mup) (replaces dycore.py:209-218
 * domain_mass_measure, identity map factors).  out is one device
 * double. */
int32_t arw_domain_mass(const arw_domain *d, double *out, void *stream);

/* Per-kernel timing of arw_step with HIP events on the launch stream.
 * Not part of the reference interface (the reference's equivalent is the
 * CUDA-event section profiler of tools/benchmark_seeded_step.py:384-420).
 * arw_profile_begin creates a pool of max_events events; every kernel
 * issued by arw_step afterwards records one.  arw_profile_end waits for
 * the last event, writes the summed milliseconds and launch counts per
 * kernel id (0 .. arw_kernel_count()-1) and frees the pool. */
int32_t arw_profile_begin(int32_t max_events);
int32_t arw_profile_end(double *ms, int32_t *launches, int32_t n_kernels);
const char *arw_kernel_name(int32_t id);
int32_t arw_kernel_count(void);

/* FP64 build of the same entry points (see arw64_domain). */
int32_t arw64_abi_version(void);
const char *arw64_last_error(void);
int32_t arw64_step(const arw64_domain *d, const arw_params *prm,
                   void *stream);
/* arw_step_noalloc of the FP64 build: no counterpart in the reference,
 * for callers that cannot allocate; its own 64 MiB array per GPU (half
 * the points of the FP32 build's limit). */
int32_t arw64_step_advective(const arw64_domain *d, const arw_params *prm,
                             void *stream);
int32_t arw64_step_noalloc(const arw64_domain *d, const arw_params *prm,
                           void *stream);
int32_t arw64_health_gate(const arw_health_field *fields, int32_t nfields,
                          uint64_t chunk, uint32_t nchunks, uint64_t *result,
                          void *stream);
int32_t arw64_calc_p_alpha(const arw64_domain *d, void *stream);
int32_t arw64_stage_fluxes(const arw64_domain *d, const arw_params *prm,
                           double *ru, double *rv, void *stream);
int32_t arw64_stage_prepare(const arw64_domain *d, const arw_params *prm,
                            double dtau, int32_t do_eos, void *stream);
int32_t arw64_acoustic_substep(const arw64_domain *d,
                               const arw_params *prm, double dtau,
                               int32_t first, void *stream);
int32_t arw64_calc_coefs(const arw64_domain *d, const arw_params *prm,
                         double dtau, void *stream);
int32_t arw64_small_step_finish(const arw64_domain *d, void *stream);
int32_t arw64_set_w_surface(const arw64_domain *d, const arw_params *prm,
                            void *stream);
int32_t arw64_health(const arw64_domain *d, double *out, void *stream);
int32_t arw64_domain_mass(const arw64_domain *d, double *out,
                          void *stream);
int32_t arw64_flux_div_scalar(const double *q, const double *ru,
                              const double *rv, const double *rw,
                              double *tend, const double *rdnw,
                              const double *fnm, const double *fnp,
                              double dx_inv, double dy_inv, int32_t nz,
                              int32_t ny, int32_t nx, void *stream);

int32_t arw64_flux_div_u(const double *u, const double *ru,
                         const double *rv, const double *rw, double *tend,
                         const double *rdnw, const double *fnm,
                         const double *fnp, double dx_inv, double dy_inv,
                         int32_t nz, int32_t ny, int32_t nx, void *stream);
int32_t arw64_flux_div_v(const double *v, const double *ru,
                         const double *rv, const double *rw, double *tend,
                         const double *rdnw, const double *fnm,
                         const double *fnp, double dx_inv, double dy_inv,
                         int32_t nz, int32_t ny, int32_t nx, void *stream);
int32_t arw64_flux_div_w(const double *w, const double *ru,
                         const double *rv, const double *rw, double *tend,
                         const double *rdn, const double *fnm,
                         const double *fnp, double dx_inv, double dy_inv,
                         int32_t nz, int32_t ny, int32_t nx, void *stream);
int32_t arw64_flux_div_scalar_adjoint(const double *q, const double *ru,
                                      const double *rv, const double *rw,
                                      const double *lam, double *g_q,
                                      double *g_ru, double *g_rv,
                                      double *g_rw, const double *rdz,
                                      const double *fnm, const double *fnp,
                                      double dx_inv, double dy_inv,
                                      int32_t nz, int32_t ny, int32_t nx,
                                      void *stream);
int32_t arw64_flux_div_u_adjoint(const double *q, const double *ru,
                                 const double *rv, const double *rw,
                                 const double *lam, double *g_q, double *g_ru,
                                 double *g_rv, double *g_rw,
                                 const double *rdz, const double *fnm,
                                 const double *fnp, double dx_inv,
                                 double dy_inv, int32_t nz, int32_t ny,
                                 int32_t nx, void *stream);
int32_t arw64_flux_div_v_adjoint(const double *q, const double *ru,
                                 const double *rv, const double *rw,
                                 const double *lam, double *g_q, double *g_ru,
                                 double *g_rv, double *g_rw,
                                 const double *rdz, const double *fnm,
                                 const double *fnp, double dx_inv,
                                 double dy_inv, int32_t nz, int32_t ny,
                                 int32_t nx, void *stream);
int32_t arw64_flux_div_w_adjoint(const double *q, const double *ru,
                                 const double *rv, const double *rw,
                                 const double *lam, double *g_q, double *g_ru,
                                 double *g_rv, double *g_rw,
                                 const double *rdz, const double *fnm,
                                 const double *fnp, double dx_inv,
                                 double dy_inv, int32_t nz, int32_t ny,
                                 int32_t nx, void *stream);

typedef struct arw64_tracer_ckpt {
    double *ru_m[3], *rv_m[3], *ww_m[3], *mup_new[3];
    double *mup_t;
    double dt[3];
} arw64_tracer_ckpt;
int32_t arw64_step_ckpt(const arw64_domain *d, const arw_params *prm,
                        arw64_tracer_ckpt *ck, void *stream);
int32_t arw64_tracer_step(const arw64_domain *d, const arw_params *prm,
                          const arw64_tracer_ckpt *ck, const double *q_in,
                          double *q_out, double *work, int32_t clamp,
                          void *stream);
int32_t arw64_tracer_adjoint(const arw64_domain *d, const arw_params *prm,
                             const arw64_tracer_ckpt *ck,
                             const double *lam_in, double *lam_out,
                             double *work, void *stream);

typedef struct arw64_acoustic_adj {
    const double *l_u_pp, *l_v_pp, *l_w_pp, *l_th_pp, *l_ph_pp, *l_mu_pp,
        *l_p_pp, *l_p_pp_old, *l_al_pp, *l_ww_pp;
    double *g_u_pp, *g_v_pp, *g_w_pp, *g_th_pp, *g_ph_pp, *g_mu_pp,
        *g_p_pp, *g_p_pp_old, *g_al_pp;
    double *g_ru_t, *g_rv_t, *g_rw_t, *g_rth_t, *g_rph_t, *g_rmu_t;
} arw64_acoustic_adj;
int32_t arw64_acoustic_substep_adjoint(const arw64_domain *d,
                                       const arw_params *prm, double dtau,
                                       int32_t first,
                                       const arw64_acoustic_adj *adj,
                                       double *work, void *stream);
int32_t arw64_acoustic_adjoint_work(int32_t nz, int32_t ny, int32_t nx,
                                    int64_t *elements);

typedef struct arw64_acoustic_ref {
    double *mut, *c2a, *alt, *thr, *ph, *w, *grf;
} arw64_acoustic_ref;
typedef struct arw64_acoustic_fwd {
    const double *u_pp, *v_pp, *w_pp, *th_pp, *mu_pp, *al_pp, *ww_pp;
} arw64_acoustic_fwd;
int32_t arw64_acoustic_substep_adjoint_ref(const arw64_domain *d,
                                         const arw_params *prm, double dtau,
                                         int32_t first,
                                         const arw64_acoustic_adj *adj,
                                         const arw64_acoustic_ref *acc,
                                         const arw64_acoustic_fwd *fwd,
                                         const double *lu, int32_t flags,
                                         double *work, void *stream);
int32_t arw64_acoustic_adjoint_factors(const arw64_domain *d,
                                     const arw_params *prm, double dtau,
                                     double *lu, void *stream);
int32_t arw64_acoustic_reference_gradients(const arw64_domain *d,
                                         const arw_params *prm,
                                         const arw64_acoustic_ref *acc,
                                         double *g_u, double *g_v, double *g_w,
                                         double *g_thp, double *g_php, double *g_mup,
                                         double *g_qv, int32_t flags, double *work,
                                         void *stream);
int32_t arw64_set_w_surface_adjoint(const arw64_domain *d, const arw_params *prm,
                                  const double *lam_w0, double *g_u, double *g_v,
                                  void *stream);
int32_t arw64_scalar_stage(const arw64_domain *d, const arw_params *prm,
                         double dt_stage, int32_t clamp, const double *q,
                         const double *q0, const double *ru, const double *rv,
                         const double *ww, const double *mup_t,
                         const double *mup_new, double *q_new, void *stream);
int32_t arw64_scalar_stage_adjoint(const arw64_domain *d, const arw_params *prm,
                                 double dt_stage, int32_t clamp,
                                 const double *q, const double *q0, const double *ru,
                                 const double *rv, const double *ww,
                                 const double *mup_t, const double *mup_new,
                                 const double *lam, double *g_q, double *g_q0,
                                 double *g_ru, double *g_rv, double *g_ww,
                                 double *g_mup_t, double *g_mup_new, double *work,
                                 void *stream);
int32_t arw64_pd_stage(const arw64_domain *d, const arw_params *prm,
                       double dt_stage, const double *q, const double *q0,
                       const double *ru, const double *rv, const double *ww,
                       const double *mup_t, const double *mup_new,
                       double *q_new, double *work, void *stream);
int32_t arw64_pd_stage_adjoint(const arw64_domain *d, const arw_params *prm,
                               double dt_stage, const double *q,
                               const double *q0, const double *ru,
                               const double *rv, const double *ww,
                               const double *mup_t, const double *mup_new,
                               const double *lam, double *g_q, double *g_q0,
                               double *g_ru, double *g_rv, double *g_ww,
                               double *g_mup_t, double *g_mup_new,
                               double *work, void *stream);

typedef struct arw64_stage_adj {
    const double *l_ru_t, *l_rv_t, *l_rw_t, *l_rth_t, *l_rph_t;
    double *g_u, *g_v, *g_w, *g_thp, *g_php, *g_mup;
    double *g_qv, *g_qc, *g_qr;     /* moist states: the stage scalars' */
    const double *l_ru, *l_rv, *l_ww; /* optional: adjoints of the stage fluxes */
    int32_t flags;   /* ARW_ADJ_ACCUMULATE: add to g_* instead of writing */
} arw64_stage_adj;
int32_t arw64_stage_tendencies_adjoint(const arw64_domain *d,
                                       const arw_params *prm,
                                       const arw64_stage_adj *adj,
                                       double *work, void *stream);
int32_t arw64_stage_adjoint_work(int32_t nz, int32_t ny, int32_t nx,
                                 int64_t *elements);

typedef struct arw64_init_adj {
    const double *l_u_pp, *l_v_pp, *l_w_pp, *l_th_pp, *l_ph_pp, *l_mu_pp,
        *l_al_pp, *l_p_pp, *l_p_pp_old;
    double *g_u0, *g_v0, *g_w0, *g_thp0, *g_php0, *g_mup0;
    double *g_u, *g_v, *g_w, *g_thp, *g_php, *g_mup;
    double *g_qv;                   /* moist states: the stage qv's */
    int32_t flags;   /* ARW_ADJ_ACCUMULATE: add to g_* instead of writing */
} arw64_init_adj;
typedef struct arw64_finish_adj {
    const double *l_u, *l_v, *l_w, *l_thp, *l_php, *l_mup;
    double *g_u, *g_v, *g_w, *g_thp, *g_php, *g_mup;
    double *g_u_pp, *g_v_pp, *g_w_pp, *g_th_pp, *g_ph_pp, *g_mu_pp;
} arw64_finish_adj;
int32_t arw64_small_step_init_adjoint(const arw64_domain *d,
                                      const arw64_init_adj *adj,
                                      double *work, void *stream);
int32_t arw64_small_step_finish_adjoint(const arw64_domain *d,
                                        const arw64_finish_adj *adj,
                                        double *work, void *stream);
int32_t arw64_small_step_adjoint_work(int32_t nz, int32_t ny, int32_t nx,
                                      int64_t *elements);

typedef struct arw64_ensemble arw64_ensemble;
int32_t arw64_ensemble_create(const arw64_domain *members, int32_t n,
                              const arw_params *prm, const int32_t *order,
                              int32_t groups, int32_t col_tile,
                              arw64_ensemble **out);
int32_t arw64_ensemble_step(arw64_ensemble *e, int32_t nsteps,
                            void *stream);
int32_t arw64_ensemble_capture(arw64_ensemble *e, int32_t nsteps,
                               int32_t join_every);
int32_t arw64_ensemble_launch(arw64_ensemble *e, void *stream);
int32_t arw64_ensemble_destroy(arw64_ensemble *e);

#ifdef __cplusplus
}
#endif

#endif /* ARWEN_DYCORE_H */
