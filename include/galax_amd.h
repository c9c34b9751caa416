/*
 * galax_amd.h -- C ABI of libgalaxhip.so, the MI355X (gfx950) orbit engine.
 *
 * Units everywhere are galax's "galactic" system: kpc, Myr, Msun, rad.
 * G = 4.498502151469553e-12 kpc^3 Msun^-1 Myr^-2.
 *
 * Conventions (all entry points):
 *   - every pointer argument that names an array is a DEVICE pointer owned by
 *     the caller (torch.Tensor.data_ptr() in the Python layer), except the
 *     component/track tables handed to gx_potential_create, which are host
 *     arrays copied into the handle;
 *   - `stream` is a hipStream_t (NULL = legacy default stream); every compute
 *     call is asynchronous on it and returns after the launch;
 *   - return value 0 = ok, < 0 = argument or launch error (see GX_E*); nothing
 *     throws across the ABI;
 *   - handles are immutable after creation: concurrent calls that share a
 *     handle are safe as long as they use distinct output buffers.
 *
 * Each entry point names the reference interface it replaces (paths are
 * relative to the galax source tree, src/galax/...).
 */
#ifndef GALAX_AMD_H
#define GALAX_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- errors */
#define GX_OK 0
#define GX_EINVAL (-1)   /* bad argument (null pointer, bad size, bad enum) */
#define GX_ELAUNCH (-2)  /* HIP launch / runtime error */
#define GX_ENOMEM (-3)   /* device allocation failed (create calls only) */
#define GX_ETOOBIG (-4)  /* more than a fixed-capacity entry point holds */

/* ------------------------------------------------------------ components */
/* Potential models (potential/_src/builtin/...).  Parameter block p[]:
 *   GX_KEPLER          p0 = G*m_tot                         kepler.py:34-94
 *   GX_HERNQUIST       p0 = G*m_tot, p1 = r_s               hernquist.py:27-82
 *   GX_MIYAMOTO_NAGAI  p0 = G*m_tot, p1 = a, p2 = b         miyamotonagai.py:30-78
 *   GX_NFW             p0 = G*m,     p1 = r_s               nfw/base.py:36-100,327-338
 *   GX_PLUMMER         p0 = G*m_tot, p1 = r_s               plummer.py:33-132
 *   GX_MONARI_BAR      p0 = alpha, p1 = R0, p2 = v0 [kpc/Myr], p3 = Rb,
 *                      p4 = phi_b [rad], p5 = Omega [rad/Myr] monari2016.py:29-138
 *   GX_NULL            no parameters (Phi = 0)              null.py
 *   GX_ISOCHRONE       p0 = G*m_tot, p1 = r_s               isochrone.py:61-90
 *   GX_JAFFE           p0 = G*m_tot, p1 = r_s               jaffe.py:34-60
 *   GX_LOGARITHMIC     p0 = v_c [kpc/Myr], p1 = r_s         logarithmic.py:19-53
 *   GX_LMJ09_LOG       p0 = v_c, p1 = r_s, p2..p4 = q1, q2, q3 (axis ratios),
 *                      p5 = phi [rad] (rotation about z)     logarithmic.py:56-108
 *   GX_LONGMURALI_BAR  p0 = G*m_tot, p1 = a, p2 = b, p3 = c,
 *                      p4 = alpha [rad] (bar angle)          longmurali.py:19-91
 *   GX_POWERLAW_CUTOFF p0 = G*m_tot, p1 = alpha (0 <= alpha < 2),
 *                      p2 = r_c                              powerlawcutoff.py:31-117
 *   GX_BURKERT         p0 = G*m, p1 = r_s                    burkert.py:150-230
 *   GX_KUZMIN          p0 = G*m_tot, p1 = r_s (|z| in Phi)   kuzmin.py:60-85
 *   GX_SATOH           p0 = G*m_tot, p1 = a, p2 = b          satoh.py:50-69
 *   GX_STONE           p0 = G*m_tot, p1 = r_c, p2 = r_h      stoneostriker15.py:90-147
 *   GX_TRIAXIAL_HERNQUIST p0 = G*m_tot, p1 = r_s, p2 = q1, p3 = q2
 *                      (Phi of r' = sqrt(x^2 + (y/q1)^2 + (z/q2)^2))
 *                                                            hernquist.py:89-182
 *   GX_TRIAXIAL_NFW    p0 = G*m, p1 = r_s, p2 = q1, p3 = q2 (density axis
 *                      ratios), p4 = Gauss-Legendre order (1..128; the
 *                      reference's default is 50)            nfw/triaxial.py:29-230
 *   GX_LEESUTO_NFW     p0 = G*m, p1 = r_s, p2..p4 = a1 >= a2 >= a3
 *                                                            nfw/leesuto.py:29-164
 *   GX_VOGELSBERGER_NFW p0 = G*m, p1 = r_s, p2 = q1 (q2^2 = 3 - q1^2),
 *                      p3 = a_r                              nfw/vogelsburger08.py:27-110
 *   GX_HARMONIC_OSCILLATOR p0 = omega [1/Myr] (Phi = omega^2 r^2 / 2)
 *                                                            example.py:17-100
 *   GX_HENON_HEILES    p0 = coeff [1/kpc], p1 = timescale [Myr]
 *                      (Phi = (R^2/2 + coeff (x^2 y - y^3/3)) / t_s^2)
 *                                                            example.py:103-176
 *   GX_HARDCUTOFF_NFW  p0 = G*m, p1 = r_s, p2 = r_t > 0 (NFW inside r_t,
 *                      point mass M_NFW(<r_t) outside)       nfw/truncated.py:37-330
 *   GX_GNFW            p0 = G*m, p1 = r_s, p2 = gamma in [0, 2) (inner slope)
 *                                                            nfw/generalized.py:30-330
 *   GX_GAUSSIAN        p0 = G*m_tot, p1 = r_s (Phi = -G m erf(r/(sqrt2 r_s))/r)
 *                                                            gaussian/base.py:40-259
 *   GX_TRIAXIAL_GAUSSIAN p0 = G*m_tot, p1 = r_s, p2 = q1, p3 = q2 (density
 *                      axis ratios; q1 = 1 is the axisymmetric model),
 *                      p4 = Gauss-Legendre order (1..128, default 50)
 *                                      gaussian/triaxial.py:26-240, axisymmetric.py
 */
enum gx_kind {
  GX_NULL = 0,
  GX_KEPLER = 1,
  GX_HERNQUIST = 2,
  GX_MIYAMOTO_NAGAI = 3,
  GX_NFW = 4,
  GX_PLUMMER = 5,
  GX_MONARI_BAR = 6,
  GX_ISOCHRONE = 7,
  GX_JAFFE = 8,
  GX_LOGARITHMIC = 9,
  GX_LMJ09_LOG = 10,
  GX_LONGMURALI_BAR = 11,
  GX_POWERLAW_CUTOFF = 12,
  GX_BURKERT = 13,
  GX_KUZMIN = 14,
  GX_SATOH = 15,
  GX_STONE = 16,
  GX_TRIAXIAL_HERNQUIST = 17,
  GX_TRIAXIAL_NFW = 18,
  GX_LEESUTO_NFW = 19,
  GX_VOGELSBERGER_NFW = 20,
  GX_HARMONIC_OSCILLATOR = 21,
  GX_HENON_HEILES = 22,
  GX_HARDCUTOFF_NFW = 23,
  GX_GNFW = 24,
  GX_GAUSSIAN = 25,
  GX_TRIAXIAL_GAUSSIAN = 26
};

/* Where the component sits (potential/_src/xfm/translate.py:29-199 —
 * TranslatedPotential evaluates Phi_base(x - Delta(t))).  Frame block f[]:
 *   GX_FRAME_NONE      centred on the origin.
 *   GX_FRAME_OFFSET    Delta = (f0, f1, f2), constant.
 *   GX_FRAME_ROTATING  Delta(t) = Rz(f3 * t) . (f0, f1, f2): a centre fixed in
 *                      a frame rotating rigidly about z at f3 rad/Myr (a
 *                      circular orbit when f2 = 0).
 *   GX_FRAME_TRACK     Delta(t) = cubic Hermite interpolant of track
 *                      number `track` (positions and velocities at knots). */
enum gx_frame {
  GX_FRAME_NONE = 0,
  GX_FRAME_OFFSET = 1,
  GX_FRAME_ROTATING = 2,
  GX_FRAME_TRACK = 3
};

typedef struct gx_component {
  int32_t kind;     /* enum gx_kind */
  int32_t frame;    /* enum gx_frame */
  int32_t track;    /* track index when frame == GX_FRAME_TRACK, else 0 */
  int32_t reserved; /* must be 0 */
  double p[8];      /* model parameters (see gx_kind) */
  double f[6];      /* frame parameters (see gx_frame) */
} gx_component;

/* A centre trajectory, in one of two forms.
 * Knots (rec == NULL): n knots (1 <= n <= INT32_MAX), strictly increasing
 *   t[n], positions q[n*3] and velocities v[n*3] (row-major), interpolated
 *   by cubic Hermite polynomials.
 * Dense solution (rec != NULL): the dense output of one integrated orbit,
 *   n accepted-step records (1 <= n <= INT32_MAX) of GX_DENSE_RECORD
 *   doubles in integration order, as gx_integrate_dense writes them for one
 *   orbit, from method `method` (gx_method); consecutive records must be
 *   contiguous in time (each starts where the previous one ends).  The
 *   position at t is the method's dense output over the step holding t
 *   (diffrax's DenseInterpolation, which Orbit.interpolant evaluates); t,
 *   q and v are ignored.
 * Outside the covered interval the end position is held. */
typedef struct gx_track {
  int64_t n;
  const double* t;
  const double* q;
  const double* v;
  const double* rec;
  int32_t method;
  int32_t reserved; /* 0 */
} gx_track;

typedef struct gx_pot gx_pot; /* opaque handle; lives on one device */

/* Build a device-resident potential from a host component table.
 * Replaces the construction of a CompositePotential / AbstractPotential
 * pytree (potential/_src/composite.py:33-111, base_multi.py:39-82): the
 * composite is the sum of its components, evaluated in table order.
 * Host arrays; copied.  At most GX_MAX_COMPONENTS components. */
#define GX_MAX_COMPONENTS 256
#define GX_MAX_TRACKS 8
#define GX_MAX_GL_ORDER 128 /* GX_TRIAXIAL_NFW quadrature order limit */
int gx_potential_create(const gx_component* comps, int32_t ncomp,
                        const gx_track* tracks, int32_t ntracks,
                        gx_pot** out);
int gx_potential_destroy(gx_pot* pot);

/* ----------------------------------------------------------- field calls */
/* Positions xyz[n*3] (row-major), times t[n] when t_stride == 1 or a single
 * time t[0] when t_stride == 0.  Outputs row-major.
 *
 * gx_acceleration  replaces AbstractPotential.acceleration /
 *                  gp.acceleration (potential/_src/base.py:250-255,
 *                  register_funcs.py:327-344), i.e. -grad Phi.
 * gx_gradient      replaces AbstractPotential._gradient (base.py:170-179).
 * gx_potential_eval replaces AbstractPotential._potential / gp.potential
 *                  (base.py:113-136).
 * gx_hessian       replaces AbstractPotential._hessian (base.py:230-239);
 *                  out[n*9] row-major 3x3. */
int gx_acceleration(const gx_pot* pot, const double* xyz, const double* t,
                    int32_t t_stride, int64_t n, double* acc, void* stream);
int gx_gradient(const gx_pot* pot, const double* xyz, const double* t,
                int32_t t_stride, int64_t n, double* grad, void* stream);
int gx_potential_eval(const gx_pot* pot, const double* xyz, const double* t,
                      int32_t t_stride, int64_t n, double* phi, void* stream);
int gx_hessian(const gx_pot* pot, const double* xyz, const double* t,
               int32_t t_stride, int64_t n, double* hess, void* stream);

/* ----------------------------------------------------------- integration */
enum gx_method { GX_DOPRI8 = 0, GX_DOPRI5 = 1 };
enum gx_controller { GX_CONSTANT = 0, GX_PID = 1 };

/* Solver configuration: mirrors OrbitSolver(solver, stepsize_controller,
 * max_steps) (dynamics/_src/orbit/solver.py:121-141) with diffrax's
 * PIDController / ConstantStepSize fields.  Unset optional bounds use
 * dtmin = 0, dtmax = +inf.  dt0 <= 0 selects the automatic initial step. */
typedef struct gx_solver_cfg {
  int32_t method;      /* enum gx_method */
  int32_t controller;  /* enum gx_controller */
  int64_t max_steps;   /* attempted-step cap per orbit; <= 0 = "unbounded",
                          which is a safety cap of 10,000,000 */
  double dt0;          /* first step (required for GX_CONSTANT) */
  double rtol, atol;
  double pcoeff, icoeff, dcoeff;
  double safety, factormin, factormax;
  double dtmin, dtmax;
  int32_t force_dtmin; /* accept a step taken at dtmin even if it fails */
  int32_t reserved;
} gx_solver_cfg;

/* Per-orbit status, mirroring diffrax RESULTS. */
enum gx_status {
  GX_SUCCESS = 0,
  GX_MAX_STEPS_REACHED = 1,
  GX_DT_MIN_REACHED = 2,
  GX_NONFINITE = 3
};

/* How save times are given.
 *   GX_TS_SHARED    ts[T], the same for every orbit (evaluate_orbit's `t`);
 *   GX_TS_PER_ORBIT ts[n*T] row-major (experimental integrate_orbit's (B,T)
 *                   saveat, dynamics/_src/experimental/integrate.py:446-570);
 *   GX_TS_LINSPACE  ts ignored; orbit i saves at linspace(t0[i], t1[i], T);
 *   GX_TS_FINAL     ts ignored, T must be 1; save the state at t1[i]. */
enum gx_ts_mode {
  GX_TS_SHARED = 0,
  GX_TS_PER_ORBIT = 1,
  GX_TS_LINSPACE = 2,
  GX_TS_FINAL = 3
};

/* Integrate n independent orbits of Hamilton's equations dq/dt = p,
 * dp/dt = -grad Phi(q, t) (dynamics/_src/orbit/field_hamiltonian.py:240-249)
 * from t0[i] to t1[i] (t_stride 0: one shared pair t0[0], t1[0]), starting
 * from w0[n*6] = (x,y,z,vx,vy,vz), writing out[n*T*6] at the save times.
 * Each orbit takes its own step sequence (the batched-t0/t1 semantics of
 * orbit/solver.py:854-1039).  The state at a save time inside an accepted
 * step is the method's dense output over that step (diffrax's interpolant
 * for SaveAt(ts), legacy/integrator.py:227-230).  With no dt0, a PID
 * controller that sets dtmin starts from a first step of dtmin; otherwise
 * the automatic first step is diffrax's.  Replaces the diffeqsolve call
 * inside OrbitSolver.solve
 * (orbit/solver.py:774-1039) and thus the hot loop of gd.evaluate_orbit
 * (legacy/funcs.py:42-213).
 *
 * Optional outputs (may be NULL): status[n] (enum gx_status), nsteps[n]
 * (attempted steps), naccept[n] (accepted steps).
 * `order` (optional, may be NULL) is a permutation of 0..n-1 giving the order
 * in which the persistent work queue hands out orbits; longest-first order
 * shortens the tail of a launch of uneven orbits.
 * `workspace` must point to at least gx_integrate_workspace_bytes(n) bytes of
 * device memory (256-byte aligned) that no concurrent call uses; it holds the
 * work-queue counter and each orbit's first step and start acceleration. */
int64_t gx_integrate_workspace_bytes(int64_t n);
int gx_integrate(const gx_pot* pot, const gx_solver_cfg* cfg,
                 const double* w0, const double* t0, const double* t1,
                 int32_t t_stride, const double* ts, int32_t T,
                 int32_t ts_mode, int64_t n, const int64_t* order,
                 double* out, int32_t* status, int64_t* nsteps,
                 int64_t* naccept, void* workspace, void* stream);

/* gx_integrate without a workspace argument and without any allocation, for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops).  No counterpart in the reference.  The workspace is
 * a fixed 4 MiB array in device memory, one per GPU, held until the process
 * ends, so n is at most gx_integrate_noalloc_max_orbits() (131,064); larger
 * n returns GX_ETOOBIG.  Every call on a GPU shares that GPU's array: calls
 * that overlap on one GPU (other streams, other host threads) race, so a
 * caller serialises them.  The first call on a device looks the array up
 * (hipGetSymbolAddress); no call allocates.  Otherwise as gx_integrate. */
int64_t gx_integrate_noalloc_max_orbits(void);
int gx_integrate_noalloc(const gx_pot* pot, const gx_solver_cfg* cfg,
                         const double* w0, const double* t0, const double* t1,
                         int32_t t_stride, const double* ts, int32_t T,
                         int32_t ts_mode, int64_t n, const int64_t* order,
                         double* out, int32_t* status, int64_t* nsteps,
                         int64_t* naccept, void* stream);

/* Dense solutions (diffrax's dense interpolation, SaveAt(dense=True); the
 * Orbit interpolant of legacy/funcs.py:210 and orbit/interp.py:42-145).
 * gx_integrate_dense is gx_integrate that also writes, for every accepted
 * step of orbit i, a record of GX_DENSE_RECORD doubles at
 * rec[(i * rec_cap + k) * GX_DENSE_RECORD]: the step's start time t and
 * signed step h (physical time), q and v there, and its stage accelerations
 * (13 stages and the end stage for Dopri8, 7 for Dopri5).  nrec[i] receives
 * the records written; an orbit whose accepted steps outnumber rec_cap stops
 * with GX_MAX_STEPS_REACHED.  gx_dense_eval evaluates the method's dense
 * output from such records at times t (per_orbit 0: t[T] for every orbit;
 * 1: t[n*T]) into out[n*T*6]; a time outside an orbit's integrated interval
 * gives NaN.  No separate reference entry point: diffrax keeps the same
 * per-step data in its DenseInterpolation. */
#define GX_DENSE_RECORD 50
int gx_integrate_dense(const gx_pot* pot, const gx_solver_cfg* cfg,
                       const double* w0, const double* t0, const double* t1,
                       int32_t t_stride, const double* ts, int32_t T,
                       int32_t ts_mode, int64_t n, const int64_t* order,
                       double* out, int32_t* status, int64_t* nsteps,
                       int64_t* naccept, double* rec, int64_t rec_cap,
                       int64_t* nrec, void* workspace, void* stream);
int gx_dense_eval(int32_t method, const double* rec, int64_t rec_cap,
                  const int64_t* nrec, int64_t n, const double* t,
                  int32_t per_orbit, int32_t T, double* out, void* stream);
/* The same records from a shared-step solve (gx_integrate_shared below):
 * every orbit's records share the batch's steps. */
int gx_integrate_shared_dense(const gx_pot* pot, const gx_solver_cfg* cfg,
                              const double* w0, double t0, double t1,
                              const double* ts, int32_t T, int64_t n,
                              double* out, int32_t* status, int64_t* nsteps,
                              int64_t* naccept, double* rec, int64_t rec_cap,
                              int64_t* nrec, void* workspace, void* stream);

/* Integrate n orbits as ONE system with a shared step sequence: galax's
 * Integrator with scalar start and end times solves a batch (B, 6) as one
 * ODE (legacy/integrator.py:269-355, orbit/solver.py:774-812; the second
 * stage of gd.evaluate_orbit, legacy/funcs.py:42-213) -- one step size, one
 * accept/reject decision per step from the RMS error over all 6n components,
 * the automatic first step from norms over the whole batch.  From t0 to t1,
 * saving at ts[T] (ordered in the direction of integration).  Under
 * GX_CONSTANT the result equals gx_integrate's.
 *
 * status[n], nsteps[n], naccept[n] (optional) receive the batch's values for
 * every orbit.  `workspace`: gx_integrate_shared_workspace_bytes(n) bytes of
 * device memory (256-byte aligned) that no concurrent call uses.
 * Synchronous: when every workgroup of the batch can be resident at once,
 * one cooperative launch runs the whole solve (a grid barrier per attempted
 * step); otherwise the host launches one kernel per attempted step, in
 * batches, reading a 4-byte flag after each batch.  Either way the call
 * returns when the solve has ended; the results are then on `stream`. */
int64_t gx_integrate_shared_workspace_bytes(int64_t n);
int gx_integrate_shared(const gx_pot* pot, const gx_solver_cfg* cfg,
                        const double* w0, double t0, double t1,
                        const double* ts, int32_t T, int64_t n, double* out,
                        int32_t* status, int64_t* nsteps, int64_t* naccept,
                        void* workspace, void* stream);

/* ------------------------------------------------------------- adjoint */
/* Discrete adjoint of fixed-step (GX_CONSTANT) Dopri8/Dopri5 integration of
 * n independent orbits, as gx_integrate would run them with the same
 * arguments.  Given cotangents cot[n*T*6] of the saved states, returns the
 * cotangents of the initial states w0_bar[n*6] and, per orbit, of the
 * potential's table parameters theta_bar[n*ntheta] laid out as
 * (p0, p1, p2) of each component in table order (p0 = G m; p1, p2 the
 * component's lengths, e.g. Miyamoto-Nagai a, b; Hernquist/NFW/Plummer
 * r_s).  Supported potentials: the Milky Way composition (Miyamoto-Nagai,
 * NFW, Hernquist, Hernquist, untranslated) and any single untranslated
 * Kepler/Hernquist/Miyamoto-Nagai/NFW/Plummer component; ntheta <= 12.
 * State is checkpointed every K steps; max_steps bounds the steps of any
 * orbit (it sizes the workspace, gx_adjoint_workspace_bytes(n, max_steps,
 * K) bytes, 256-byte aligned).  status[n] (optional) is GX_SUCCESS or
 * GX_MAX_STEPS_REACHED.  Replaces differentiating evaluate_orbit /
 * MockStreamGenerator.run with jax (dynamics/_src/orbit/solver.py:133,
 * tests/functional/test_mockstreamgenerator.py:24-116). */
int64_t gx_adjoint_workspace_bytes(int64_t n, int64_t max_steps, int32_t K);
int gx_adjoint(const gx_pot* pot, const gx_solver_cfg* cfg, const double* w0,
               const double* t0, const double* t1, int32_t t_stride,
               const double* ts, int32_t T, int32_t ts_mode, int64_t n,
               const double* cot, double* w0_bar, double* theta_bar,
               int32_t ntheta, int64_t max_steps, int32_t K, int32_t* status,
               void* workspace, void* stream);

/* Launch shape of the persistent integrator: resident workgroups per CU
 * that a launch may take (capped by the kernel's occupancy; 0 keeps the
 * current value).  `block` must be 0 or 256 (the kernel's fixed workgroup
 * size).  Returns the previous values through the pointers when they are
 * non-NULL.  A process-wide setting: not thread-safe with concurrent
 * launches.  No reference counterpart (a tuning control). */
int gx_set_launch_shape(int32_t block, int32_t blocks_per_cu,
                        int32_t* prev_block, int32_t* prev_blocks_per_cu);

/* Each persistent integrator launch takes 1/parts of the device's resident
 * capacity (parts >= 1; 0 keeps the current value; default 1).  A launch is
 * bounded by its longest orbit's serial step chain, during which most of
 * the device falls idle; a caller that keeps `parts` independent batches in
 * flight on as many streams runs them side by side.  The hardware queues of
 * a process bound the useful value (4 by default).  Returns the previous
 * value through `prev_parts` when non-NULL.  Process-wide, like
 * gx_set_launch_shape.  No reference counterpart (a tuning control). */
int gx_set_launch_fraction(int32_t parts, int32_t* prev_parts);

/* The persistent integrator's work-queue policy.  head_lanes (>= 0; < 0
 * keeps the current value; default 4096): in a launch of at least four
 * times as many lanes, the first head_lanes lanes start on the first
 * head_lanes tickets (the longest orbits of a longest-first order), run at
 * top wave priority and take no second ticket; 0 turns this off.  chunk
 * (1..64; 0 keeps; default 1): tickets a lane takes per atomic.  Results do
 * not depend on either.  No reference counterpart (a tuning control, set by
 * galax_amd.tune). */
int gx_set_queue_policy(int64_t head_lanes, int32_t chunk,
                        int64_t* prev_head_lanes, int32_t* prev_chunk);

/* Library and device information for tooling: number of compute units of
 * the current device, and the VGPR count of the integrator kernels. */
int gx_device_info(int32_t* num_cu, int32_t* warp_size);

/* ------------------------------------------------------------- surrogate */
/* The learned acceleration surrogate has no counterpart in the reference:
 * galax evaluates every field exactly (potential/_src/base.py:170-179).
 * These entry points let a caller train an MLP on the perturbing part of a
 * field (here a bar plus co-rotating subhaloes) and integrate with it.
 * Matrices are row-major BF16 (uint16 storage) unless stated; all pointers
 * are device pointers; every call is asynchronous on `stream`.
 *
 * gx_mlp_gemm: C(m, n) = sum_k A(m, k) B(k, n) on the BF16 matrix cores with
 * FP32 accumulation, where A(m, k) = A[m*sam + k*sak] and B(k, n) =
 * B[k*sbk + n*sbn].  epi selects the epilogue:
 *   GX_EPI_STORE      C[m*ldc + n] = acc (+= if accumulate; with ksplit > 1
 *                     the K range is split over ksplit workgroups that add
 *                     atomically, so C must hold the initial value);
 *   GX_EPI_BIAS       C[m*ldc + n] = acc + bias[n];
 *   GX_EPI_BIAS_SILU  z = acc + bias[n]; C[m*ldc + n] = z (if C != NULL);
 *                     H[m*ldh + n] = bf16(silu(z));
 *   GX_EPI_DSILU      H[m*ldh + n] = bf16(acc * silu'(Zin[m*ldzin + n])). */
#define GX_EPI_STORE 0
#define GX_EPI_BIAS 1
#define GX_EPI_BIAS_SILU 2
#define GX_EPI_DSILU 3
int gx_mlp_gemm(const void* A, int64_t sam, int64_t sak, const void* B,
                int64_t sbk, int64_t sbn, int64_t M, int64_t N, int64_t K,
                int32_t epi, int32_t accumulate, int32_t ksplit, const float* bias,
                float* C, int64_t ldc, void* H, int64_t ldh, const float* Zin,
                int64_t ldzin, void* stream);
/* The MLP's whole inference pass in one kernel, for hidden width 256:
 * Y[m*ldy + o] = layer_depth(... silu(layer_0(X[m*ldx + 0..nin-1]))), with
 * W[l] the BF16 weights [out][in] (W[0]: 256 x nin; W[1..depth-1]:
 * 256 x 256; W[depth]: nout x 256; the last two 16-byte aligned) and b[l]
 * the FP32 biases.  Activations stay on chip (rounded to BF16 between
 * layers, as the layer-by-layer gx_mlp_gemm path rounds them).  W and b are
 * host arrays of depth + 1 device pointers; nin <= 8, nout <= 16,
 * depth <= 8. */
int gx_mlp_forward_fused(const void* X, int32_t ldx, int32_t nin, int64_t n,
                         int32_t depth, const void* const* W,
                         const float* const* b, int32_t nout, float* Y,
                         int32_t ldy, void* stream);
/* out[n] += sum_m dz[m*ld + n] (bias gradients; out must hold the start). */
int gx_mlp_colsum(const void* dz, int64_t ld, int64_t M, int64_t N, float* out,
                  void* stream);
/* One bias-corrected Adam step (step >= 1) on FP32 parameters p with
 * gradient g and moments m, v; pb (optional) receives the BF16 copy. */
int gx_mlp_adam(float* p, const float* g, float* m, float* v, void* pb,
                int64_t n, float lr, float b1, float b2, float eps, int64_t step,
                void* stream);
/* Pseudo-Huber loss of Y[:, 0:3] (row stride ldy) against T[B][3]:
 * *loss += sum delta^2 (sqrt(1 + (r/delta)^2) - 1) (loss must hold the
 * start), dY[:, 0:3] = bf16(scale * r / sqrt(1 + (r/delta)^2)), other
 * columns of dY (row stride ldd) zero. */
int gx_mlp_huber(const float* Y, int64_t ldy, const float* T, int64_t B, float delta,
                 float scale, void* dY, int64_t ldd, double* loss, void* stream);
/* Network inputs at points xyz[n][3] and times t[0] + dt (t_stride 0) or
 * t[i] + dt (t_stride 1): X[i] = bf16((x', y', z, R, r) / L, 0...) with
 * (x', y') the position rotated by -omega t, rows of width ldx >= 5. */
int gx_sur_inputs(const double* xyz, const double* t, int32_t t_stride, double dt,
                  int64_t n, double inv_len, double omega, void* X, int64_t ldx,
                  void* stream);
/* Training targets T[i] = float(Rz(-omega t_i) (a_full - a_base)_i / scale). */
int gx_sur_targets(const double* t, int32_t t_stride, int64_t n, const double* a_full,
                   const double* a_base, double omega, double inv_scale, float* T,
                   void* stream);
/* Surrogate field: acc[i] = a_base(x_i) + scale Rz(omega t) Y[i][0:3], or
 * the exact field of `full` where |x_i| > r_max or flag[i] != 0 (flag may be
 * NULL).  `base` must be an untranslated table without bars. */
int gx_sur_combine(const gx_pot* base, const gx_pot* full, const double* xyz,
                   const double* t, int32_t t_stride, double dt, int64_t n,
                   const float* Y, int64_t ldy, double scale, double omega,
                   double r_max, const int32_t* flag, double* acc, void* stream);
/* Gate check: flag[i] = 1 (sticky) where |a_sur - a_full| / |a_full| > tol
 * at xyz[i]; err[i] (optional) receives the relative error. */
int gx_sur_check(const gx_pot* full, const double* xyz, const double* t,
                 int32_t t_stride, int64_t n, const double* a_sur, double tol,
                 int32_t* flag, float* err, void* stream);
/* Stage nodes c_i of a method (time of stage i = t + c_i h): writes up to
 * cap of them to c (host memory, may be NULL) and returns the stage count. */
int gx_method_nodes(int32_t method, double* c, int32_t cap);
/* Lock-step fixed-step Dopri8 for batched fields: stage positions
 * Q = q + h (rs_stage p + h sum_l (A^2)_{stage,l} G[l]) from the stage
 * accelerations G[13][n][3]; and the step update of q, p (and t[0] += h). */
int gx_lockstep_stage(int32_t stage, const double* q, const double* p,
                      const double* G, int64_t n, double h, double* Q, void* stream);
int gx_lockstep_update(double* q, double* p, const double* G, int64_t n, double h,
                       double* t, void* stream);

/* ----------------------------------------------------------------- N-body */
/* Direct-summation N-body field with an external potential, replacing
 * NBodyField (dynamics/_src/orbit/field_nbody.py:94-131): for bodies at
 * xyz[n][3] with gm[n] = G m_i and softening eps,
 *   acc_j = sum_{i != j} gm_i (x_i - x_j) / (|x_i - x_j|^2 + eps^2)^(3/2)
 *           - grad Phi_ext(x_j, t).
 * gx_nbody_acceleration evaluates it at one time t (any n). */
int gx_nbody_acceleration(const gx_pot* ext, const double* xyz, const double* gm,
                          double eps, double t, int64_t n, double* acc, void* stream);
/* Integrates the n-body system (1 <= n <= 4096) as ONE ODE system, as
 * OrbitSolver.solve(NBodyField, ...) does (orbit/solver.py:121-141): shared
 * steps, the PID error norm the RMS over all 6n state components, saves at
 * ts[T] (monotone in the direction t0 -> t1), out[n][T][6].  One status,
 * step and accepted-step count for the system.  workspace holds
 * gx_nbody_workspace_bytes(n) bytes. */
int64_t gx_nbody_workspace_bytes(int64_t n);
int gx_nbody_integrate(const gx_pot* ext, const double* gm, double eps,
                       const gx_solver_cfg* cfg, const double* w0, double t0, double t1,
                       const double* ts, int32_t T, int64_t n, double* out, int32_t* status,
                       int64_t* nsteps, int64_t* naccept, void* workspace, void* stream);

/* Accuracy check of the library's short FP64 primitives (no counterpart in
 * the reference; for tests): out[9i .. 9i+8] at x = x[i] are 1/x,
 * 1/sqrt(x) (x > 0), log1p(x) (x > -1), 1/x by the division routine,
 * sqrt(x), x/6, x^(1/9), x^(1/6) (x >= 0) and x^(1/8) by three reciprocal
 * square roots (x > 0). */
#define GX_MATH_CHECK_OUTPUTS 9
int gx_math_check(const double* x, int64_t n, double* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* GALAX_AMD_H */
