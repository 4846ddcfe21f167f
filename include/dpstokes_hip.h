/*
 * dpstokes_hip.h -- C ABI of the MI355X (gfx950) doubly periodic Stokes
 * mobility solver.
 *
 * The entry points replace the GPU backend that the reference's joint
 * interface binds:
 *   reference pybind module `uammd` (source/gpu/python_wrapper/uammd_python.cpp:59-129)
 *   -> C++ glue DPStokesGlue        (source/gpu/python_wrapper/uammd_interface.h:43-65)
 *   -> DPStokesUAMMD                (source/gpu/python_wrapper/uammd_wrapper.cu:90-177)
 * and the CPU backend's FCM class (source/cpu/python/FCM.py:149-441) it sits
 * beside.  Unlike the reference glue (std::string, shared_ptr, exceptions)
 * this is a plain C ABI: POD structs, device pointers, sizes and an int
 * status; no exception or exit() crosses it.
 *
 * Conventions
 *   - All particle arrays are interleaved [x0,y0,z0,x1,...] device pointers
 *     (hipMalloc'ed or torch-owned), of the dtype named by `dtype`.
 *   - `stream` is a hipStream_t passed as void*; 0 means the null stream.
 *     Every call is asynchronous on that stream unless stated otherwise.
 *   - Positions: x and y are periodic (wrapped into [0,L) as the reference's
 *     ParticleList::update does, src/ParticleList.cpp:561-580); z must lie in
 *     [zmin, zmax] or DPS_ERR_RANGE is reported by dps_set_positions.
 */
#ifndef DPSTOKES_HIP_H
#define DPSTOKES_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define DPS_OK 0
#define DPS_ERR_ARG (-1)        /* invalid argument / parameter */
#define DPS_ERR_STATE (-2)      /* handle not initialised, positions not set */
#define DPS_ERR_RANGE (-3)      /* a particle is outside the z extent */
#define DPS_ERR_HIP (-4)        /* HIP runtime error */
#define DPS_ERR_FFT (-5)        /* rocFFT error */
#define DPS_ERR_ALLOC (-6)      /* device allocation failed */

/* geometry; mode strings of uammd_wrapper.cu:55-66 */
#define DPS_MODE_NOWALL 1
#define DPS_MODE_BOTTOM 2
#define DPS_MODE_SLIT 3
#define DPS_MODE_PERIODIC 4     /* triply periodic: uniform z, nz = Lz / hz points (Solvers.py:197-255) */

/* arithmetic */
#define DPS_PREC_DOUBLE 0       /* everything in FP64 */
#define DPS_PREC_MIXED 1        /* FP32 grids/FFT/z-transform, FP64 BVP + corrections */

/* spectral solve */
#define DPS_SOLVER_AUTO 0       /* fused lane-group kernel when nz <= 512, else DPS_SOLVER_LANE */
#define DPS_SOLVER_LANE 1       /* one wavenumber per lane, separate z-transform kernels */

/* dtype of user arrays */
#define DPS_F32 0
#define DPS_F64 1

/*
 * Parameters; fields and defaults mirror uammd.StokesParameters
 * (uammd_python.cpp:74-127, uammd_interface.h:20-41).  Lengths are physical.
 *   beta, beta_d : ES exponent beta*w (uammd convention); <=0 selects 1.785*w.
 *   alpha, alpha_d : kernel support half-width; <=0 selects w*h/2 (the CPU
 *                    reference's value, src/ParticleList.cpp:229).
 *   w_d/beta_d/alpha_d describe the derivative (dipole) kernel used for torques.
 *   solver : DPS_SOLVER_* (no counterpart in the reference; both give the same
 *            result to rounding).
 *   reg_kernel : 0 = torques through the derivative kernels (default);
 *            1 = the CPU reference's useRegKernel option (FCM.py:149, 257,
 *            372-396, 427-439): torques are spread with the plain ES kernel
 *            of the dipole width (w_d, beta_d, alpha_d; support alpha_d), half
 *            their spectral curl joins the forcing, and angular velocities
 *            are the same kernel applied to half the spectral curl of the
 *            velocity (Fourier in x, y; Chebyshev coefficient derivative in
 *            z, Fourier in z when periodic).  Uses the one-wavenumber-per-lane
 *            spectral kernels whatever `solver` says.
 */
typedef struct dps_params {
  double viscosity;
  double Lx, Ly;
  double zmin, zmax;
  int w, w_d;
  double alpha, alpha_d;
  double beta, beta_d;
  int nx, ny, nz;
  int mode;
  int precision;
  int solver;
  int reg_kernel;
} dps_params;

typedef struct dps_handle dps_handle;
typedef struct dps_batch dps_batch;

/* Create a solver for `number_particles` particles (DPStokesGlue::initialize,
   uammd_wrapper.cu:119-133).  Allocates all device memory and FFT plans on the
   current HIP device.  *out is NULL on failure. */
int dps_create(const dps_params* params, int number_particles, dps_handle** out);

/* Release everything (DPStokesGlue::clear).  NULL is accepted. */
int dps_destroy(dps_handle* h);

/* Set particle positions from a device array of 3N values (DPStokesGlue::
   setPositions, uammd_wrapper.cu:136-143; FCM.SetPositions, FCM.py:270-332).
   Synchronises `stream` to report DPS_ERR_RANGE. */
int dps_set_positions(dps_handle* h, const void* positions, int dtype, void* stream);

/* Position update of a time step, kept on the device (ParticleList::update,
   ParticleList.cpp:550-753, which FCM.SetPositions calls from the second call
   on, FCM.py:326-332): positions += dt * velocities (velocities may be NULL:
   no move), x and y wrapped into the unit cell [0, Lx] x [0, Ly] by the
   reference's rule and written back in place -- the reference rewraps the
   caller's array too (ParticleList.cpp:556-558) -- then the particles are
   located as by dps_set_positions.  A caller's integrator can loop Mdot and
   this call with the positions never leaving the device.  Both arrays use
   `dtype`.  Unlike dps_set_positions it does not synchronise the stream, so
   a step loop of dps_mdot and this call can be captured in a HIP graph: a
   particle leaving the z extent (the reference exits) sets an error flag on
   the device, which stays set over further updates until dps_check_positions
   reports it; until then such a particle is located at the nearest end of
   the z extent.  Without valid positions (before dps_set_positions, or after
   a reported range error) the call synchronises and reports DPS_ERR_RANGE
   itself, as dps_set_positions does. */
int dps_update_positions(dps_handle* h, void* positions, const void* velocities, double dt, int dtype,
                         void* stream);

/* Synchronise `stream` and report whether a dps_update_positions since the
   last check (or since dps_set_positions) moved a particle out of the z
   extent: DPS_ERR_RANGE (and the positions count as unset, as after a
   failed dps_set_positions), else DPS_OK.  Clears the flag.  No counterpart
   in the reference, whose update exits at once (ParticleList.cpp:595-600). */
int dps_check_positions(dps_handle* h, void* stream);

/* Mobility product (DPStokesGlue::Mdot, uammd_wrapper.cu:146-170; FCM.Mdot,
   FCM.py:334-441): velocities = M_tt F + M_tr T, angular = M_rt F + M_rr T.
   torques and angular may be NULL (then no torque is applied and angular
   velocities are not computed), matching the reference's `torques.size < 3N`
   rule (uammd_python.cpp:51-53).  All four arrays use `dtype`. */
int dps_mdot(dps_handle* h, const void* forces, const void* torques, void* velocities,
             void* angular, int dtype, void* stream);

/* Adjoint of dps_mdot at the current positions (no counterpart in the
   reference; the differentiable form of the path).  With
   L = <grad_velocities, U> + <grad_angular, W> and (U, W) = Mdot(forces,
   torques):
     grad_forces, grad_torques = dL/dF, dL/dT = Mdot(grad_velocities,
       grad_angular), the solve taken as self-adjoint (exact when triply
       periodic; doubly periodic, up to the discretisation's asymmetry);
       grad_torques needs grad_forces and grad_angular (zeros when W does not
       enter L);
     grad_positions = dL/dx (3N):
       -sum_i gU[p,i] D_k u_i(x_p) - sum_i F[p,i] D_k v_i(x_p)
       + 1/2 sum eps_abc (gW[p,a] J_kb u_c(x_p) - T[p,c] J_kb v_a(x_p)),
       with u the velocity grid of (F, T), v that of (gU, gW), D_k
       interpolation with the k-th derivative of the force kernel and J_kb
       with the (k, b) second derivative of the torque kernel, as functions
       of the particle position (a clamped offset contributes no derivative
       along its axis); two spreads, two solves and one or two gradient
       interpolations.  With reg_kernel = 1 torque terms are DPS_ERR_ARG.
       With grad_positions, grad_torques (dL/dT) needs no grad_angular.
   Outputs may be NULL (not computed).  All arrays use `dtype`. */
int dps_mdot_adjoint(dps_handle* h, const void* forces, const void* torques, const void* grad_velocities,
                     const void* grad_angular, void* grad_forces, void* grad_torques, void* grad_positions,
                     int dtype, void* stream);

/* dps_mdot_adjoint without any device allocation.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc),
   for which the first position-gradient call of dps_mdot_adjoint is unusable
   because it allocates the adjoint's second velocity grid and derivative
   records.  Same arguments, results and errors as dps_mdot_adjoint; those
   buffers come instead from one static scratch array per GPU of
   DPS_NOALLOC_SCRATCH_BYTES bytes, set up by dps_create and held until the
   process ends.  Solvers whose adjoint needs more (second grid 3 nx ny nz
   reals plus records of a few hundred reals per particle: the 256^2 x 96
   slit workload in FP64 fits, the 1024^2 x 48 monolayer does not) get
   DPS_ERR_ARG.  Every call on one GPU shares that GPU's array: calls that
   overlap on one GPU (other streams, other threads) race, and must be
   ordered by the caller.  The array keeps the last call's contents; nothing
   in it needs to start at zero.  The handle's device must be current. */
#define DPS_NOALLOC_SCRATCH_BYTES 268435456
int dps_mdot_adjoint_noalloc(dps_handle* h, const void* forces, const void* torques,
                             const void* grad_velocities, const void* grad_angular, void* grad_forces,
                             void* grad_torques, void* grad_positions, int dtype, void* stream);

/* Stage entry points (for stage-level parity tests and the adjoint):
   spread forces (+torque curl) to the real grid, returned in the reference's
   interior layout (Ny, Nx, Nz, 3) FP64; solve grid -> velocity grid (same
   layout, FP64); interpolate a velocity grid (same layout) to particles.
   With reg_kernel = 1 the torque curl lives in spectral space only: torques
   to dps_stage_spread and angular to dps_stage_interp are DPS_ERR_ARG. */
int dps_stage_spread(dps_handle* h, const void* forces, const void* torques, int dtype,
                     double* grid_out, void* stream);
int dps_stage_solve(dps_handle* h, const double* grid_in, double* grid_out, void* stream);
int dps_stage_interp(dps_handle* h, const double* grid_in, void* velocities, void* angular,
                     int dtype, void* stream);

/* Batched form (no counterpart in the reference; many independent
   suspensions as one launch).  A batch groups existing handles (each with its
   own box, grid and particle count; they stay owned by the caller and must
   outlive the batch).  dps_batch_mdot runs every instance's dps_mdot: the
   first call captures them, forked over `nstreams` internal streams (default
   priority), into one HIP graph, and later calls with the same pointers and
   dtype replay it with one launch (other pointers re-capture).  Each array
   argument holds one pointer per instance; torques / angular (the arrays or
   their entries) may be NULL.  Position updates of the instances (same
   buffers) are seen by replays.  dps_batch_tune times the spreading z-chunk
   candidates of every instance on its current positions and inputs, keeps
   the fastest (chunk_out: one int per instance, may be NULL) and drops the
   captured graph. */
int dps_batch_create(dps_handle* const* handles, int n, int nstreams, dps_batch** out);
int dps_batch_destroy(dps_batch* b);
int dps_batch_mdot(dps_batch* b, const void* const* forces, const void* const* torques, void* const* velocities,
                   void* const* angular, int dtype, void* stream);
int dps_batch_tune(dps_batch* b, const void* const* forces, const void* const* torques, int dtype, void* stream,
                   int* chunk_out);
/* Apply spreading z chunks chosen earlier (one int per instance, e.g. what
   dps_batch_tune returned and the caller stored), without timing: each must
   fit its instance (with the torque pass when `torques` is nonzero and the
   instance has a dipole kernel), else DPS_ERR_ARG and nothing changes.
   Rebuilds the entry lists and drops the captured graph.  No counterpart in
   the reference: it lets a caller reuse tuned choices across runs (the
   Python layer's on-disk tuning cache). */
int dps_batch_set_chunks(dps_batch* b, const int* chunks, int torques, void* stream);

/* Query derived sizes: nx, ny, nz, number of particles, Nzeff of the monopole
   and dipole z-extended grids, z support widths.  `out` holds 8 ints. */
int dps_get_info(const dps_handle* h, int* out);

/* Per-kernel timings of the last dps_mdot in microseconds (HIP events on the
   call's stream, enabled by dps_set_profiling).  `out` holds 8 doubles:
   locate+sort, spread, fft forward, z forward, bvp, z inverse, fft inverse,
   interp; z forward and z inverse read 0 where one fused kernel runs them
   with the bvp.  Returns DPS_ERR_STATE if profiling is off. */
int dps_set_profiling(dps_handle* h, int enable);
int dps_get_timings(dps_handle* h, double* out);

/* "single" or "double" for the handle's grid arithmetic (uammd getPrecision,
   uammd_wrapper.cu:214-221).  Never NULL. */
const char* dps_get_precision(const dps_handle* h);

/* Last error message of this thread (static storage). */
const char* dps_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* DPSTOKES_HIP_H */
