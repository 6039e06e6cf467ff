/*
 * getm_amd: MI355X (gfx950) tracer-transport step for pygetm.
 *
 * C-ABI drop-in for the hot-path entry points that pygetm's Cython layer binds
 * (python/_pygetm.pyx:24-66, implemented bind(c) in python/src/wrapper.F90).
 * Each entry point keeps the reference's name and argument meaning; the
 * differences are stated per function and are the same throughout:
 *
 *   - every field pointer is a DEVICE pointer (HBM) in the reference layout:
 *     C-order (nz, ny_, nx_) = Fortran (nx_, ny_, nz), i fastest, halo
 *     included; interface fields (w, nuh) have nz+1 levels, level 0 = bottom;
 *     masks are int32, everything else float64;
 *   - a trailing `stream` (a hipStream_t, may be NULL for the null stream):
 *     work is enqueued on it and the call returns without synchronising;
 *   - an int status instead of `void`: GETM_OK or one of the GETM_ERR_*
 *     codes below (the reference has no status and silently ignores an
 *     invalid scheme; here the create call returns NULL instead).
 *
 * Fields are caller-owned device buffers that are registered with a grid by
 * their pygetm names (grid_set_array), where the reference's Fortran grid owns
 * them and hands out pointers (get_array, wrapper.F90:70-125).
 *
 * Objects keep no global state and may be used from several host threads as
 * long as one object is not used concurrently on two streams.
 */
#ifndef GETM_AMD_H
#define GETM_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    GETM_OK = 0,
    GETM_ERR_ARG = 1,          /* null handle/pointer, bad direction, shape mismatch */
    GETM_ERR_SCHEME = 2,       /* scheme code outside 1..6 */
    GETM_ERR_NOT_PREPARED = 3, /* vertical diffusion applied before prepare (diffusion.F90:154) */
    GETM_ERR_HIP = 4,          /* a HIP allocation or launch failed */
    GETM_ERR_CAPACITY = 5      /* a size exceeds what the entry point supports */
};

/* Advection schemes, pygetm.operators.AdvectionScheme (operators.py:11-18). */
enum {
    GETM_HSIMT = 1,
    GETM_MUSCL = 2,
    GETM_P2_PDM = 3,
    GETM_SPLMAX13 = 4,
    GETM_SUPERBEE = 5,
    GETM_UPSTREAM = 6
};

/* ---- grids ------------------------------------------------------------- */

/* Replaces create_grid (python/src/wrapper.F90:48-59).  Fortran index bounds;
 * padded extents are nx_ = imax-imin+1+2*halox, ny_ likewise, nz = kmax-kmin+1.
 * Returns NULL for empty ranges, negative halos, nx_ or ny_ >= 2^24, or a layer
 * of FP64 values (nx_*ny_*8 bytes) of 4 GiB or more. */
void* create_grid(int imin, int imax, int jmin, int jmax, int kmin, int kmax, int halox, int haloy);

/* Replaces grid_finalize (wrapper.F90:61-68). */
void grid_finalize(void* grid);

/* Counterpart of get_array (wrapper.F90:70-125) with ownership inverted: the
 * caller registers a device buffer under the reference's field name.  Names
 * used by the transport step: "mask" (int32 2D), "dx", "dy", "idx", "idy",
 * "area", "iarea" (2D), "hn", "ho" (3D). */
int grid_set_array(void* grid, const char* name, void* device_ptr);

/* ---- advection ------------------------------------------------------------ */

/* Replaces advection_create (wrapper.F90:188-202).  NULL for an invalid scheme
 * or a NULL grid. */
void* advection_create(int scheme, void* tgrid);

/* Replaces advection_finalize (wrapper.F90:204-211). */
void advection_finalize(void* adv);

/* Replaces advection_uv_calculate (wrapper.F90:213-254): one x (direction 1,
 * u2d, advection.F90.template:23-121) or y (direction 2, v2d, :123-221)
 * half-sweep over nk layers, in place on h and var.  uvgrid is the U grid for
 * direction 1 (reads mask, idx, dy) and the V grid for direction 2 (mask, dx,
 * idy).  Ah (2D on uvgrid) may be NULL: no horizontal diffusion.  The FP64
 * result equals the reference bit for bit. */
int advection_uv_calculate(int direction, int nk, void* adv, void* tgrid, void* uvgrid, const double* uv,
                           const double* Ah, double timestep, double* h, const double* huv, double* var,
                           void* stream);

/* Replaces advection_w_calculate (wrapper.F90:256-273): vertical sweep with
 * CFL sub-stepping (w3d, advection.F90.template:223-374), in place. */
int advection_w_calculate(void* adv, void* tgrid, const double* w, const double* w_var, double timestep,
                          double* h, double* var, void* stream);

/* No reference counterpart (it is a sequence of the calls above in
 * Advection.apply_3d, operators.py:95-133): the five stages
 * v dt/2, u dt/2, w dt, u dt/2, v dt/2 (ufirst = 0) or u-v-w-v-u (ufirst = 1)
 * on one tracer with h starting from its current contents, enqueued without
 * intermediate copies.  w_var may equal w. */
int advection_apply_3d(void* adv, void* tgrid, void* ugrid, void* vgrid, const double* u, const double* v,
                       const double* w, const double* w_var, const double* Ah_u, const double* Ah_v,
                       double timestep, double* h, double* var, int ufirst, void* stream);

/* Sub-stepping statistics of the z sweep since the last reset (no reference
 * counterpart; the reference caps at 200 iterations silently,
 * advection.F90.template:298-310).  Synchronises the stream. */
int advection_get_stats(void* adv, int* max_iters, int* capped_columns, int reset, void* stream);

/* ---- vertical diffusion ---------------------------------------------------- */

/* Replaces vertical_diffusion_create (wrapper.F90:127-138). */
void* vertical_diffusion_create(void* tgrid);

/* Replaces vertical_diffusion_finalize (wrapper.F90:140-147). */
void vertical_diffusion_finalize(void* vd);

/* Replaces c_vertical_diffusion_prepare (wrapper.F90:149-161,
 * diffusion.F90:135-204).  nx, ny are the padded extents nx_, ny_; nuh has
 * nz+1 levels of which 1..nz-1 are read. */
int c_vertical_diffusion_prepare(void* vd, int nx, int ny, int nz, double molecular, const double* nuh,
                                 double timestep, double cnpar, const int* mask, const double* ho,
                                 const double* hn, void* stream);

/* Replaces c_vertical_diffusion_apply (wrapper.F90:163-186,
 * diffusion.F90:206-314).  ea2/ea4 may be NULL. */
int c_vertical_diffusion_apply(void* vd, int nx, int ny, int nz, const int* mask, const double* ho,
                               const double* hn, double* var, const double* ea2, const double* ea4,
                               void* stream);

/* ---- batched tracer step ------------------------------------------------- */

/* Tracer storage/arithmetic precision of a transport object.
 * GETM_F32C (compensated FP32): tracers are stored as FP32 values plus int16
 * residuals that extend them by 15 bits; the tracer buffer of transport_step,
 * transport_inventory and transport_conserve is then ntracers FP32 planes
 * (ntracers x nz x ny_ x nx_) immediately followed by ntracers int16 planes of
 * the same shape, and every call on one buffer passes the same ntracers. */
enum { GETM_F64 = 1, GETM_F32 = 2, GETM_F32C = 3 };

/* Rounding of the FP32 tracer stores of the batched step. */
enum { GETM_ROUND_NEAREST = 0, GETM_ROUND_STOCHASTIC = 1 };

/* No single reference counterpart: replaces the loop of
 * TracerCollection.advance (python/pygetm/tracer.py:193-276), i.e.
 * Advection.apply_3d_batch (operators.py:135-193) followed by
 * VerticalDiffusion prepare/apply per tracer, for T tracers that share one
 * velocity field, one layer-thickness evolution (starting from the T grid's
 * "ho") and one diffusion matrix (dzo = dzn = the T grid's "hn").
 * ugrid/vgrid must have "hn" (face thicknesses) registered.
 * precision GETM_F64: tracers are double and results equal the per-tracer
 * calls bit for bit.  GETM_F32: tracers are float, stencil/limiter/tridiagonal
 * arithmetic is FP32, the thickness update is FP64.  GETM_F32C: tracers are
 * FP32 plus an int16 residual each (see above); limiter and tridiagonal
 * elimination run in FP32 on differences of the FP64-decoded values, fluxes,
 * cell updates and the thickness in FP64, and each launch rounds its results
 * once (to 2^-16 of an FP32 ulp), so neither rounding errors nor inventory
 * defects accumulate. */
void* transport_create(void* tgrid, void* ugrid, void* vgrid, int scheme, double cnpar, int precision, int ufirst);
void transport_finalize(void* tr);

/* One step for ntracers stacked tracers (ntracers x nz x ny_ x nx_, float or
 * double as created), in place.  w is used as w_var for every tracer; Ah_u and
 * Ah_v are both NULL (no horizontal diffusion) or both set.  h_out (FP64
 * nz x ny_ x nx_, may be NULL) receives the advected thickness, the
 * reference's Advection.h after apply_3d_batch. */
int transport_step(void* tr, int ntracers, void* tracers, const double* u, const double* v, const double* w,
                   const double* nuh, const double* Ah_u, const double* Ah_v, double timestep, double molecular,
                   double* h_out, void* stream);

/* As transport_step, with a vertical velocity and a molecular diffusivity per
 * tracer: tracer t is advected in the w stage by w_var[t] (the reference's
 * get_w / combined_w, python/pygetm/tracer.py:223-231 and operators.py:108-109;
 * the thickness still follows w) and diffused with molecular[t] (tracer.py:
 * 273-275).  w_var is a host array of ntracers device pointers (NULL entries,
 * or w_var NULL: w); molecular a host array of ntracers values.  The w stage
 * and the diffusion run once per run of consecutive tracers with equal values,
 * so the FP64 form stays bit-identical to the reference's per-tracer sequence. */
int transport_step_mixed(void* tr, int ntracers, void* tracers, const double* u, const double* v, const double* w,
                         const double* const* w_var, const double* nuh, const double* Ah_u, const double* Ah_v,
                         double timestep, const double* molecular, double* h_out, void* stream);

/* No reference counterpart: launch shape of the batched kernels -- the height
 * of the horizontal-stage tiles (4, 8 or 16 rows of 64 columns, and 12 for the
 * compensated form; 0 = the default: 16 rows, two cells per lane, for the FP32
 * forms with 16 or more tracers and more than 8 per block and for the FP64
 * and compensated forms without horizontal diffusion, the compensated form's
 * y-first pair on 12 rows with 16 or more tracers; 8 rows otherwise) and, for
 * the
 * three kernel kinds (horizontal stage pairs, w stage, FP32 diffusion), the
 * tracers per block (the tracer batches of one step are spread over the grid;
 * 0 = all tracers in every block; defaults 0, 16 (64 for GETM_F32C) and 16).
 * Results do not depend on these values. */
int transport_set_tuning(void* tr, int tile_y, int tracers_per_block_h, int tracers_per_block_w,
                         int tracers_per_block_d);

/* No reference counterpart: the FP32 column solve of the batched step.  Columns
 * with 2 <= nz <= pcr_max_nz (at most 32) are solved by wave-level parallel
 * cyclic reduction (one lane per level, for shallow columns), the others by per-lane Thomas
 * elimination (the reference's algorithm, diffusion.F90:276-313); 0 = Thomas
 * only (default).  The two agree to FP32 rounding, not bit for bit. */
int transport_set_solver(void* tr, int pcr_max_nz);

/* No reference counterpart (the reference computes in FP64): rounding of the
 * FP32 tracer updates.  GETM_ROUND_STOCHASTIC (the default) rounds every stored
 * c + dc up or down with probability given by the exact residual, so the
 * expected stored value is exact and rounding errors do not accumulate in a
 * steady flow (with rounding to nearest, an increment below half an ulp is lost
 * every step).  The random bits are a hash of (step count, stage, tracer index,
 * cell index): results are deterministic and independent of launch shapes and
 * batching.  The step count is read from step_counter (a device uint64 that each
 * transport_step advances by one; NULL: the object's own counter, zero at
 * creation), so a caller that moves a domain between objects keeps its stream
 * by passing the same counter.  No effect on GETM_F64. */
int transport_set_rounding(void* tr, int mode, unsigned long long* step_counter);

/* Tracer sources of the FP32 form's next steps, the counterpart of the ea4 term
 * that TracerCollection._diffuse passes to the diffusion (python/pygetm/tracer.py:
 * 263-272, diffusion.F90:269-270): a device array of ntracers x nz x ny_ x nx_
 * FP64 values, (source + surface flux in the top layer) * dt * source_scale per
 * tracer, read by every transport_step until replaced; NULL clears it.  FP32
 * transports only (GETM_ERR_ARG for GETM_F64, which runs tracers with sources
 * through the per-tracer operators); the column solve is then the per-lane
 * elimination whatever transport_set_solver chose. */
int transport_set_sources(void* tr, const double* ea4);

/* Tracer inventories (the reference's domain integrals of tracers,
 * simulation.py:1408-1466): inventory[t] = sum over wet cells of
 * area * h * c for each of ntracers stacked tracers, accumulated in FP64
 * with compensation in a fixed order (deterministic).  h = NULL uses the T
 * grid's "hn"; the T grid must have "area" registered.  inventory is a device
 * array of ntracers doubles. */
int transport_inventory(void* tr, int ntracers, const void* tracers, const double* h, double* inventory,
                        void* stream);

/* No reference counterpart (the FP64 reference conserves by construction):
 * for FP32 tracers, restores each tracer's inventory (with "hn") to target[t]
 * (device, ntracers doubles) by changing a few cells by one ulp, chosen by
 * rounding a mass-proportional prefix sum over a fixed cell order.  Leaves the
 * inventory within one cell's rounding of the target.  GETM_F32C: multiplies
 * every wet cell by target[t] / inventory in FP64 (exact to the storage
 * rounding).  No-op for FP64. */
int transport_conserve(void* tr, int ntracers, void* tracers, const double* target, void* stream);

/* No reference counterpart: transport_conserve for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results; the working
 * memory is a 32 MiB array inside the library, one per GPU, held until the
 * process ends, instead of buffers the transport object grows on demand.
 * Needs B = ceil(nx_ / 64) * ceil(ny_ / 4) blocks of partial sums per tracer:
 * GETM_F32C uses 4 B + 1 doubles per tracer, GETM_F32 8 B + 3; more than
 * 4 Mi doubles in total returns GETM_ERR_CAPACITY (e.g. 1028 x 772 cells with
 * halos: up to 319 GETM_F32C tracers, 159 GETM_F32 tracers).  Calls that
 * overlap on one GPU (other streams or threads) share the array and race;
 * calls ordered on one stream do not. */
int transport_conserve_noalloc(void* tr, int ntracers, void* tracers, const double* target, void* stream);

/* The tracer part of the reference's Simulation.check_finite
 * (python/pygetm/simulation.py:348-410): for each of ntracers stacked tracers
 * (the object's precision; GETM_F32C: the FP32 values), result[2 t] = the
 * number of non-finite values in interior cells whose T mask is 1 or 2 and
 * result[2 t + 1] = the index within the tracer, k * ny_ * nx_ + j * nx_ + i,
 * of the first of them (INT64_MAX if none).  result: device, 2 ntracers int64;
 * the call initialises it (by a kernel, so it may be graph-captured). */
int transport_check_finite(void* tr, int ntracers, const void* tracers, long long* result, void* stream);

/* z-sweep sub-stepping statistics of a transport object (as advection_get_stats). */
int transport_get_stats(void* tr, int* max_iters, int* capped_columns, int reset, void* stream);

/* ---- ensembles of independent domains -------------------------------------- */

/* No reference counterpart: many independent domains (each its own T/U/V
 * grids, nx, ny, nz <= 64, mask and ntracers[d] FP32 tracers) stepped together
 * with one launch per kernel kind, i.e. TracerCollection.advance of every
 * domain (tracer.py:193-276) in four launches.  Per-domain arithmetic does not
 * depend on the other domains: results equal transport_step on each domain
 * with GETM_F32, bit for bit.  Usage: create, bind every domain, commit (all
 * allocation and table uploads, synchronous), then ensemble_step any number of
 * times -- it only enqueues kernels, so it can be captured in a hipGraph. */
void* ensemble_create(int ndomains, void* const* tgrids, void* const* ugrids, void* const* vgrids,
                      const int* ntracers, int scheme, double cnpar, int ufirst);
void ensemble_finalize(void* ens);

/* Fields of domain d: tracers (ntracers[d] x nz x ny_ x nx_ float), u, v (on the
 * domain's U/V grids), w, nuh (nz+1 levels), time step and molecular
 * diffusivity.  Call ensemble_commit after (re)binding. */
int ensemble_bind(void* ens, int domain, float* tracers, const double* u, const double* v, const double* w,
                  const double* nuh, double timestep, double molecular);

/* Launch shape as transport_set_tuning (tile_y 0: the default, 16 rows, or 8
 * with 128-column tiles); takes effect at the next commit. */
int ensemble_set_tuning(void* ens, int tile_y, int tracers_per_block_h, int tracers_per_block_w,
                        int tracers_per_block_d);

/* Width of the horizontal-stage tiles of an ensemble: 32, 64 (the default) or
 * 128 columns; with the tile height it gives 32 x {8, 16}, 64 x {4, 8, 16} or
 * 128 x {4, 8} (ensemble_commit returns GETM_ERR_ARG for another pair).  Takes
 * effect at the next commit; results do not depend on it.  No counterpart in
 * the reference (a launch shape of this implementation, searched by the
 * autotuner). */
int ensemble_set_tile_x(void* ens, int tile_x);

/* Column solver as transport_set_solver, per domain by its nz; takes effect at
 * the next commit. */
int ensemble_set_solver(void* ens, int pcr_max_nz);

/* Rounding as transport_set_rounding; step_counters: ndomains device counters
 * (distinct), or NULL entries / a NULL array for the ensemble's own (zero at
 * creation).  Results equal transport_step on each domain started from the same
 * count.  Takes effect at the next commit. */
int ensemble_set_rounding(void* ens, int mode, unsigned long long* const* step_counters);

/* Allocates scratch, builds and uploads the work tables (synchronous). */
int ensemble_commit(void* ens, void* stream);

/* One step of every domain. */
int ensemble_step(void* ens, void* stream);

/* Device pointer of domain d's advected thickness after the last step. */
const double* ensemble_thickness(void* ens, int domain);

/* ---- instrumentation ---------------------------------------------------- */

/* Kernel identifiers for getm_profile_read. */
enum {
    GETM_K_ADV_UV_F64 = 0,
    GETM_K_ADV_W_F64 = 1,
    GETM_K_VDIFF_PREPARE_F64 = 2,
    GETM_K_VDIFF_APPLY_F64 = 3,
    GETM_K_HPAIR_VU = 4,  /* batched stages v dt/2 + u dt/2 (or u + v) */
    GETM_K_HPAIR_UV = 5,  /* batched stages u dt/2 + v dt/2 (or v + u) */
    GETM_K_WCOL = 6,      /* batched w stage */
    GETM_K_VDIFF_F32 = 7, /* batched FP32 vertical diffusion */
    GETM_K_INVENTORY = 8, /* tracer inventories */
    GETM_K_FIX = 9,       /* FP32 inventory fixer */
    GETM_NUM_KERNELS = 16
};

/* Input producer of the step: vertical velocity at layer interfaces from layer
 * transports pk (U grid) and qk (V grid) and the thickness change ho -> hn,
 * replacing c_w_momentum_3d (python/src/momentum.F90:191-216, bound at
 * python/_pygetm.pyx:54, called by Momentum.update_diagnostics,
 * momentum.py:879-883).  Same arguments: nx, ny the array extents, imin..jmax
 * the Fortran (1-based) point range (_pygetm.pyx:454 passes halo+1 ..
 * halo+n+1), w (nz+1, ny, nx) updated for interfaces 1 .. nz-1 of mask == 1
 * points, w_0 and w_nz untouched.  Bit-identical to the reference. */
int c_w_momentum_3d(int nx, int ny, int nz, int imin, int imax, int jmin, int jmax, const int* mask,
                    const double* dyu, const double* dxv, const double* iarea, const double* ho, const double* hn,
                    const double* pk, const double* qk, double timestep, double* w, void* stream);

/* Replaces c_vertical_advection_to_sources (python/src/wrapper.F90:447-486,
 * declared at python/_pygetm.pyx:65, called by python/pygetm/fabm.py:368):
 * first-order upstream transfer of a vertical velocity w (sinking < 0, floating
 * > 0) of concentration c into the sources s of the layers it moves material
 * between, in place.  Arrays are the Fortran (nx, ny, nz) -- C (nz, ny, nx), k = 0
 * at the surface -- with nx, ny the extents including the halos halox, haloy, as
 * the reference passes them.  If w is 0 at every interior point, s is not
 * touched.  Bit-identical to the reference.  The activity test uses a flag that
 * the library holds per device: calls on one device must be ordered (one stream,
 * or synchronised). */
int c_vertical_advection_to_sources(int nx, int ny, int nz, int halox, int haloy, const int* mask, const double* c,
                                    const double* w, const double* h, double* s, void* stream);

/* Replaces c_elevation2depth (python/src/wrapper.F90, declared at
 * python/_pygetm.pyx:49, called by update_depth, simulation.py:1575-1577):
 * D = max(H + z, Dmin) where mask != 0, over n points.  Bit-identical. */
int c_elevation2depth(int n, const double* z, const double* H, double Dmin, const int* mask, double* D, void* stream);

/* Replaces c_thickness2vertical_coordinates (python/src/wrapper.F90, declared at
 * python/_pygetm.pyx:47, called at simulation.py:1596-1599): layer-centre depths
 * zc (nz levels) and interface depths zf (nz + 1 levels) of the layer
 * thicknesses h above the bottom at -bottom_depth, where mask != 0; nx, ny the
 * extents including halos.  Bit-identical. */
int c_thickness2vertical_coordinates(int nx, int ny, int nz, const int* mask, const double* bottom_depth,
                                     const double* h, double* zc, double* zf, void* stream);

/* Replaces c_alpha (python/src/wrapper.F90:428-435, declared at
 * python/_pygetm.pyx:48, called at simulation.py:1569-1570): the shallow-water
 * damping alpha = max(0, min(1, (D - Dmin) / (Dcrit - Dmin))) where mask != 0,
 * over n points.  Bit-identical. */
int c_alpha(int n, const double* D, double Dmin, double Dcrit, const int* mask, double* alpha, void* stream);

/* Replaces c_update_gvc (python/src/wrapper.F90:567-601, declared at
 * python/_pygetm.pyx:66, called by vertical_coordinates.GVC.__call__): layer
 * thicknesses h (nz levels) of generalized vertical coordinates from the water
 * depth D where mask != 0; kk is the 1-based reference layer as the reference's
 * wrapper passes it (_pygetm.pyx:583-585) and dbeta is a device array of nz
 * zoomed sigma fractions.  Bit-identical. */
int c_update_gvc(int nx, int ny, int nz, double dsigma, const double* dbeta, double Dgamma, int kk, const double* D,
                 const int* mask, double* h, void* stream);

/* Layer thicknesses of sigma coordinates, h_k = dga[k] * D where mask != 0: the
 * numpy product of vertical_coordinates.Sigma.__call__
 * (python/pygetm/vertical_coordinates.py:106-113), which has no compiled
 * counterpart in the reference.  dga is a device array of nz fractions. */
int c_update_sigma(int nx, int ny, int nz, const double* dga, const double* D, const int* mask, double* h,
                   void* stream);

/* Replace grid_interp_x / grid_interp_y / grid_interp_z / grid_interp_xy
 * (python/src/interpolation.F90:9-72, declared at python/_pygetm.pyx:28-31, bound
 * at _pygetm.pyx:180-190 and chosen by Grid.interpolator, core.py:407-462):
 * target(i + off) = 0.5 (source(i) + source(i + 1)) along x, y or z, and the
 * four-point average for xy, over i-fastest arrays (nx, ny, nz); other target
 * points are not written.  Bit-identical.  source and target must be distinct
 * arrays; offsets that would write outside the target are rejected. */
int grid_interp_x(int nx, int ny, int nz, const double* source, double* target, int ioffset, void* stream);
int grid_interp_y(int nx, int ny, int nz, const double* source, double* target, int joffset, void* stream);
int grid_interp_z(int nx, int ny, int nz1, int nz2, const double* source, double* target, int koffset,
                  void* stream);
int grid_interp_xy(int nx1, int ny1, int nx2, int ny2, int nz, const double* source, double* target, int ioffset,
                   int joffset, void* stream);

/* Replaces c_reconstruct_transport_change (python/src/momentum.F90:323-340,
 * declared at python/_pygetm.pyx:59, called by momentum.py:941-942, 959-960,
 * 1088, 1097): un = (un * hn - Uo) / timestep for every i and Fortran rows
 * jmin..jmax of nz layers (imin/imax are accepted and, as in the reference, not
 * used).  Bit-identical. */
int c_reconstruct_transport_change(int nx, int ny, int nz, int imin, int imax, int jmin, int jmax, double* un,
                                   const double* hn, const double* Uo, double timestep, void* stream);

/* Velocities from layer transports, uk = pk / hn over n points: the numpy
 * division of Momentum.update_diagnostics (python/pygetm/momentum.py:885-892),
 * which has no compiled counterpart in the reference.  IEEE division: the bits
 * of numpy's result, zero thicknesses included. */
int c_velocities_from_transports(long long n, const double* pk, const double* hn, double* uk, void* stream);

/* No reference counterpart (the reference accumulates cpu_time around its
 * loops, diffusion.F90:161-311).  When enabled, every kernel launch of this
 * library is bracketed by HIP events recorded on the launch stream;
 * getm_profile_read returns the summed elapsed milliseconds and the number of
 * launches of one kernel (synchronising on the recorded events) and optionally
 * resets them.  Process-wide instrumentation, off by default. */
int getm_profile_enable(int enable);
int getm_profile_read(int kernel, double* total_ms, long* launches, int reset);

#ifdef __cplusplus
}
#endif
#endif
