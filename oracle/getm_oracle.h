/*
 * FP64 CPU restatement of pygetm's 3D tracer-transport step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP product path is
 * compared against; the product never links, loads or calls it.
 *
 * Layout: every 3D field is a C-order (nz, ny_, nx_) block (Fortran
 * (nx_, ny_, nz), i fastest), 2D fields are (ny_, nx_), with nx_ = nx + 2*halox
 * and ny_ = ny + 2*haloy.  Interface fields (w, nuh) have nz+1 levels,
 * level 0 = bottom.  Masks are int32.
 *
 * All functions return 0 on success, nonzero on bad arguments.
 */
#ifndef GETM_ORACLE_H
#define GETM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* One directional half-sweep over nk layers.
 * direction 1 (x, u2d): metric1 = idxu, metric2 = dyu   (advection.F90.template:23-121)
 * direction 2 (y, v2d): metric1 = dxv,  metric2 = idyv  (advection.F90.template:123-221)
 * ah may be NULL (no horizontal diffusion). */
int oracle_advection_uv(int direction, int scheme, int nx, int ny, int nk, int halox, int haloy,
                        const int* uvmask, const double* metric1, const double* metric2,
                        const double* huv, const double* uv, const int* tmask, const double* iarea,
                        const double* ah, double dt, double* h, double* var);

/* Vertical sweep with CFL sub-stepping (advection.F90.template:223-374).
 * Returns the largest iteration count used in *max_iters if non-NULL. */
int oracle_advection_w(int scheme, int nx, int ny, int nz, int halox, int haloy,
                       const double* w, const double* w_var, const int* tmask, double dt,
                       double* h, double* var, int* max_iters);

/* theta-method matrix set-up (diffusion.F90:135-204): fills auxo, auxn (nz, ny_, nx_);
 * level k holds interface k+1 (between layers k and k+1).  nuh has nz+1 levels. */
int oracle_vdiff_prepare(int nx, int ny, int nz, int halox, int haloy, double molecular,
                         const double* nuh, double dt, double cnpar, const double* dzo,
                         const double* dzn, double* auxo, double* auxn);

/* RHS + Thomas solve per mask==1 column (diffusion.F90:206-314).  ea2/ea4 may be NULL. */
int oracle_vdiff_apply(int nx, int ny, int nz, int halox, int haloy, const int* mask,
                       const double* dzo, const double* dzn, const double* auxo,
                       const double* auxn, double* var, const double* ea2, const double* ea4);

/* One tracer-transport step for ntracers tracers (stacked (T, nz, ny_, nx_)),
 * in the order of Advection.apply_3d_batch (operators.py:135-193, ufirst=False:
 * v dt/2, u dt/2, w dt, u dt/2, v dt/2, every tracer restarting each stage
 * from the shared thickness) followed by TracerCollection._diffuse
 * (tracer.py:248-276: prepare once with dzo = dzn = hn, apply per tracer,
 * no sources).  ah_u/ah_v may be NULL.  h_out (nz, ny_, nx_) receives the
 * advected thickness (the operator's work array) if non-NULL. */
int oracle_transport_step(int scheme, int nx, int ny, int nz, int halox, int haloy,
                          const int* tmask, const int* umask, const int* vmask,
                          const double* idxu, const double* dyu, const double* dxv,
                          const double* idyv, const double* iarea,
                          const double* ho, const double* hn, const double* hu, const double* hv,
                          const double* u, const double* v, const double* w, const double* nuh,
                          const double* ah_u, const double* ah_v, double dt, double cnpar,
                          double molecular, int ntracers, double* tracers, double* h_out);

/* Vertical velocity at layer interfaces from layer transports and the layer
 * thickness change, python/src/momentum.F90:191-216 (c_w_momentum_3d): for the
 * interfaces k = 1 .. nz-1 of mask == 1 points in [imin, imax] x [jmin, jmax]
 * (0-based, inclusive),
 *   w_k = w_{k-1} - (hn_k - ho_k) / dt
 *         - (pk_k dyu - pk_k(i-1) dyu(i-1) + qk_k dxv - qk_k(j-1) dxv(j-1)) iarea,
 * with hn_k, ho_k, pk_k, qk_k the layer below interface k; w_0 and w_nz are not
 * written.  nx, ny are the array extents including halos. */
int oracle_w_momentum_3d(int nx, int ny, int nz, int imin, int imax, int jmin, int jmax, const int* mask,
                         const double* dyu, const double* dxv, const double* iarea, const double* ho,
                         const double* hn, const double* pk, const double* qk, double dt, double* w);

/* First-order upstream transfer of a vertical velocity (sinking/floating) into
 * source terms, python/src/wrapper.F90:447-486 (c_vertical_advection_to_sources).
 * Arrays are (nz, ny, nx) in C order, i.e. the Fortran (nx, ny, nz), with k = 0 at
 * the surface (the reference's surface-to-bottom ordering, upward = -1).  If w is
 * 0 at every interior point (halos excluded) of every layer, s is left untouched.
 * Else, for every interior (i, j) and k = 0 .. nz-2 with mask == 1 at k and k+1,
 * in increasing k:
 *   lw = (-1 * 0.5) (w_k + w_{k+1});  flux = lw (lw > 0 ? c_k : c_{k+1});
 *   s_k -= flux / h_k;  s_{k+1} += flux / h_{k+1}. */
int oracle_vertical_advection_to_sources(int nx, int ny, int nz, int halox, int haloy, const int* mask,
                                         const double* c, const double* w, const double* h, double* s);

/* Water depth from elevation, python/src/wrapper.F90 (c_elevation2depth):
 * D = max(H + z, Dmin) where mask != 0 (flang's MAX: a > b ? a : b), n points. */
int oracle_elevation2depth(int n, const double* z, const double* H, double Dmin, const int* mask, double* D);

/* Layer-centre and interface depths from layer thicknesses, python/src/wrapper.F90
 * (c_thickness2vertical_coordinates), where mask != 0: zc_0 = -H + 0.5 h_0,
 * zc_k = zc_{k-1} + 0.5 (h_{k-1} + h_k); zf_0 = -H, zf_k = zf_{k-1} + h_{k-1}
 * (zf has nz + 1 levels).  Arrays (nz, ny, nx) C order, k = 0 at the bottom. */
int oracle_thickness2vertical_coordinates(int nx, int ny, int nz, const int* mask, const double* H, const double* h,
                                          double* zc, double* zf);

/* Grid-to-grid averages, python/src/interpolation.F90 (grid_interp_x/_y/_z/_xy):
 * target(i + off) = 0.5 (source(i) + source(i + 1)) along x, y or z, and the
 * four-point average 0.25 (s(i,j) + s(i+1,j) + s(i,j+1) + s(i+1,j+1)) for xy, in
 * that order.  Arrays (nz, ny, nx) C order; other target points are untouched. */
int oracle_interp_x(int nx, int ny, int nz, const double* source, double* target, int ioffset);
int oracle_interp_y(int nx, int ny, int nz, const double* source, double* target, int joffset);
int oracle_interp_z(int nx, int ny, int nz1, int nz2, const double* source, double* target, int koffset);
int oracle_interp_xy(int nx1, int ny1, int nx2, int ny2, int nz, const double* source, double* target, int ioffset,
                     int joffset);

/* Change of transport from advected velocity, python/src/momentum.F90:323-340:
 * un = (un hn - Uo) idt with idt = 1/timestep, for all i and Fortran rows
 * jmin..jmax of each of the nz layers (imin/imax are not used there either). */
int oracle_reconstruct_transport_change(int nx, int ny, int nz, int jmin, int jmax, double* un, const double* hn,
                                        const double* Uo, double timestep);

/* Shallow-water damping, python/src/wrapper.F90:428-435 (c_alpha): where
 * mask != 0, alpha = max(0, min(1, (D - Dmin) / (Dcrit - Dmin))), flang's
 * MAX/MIN being the compare-selects a > b ? a : b and a < b ? a : b. */
int oracle_alpha(int n, const double* D, double Dmin, double Dcrit, const int* mask, double* alpha);

/* Generalized vertical coordinates, python/src/wrapper.F90:567-601 (c_update_gvc):
 * per point with mask != 0, alpha = (min(Dgamma / D, 1) dsigma - dbeta[kk-1]) /
 * (dsigma - dbeta[kk-1]) and h_k = D (dbeta[k] + alpha (dsigma - dbeta[k])); kk is
 * the 1-based reference layer.  D, mask (ny, nx); h (nz, ny, nx). */
int oracle_update_gvc(int nx, int ny, int nz, double dsigma, const double* dbeta, double Dgamma, int kk,
                      const double* D, const int* mask, double* h);

/* Sigma coordinates, python/pygetm/vertical_coordinates.py:106-113: h_k = dga[k] D
 * where mask != 0 (numpy's multiply(dga, D, where=mask)). */
int oracle_update_sigma(int nx, int ny, int nz, const double* dga, const double* D, const int* mask, double* h);

/* Limiter value for a given scheme, cfl and slope ratio (limiters.h:1-44);
 * exported for the closed-form tests. */
double oracle_limiter(int scheme, double cfl, double ratio);

#ifdef __cplusplus
}
#endif
#endif
