/* firtez_gpu.h -- C ABI of the MI355X implementation of the FIRTEZ-dz per-column
 * polarized radiative-transfer path (forward synthesis + analytic response
 * functions).  Plain C types only; a Fortran host binds it through
 * iso_c_binding (firtez_dz_amd/fortran/firtez_gpu_mod.f90).
 *
 * All entry points return 0 on success or one of the FZ_E* codes; the text of
 * the last error of a context is available from fz_last_error().  The
 * reference STOPs on the same conditions (file:line given per code).
 */
#ifndef FIRTEZ_GPU_H
#define FIRTEZ_GPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FZ_OK = 0,
  FZ_EEOS = 1,      /* electron-density iteration failed / N_H<0  (chemical_equilibrium.f90:259-268) */
  FZ_ENAN = 2,      /* non-finite value  (voigt.f:91-95, absorption_matrix.f90:1882-1889) */
  FZ_EION = 3,      /* ionisation stage outside 1..3  (line_opacity.f90:86-90) */
  FZ_ELAMBDA = 4,   /* wavelength below 1800 A  (cont_opacity.f90:83-86) */
  FZ_ECONFIG = 5,   /* invalid line/region/height configuration */
  FZ_EHIP = 6,      /* HIP runtime error */
  FZ_ENOMEM = 7,    /* device allocation failed */
  FZ_EARG = 8       /* invalid argument */
};

/* Response-function parameter sets */
enum {
  FZ_RF_NONE = 0,
  FZ_RF_SPHERICAL = 1, /* d/d(T|Pg, |B|, inclination, azimuth, v_los) at every height */
  FZ_RF_CARTESIAN = 2,  /* d/d(T|Pg, Pg|T, rho|T, Bx, By, Bz, v_z): DSYN1D slots 1-7 (rtesolver.f90:209-258) */
  FZ_RF_CARTESIAN_P0 = 3 /* slots 1-7 and P0 (slot 8): local response 1, propagated, then times
                            PG(K)/PG(NZ) and the propagated Pg|T response (forward.f90:543-545) */
};

/* One row of the line table.  Replaces the FORWARD_PARAM per-line arrays
 * LINE_ZN, LINE_ION, LINE_L0, LOGGF, EPLOW, SIGMA, ALPHA, SL..JU
 * (forward_param.f90:16-20) filled by NREAD_LINES_DATABASE. */
typedef struct {
  int32_t zn, ion;
  double lambda0, loggf, eplow, sigma, alpha;
  double sl, ll, jl, su, lu, ju;
} fz_line;

/* One spectral region = one LINES: control-file row (input_file.f90:253-367):
 * main line + blends (indices into the line table), WAVE grid in mA. */
typedef struct {
  int32_t nblend;
  int32_t line[4];
  int32_t nwave;
  double wave_start, wave_step;
} fz_region;

typedef struct fz_ctx fz_ctx;

/* Context on one device.  No counterpart in the reference (module globals). */
int fz_create(fz_ctx** ctx, int device);
void fz_destroy(fz_ctx* ctx);
const char* fz_last_error(const fz_ctx* ctx);

/* Lines, regions and height grid z_km[nz] (bottom first).  Replaces
 * NREAD_LINES / WAVE_INIT (input_file.f90:253-367, pre_post_duties.f90:592-621)
 * and the ZZ grid (grid_param.f90). */
int fz_configure(fz_ctx* ctx, const fz_line* lines, int nlines, const fz_region* regions,
                 int nregions, int nz, const double* z_km);
int fz_numw(const fz_ctx* ctx);

/* GAIN1D(4,NUMW) continuum normalisation from the HSRA model, computed on the
 * device; replaces GET_NORM (normalization.f90:696-754).  gain: host, 4*numw
 * floats, layout gain[4*l + s]. */
int fz_gain_hsra(fz_ctx* ctx, float* gain);
/* Use a given GAIN1D instead (e.g. the reference's). */
int fz_set_gain(fz_ctx* ctx, const float* gain);

/* Spectral line-spread function (MLSF).  lsf_kernel: LSF_KERNEL(NUMW) and
 * lsf_valid: LSF_VALID(NUML) (as int32) exactly as LSF_INIT leaves them
 * (pre_post_duties.f90:770-898): per region the unit-sum taps divided by the
 * region's pixel count, tap 0 (the centre) first.  Every later forward then
 * convolves the Stokes rows and the response functions of each valid region
 * circularly with the taps, after GAIN1D, like the FFT products of
 * forward.f90:621-657 and 814-844.  lsf_kernel NULL switches it off;
 * fz_configure resets it. */
int fz_set_lsf(fz_ctx* ctx, const double* lsf_kernel, const int32_t* lsf_valid);

/* NLTE departure coefficients (POPL, POPU of FORWARD_PARAM,
 * forward_param.f90:47, as READ_DEPARTURE_COEFFICIENTS fills them,
 * input_file.f90:2146-2166): popl POPL(NUML,NZ), popu POPU(NUML,NZ)
 * column-major double (element (r,k) at r + k*nreg) for the configured
 * regions and heights.  Every later forward uses them in the line opacity of
 * every line of region r (OPAC_LINE(POPL(r,k), POPU(r,k), ...),
 * forward.f90:949, line_opacity.f90:101-106) and in the source function of
 * the region's wavelengths (rtesolver.f90:1030-1064).  NULL: LTE (the
 * reference's default of 1, input_file.f90:504-508); fz_configure resets to
 * LTE.  Values must be finite and positive. */
int fz_set_departure(fz_ctx* ctx, const double* popl, const double* popu);

/* Spatial PSF of the coupled inversion, as READ_PSF reads it into COU_PSF
 * (pre_post_duties.f90:488-555) for the coupled-inversion buffers of
 * INIT_COUPLED_INVERSION_STATIC_VARS (coupled_inversion.f90:10-84):
 * psf PSF(NPY,NPX) column-major (element (iy,ix) at iy + npy*ix), centre at
 * (NPY/2+1, NPX/2+1), used as given (no renormalisation, like the
 * reference's FFT kernel).  Odd sides of at most 63 pixels (the reference
 * takes NPX/2 pixels either side of the centre, which an even side does not
 * hold).  NULL clears it
 * (END_COUPLED_INVERSION_*, :136-207).  The PSF is a property of the
 * context and survives fz_configure. */
int fz_set_psf(fz_ctx* ctx, const double* psf, int npy, int npx);

/* The PSF's periodic 2-D convolution of nmaps maps on the nx*ny grid
 * (the FFT product of SMOOTH_MODEL, invert3d.f90:412-488, formed directly):
 *   out(y,x) = sum PSF(Y0+dy, X0+dx) in(y-dy mod ny, x-dx mod nx),
 * or with adjoint != 0 its transpose, the correlation with +dy, +dx.
 * Maps are channel-major float, map m at [m*nx*ny + y + ny*x]; the PSF's
 * half-widths NPX/2, NPY/2 must leave 2r+1 <= nx, ny.  d_in and d_out must
 * differ.  Asynchronous on the stream. */
int fz_psf_convolve_dev(fz_ctx* ctx, int nx, int ny, int nmaps, const float* d_in, float* d_out, int adjoint,
                        void* stream);

/* Per-pixel spectral windows: each of the ncol pixels observes its own
 * sub-range of every configured region.  The reference reads one set of
 * regions for the whole map (NREAD_LINES, input_file.f90:253-367: NUML
 * regions of NUMWAVE points, NUMW = SUM(NUMWAVE) at :359); a map whose
 * pixels carry different lines or window widths is run here in one launch instead of one run per
 * line set.  windows[(col*nreg + r)*2 + {0,1}] = first wavelength index in
 * region r and count (count 0: the region is not observed in that pixel).
 * With windows set, fz_forward_dev (rf_mode NONE), fz_forward_nodes_dev,
 * fz_lm_step_dev and fz_invert3d_dev take ncol equal to the ncol given here and
 * read/write ragged spectra: pixel col's wavelengths are its windows in region
 * order, numw_col = offsets[col+1]-offsets[col] of fz_window_offsets; Stokes
 * rows [4][numw_col] start at 4*offsets[col]; observations and best synthesis
 * rows [nstk][numw_col] at nstk*offsets[col]; node response functions
 * [np][nstk][numw_col] at np*nstk*offsets[col].  Each pixel's chi^2 is
 * normalised by its own nstk*numw_col.  Not combined with an LSF, response
 * functions per height or the host-array entry points.  windows NULL switches
 * windows off; fz_configure resets them. */
int fz_set_windows(fz_ctx* ctx, int ncol, const int32_t* windows);

/* offsets[0..ncol] (int64) of the ragged spectra set by fz_set_windows. */
int fz_window_offsets(const fz_ctx* ctx, int64_t* offsets);

/* FORWARD3D on device-resident data (forward.f90:43-306, per column the
 * FORWARD1D of forward.f90:857-1029; HYDROSTATIC as set by fz_set_hydrostatic).
 *   d_atm    [12][nz][ncol] float, in/out: T, Pg, rho, Pe, mu, Bx, By, Bz, vx,
 *            vy, vz, log tau5 (rho, Pe, mu, log tau5 -- and Pg with
 *            HYDROSTATIC -- are updated like the reference updates the cube)
 *   d_stokes [ncol][4][numw] float: SYN1D*GAIN1D
 *   d_jac    [ncol][np][nz][4][numw] float (np = 5 spherical, 7 cartesian, 8 with P0), the
 *            propagated response functions times GAIN1D; NULL if rf_mode=NONE
 *   d_status [ncol] int32 or NULL: per-column FZ_E* code
 *   stream   hipStream_t (NULL = default stream); the call is asynchronous. */
int fz_forward_dev(fz_ctx* ctx, int ncol, float* d_atm, int rf_mode, float* d_stokes, float* d_jac,
                   int32_t* d_status, void* stream);

/* fz_forward_dev with MISC SETUP FULL_STOKES (forward.f90:980-983, written
 * to fs_<FILEPROFILE> by DEALLOCATE_AND_WRITE, pre_post_duties.f90:265-271):
 * d_full [ncol][nz][4][numw] float receives the Stokes vector after every
 * height, bottom (k = 0) to top, times GAIN1D; its top height is d_stokes
 * before any LSF.  Not with per-pixel windows. */
int fz_forward_full_dev(fz_ctx* ctx, int ncol, float* d_atm, int rf_mode, float* d_stokes, float* d_jac,
                        float* d_full, int32_t* d_status, void* stream);

/* LINE OPTICAL DEPTH (MTAULIN, forward.f90:966-968, 1008-1015; written to
 * tl_<FILEMODEL> by WRITE_LINTAU, pre_post_duties.f90:168-180): d_tau
 * [ncol][nz][numw] float receives log10 of the optical depth at every
 * wavelength, integrated like log tau5 from the top (tau = 1e-9) with
 * KLINTAU = ETAI / N_H / m_u / MW per height.  Heights whose 5000 A
 * continuum opacity reaches the boundary threshold contribute 0 (the
 * reference reuses the last ETAI it formed there).  Runs the EOS on d_atm
 * (updated like a forward).  Not with per-pixel windows. */
int fz_line_tau_dev(fz_ctx* ctx, int ncol, float* d_atm, float* d_tau, int32_t* d_status, void* stream);

/* FORWARD3D on host arrays in the reference's layout (forward.f90:43-306):
 * cubes(NZ,NY,NX) column-major float (TEM3D..TAU3D5; PG3D, RHO3D, PEL3D,
 * MW3D and TAU3D5 are written back as SPLIT_MODEL3D does, forward.f90:1225-1270),
 * am_i_done(NY,NX) (columns with nonzero are skipped), syn3d(NFREQ,NY,NX)
 * double with NFREQ = 4*NUMW in Stokes-major blocks, dsyn(NFREQ,IFREEP,NY,NX)
 * float with IFREEP = np*NZ (parameter-major, height within) or NULL. */
int fz_forward3d(fz_ctx* ctx, int nx, int ny, float* tem3d, float* pg3d, float* rho3d, float* pel3d,
                 float* mw3d, const float* bx3d, const float* by3d, const float* bz3d,
                 const float* vx3d, const float* vy3d, const float* vz3d, float* tau3d5,
                 const int32_t* am_i_done, int rf_mode, double* syn3d, float* dsyn);

/* Hydrostatic equilibrium from the top (RK4_INTEGRATION_LOG,
 * hydrostatic_tools.f90:150-222) for ncol columns on the device: reads T and
 * the top Pg (and Pe) of d_atm, writes Pg, rho, Pe, mu. */
int fz_hydrostatic_dev(fz_ctx* ctx, int ncol, float* d_atm, int32_t* d_status, void* stream);

/* Device memory helpers for hosts without their own allocator. */
int fz_malloc(fz_ctx* ctx, void** ptr, uint64_t bytes);
int fz_free(fz_ctx* ctx, void* ptr);
int fz_memcpy_h2d(fz_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int fz_memcpy_d2h(fz_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int fz_synchronize(fz_ctx* ctx);

/* Kernel timing: when enabled, the library records HIP events on the launch
 * stream around the microphysics launch and around the radiative-transfer
 * launches of each fz_forward_dev call; fz_last_timing waits for them and
 * returns the elapsed milliseconds of the last call. */
int fz_set_profiling(fz_ctx* ctx, int on);

/* Overlap (default off): with more than one radiative-transfer tile,
 * fz_forward_dev runs the microphysics of the next tile on a second stream
 * of the library while the current tile's radiative transfer runs on the
 * caller's stream; the call stays asynchronous and ordered on the caller's
 * stream.  With profiling on, the microphysics time is then the first tile's
 * only.  On the 288x288x96 forward with response functions the concurrent
 * kernels slow the radiative transfer more than they hide (112.5 ms per
 * step against 108).  No counterpart in the reference (its columns are
 * independent MPI work items). */
int fz_set_overlap(fz_ctx* ctx, int on);
int fz_last_timing(fz_ctx* ctx, float* micro_ms, float* rt_ms, int* rt_launches);

/* HYDROSTATIC: yes / no (input_file.f90:516-520).  With it every forward
 * re-integrates Pg from the top pressure of each column (RK4_INTEGRATION_LOG,
 * forward.f90:891-906), takes N_H and n_e from that integration, and adds the
 * hydrostatic term to dI/dT (rtesolver.f90:218-231, HYDROSTATICDER).  The P0
 * response (forward.f90:543-545) is not produced. */
int fz_set_hydrostatic(fz_ctx* ctx, int on);

/* Microturbulence xi (cm/s, 0 <= xi <= 1e7) in every line's Doppler width,
 * sqrt(2kT/m + xi^2) in the line opacity, the damping parameter and the
 * Voigt argument (and their T derivatives).  No counterpart in the
 * reference, whose width is thermal only (absorption_matrix.f90:1794-1795,
 * damping.f90:842-1012, line_opacity.f90:34-217); 0, the default, is that
 * width exactly.  A property of the context: it survives fz_configure. */
int fz_set_microturbulence(fz_ctx* ctx, double xi_cm_s);

/* Columns processed per radiative-transfer launch (scratch is sized for this
 * tile); the autotuner sets it.  Default 32768. */
int fz_set_tile(fz_ctx* ctx, int columns_per_launch, int block_threads);

/* Autotuner (no counterpart in the reference): times fz_forward_dev on the
 * caller's data with microphysics tiles of 4096 .. 262144 columns and
 * radiative-transfer launches of 2048 .. 32768 columns (within a third of the
 * free device memory), keeps the fastest of each for this context and
 * returns them.  The outputs of the last trial are a valid forward. */
int fz_autotune(fz_ctx* ctx, int ncol, float* d_atm, int rf_mode, float* d_stokes, float* d_jac,
                int32_t* d_status, void* stream, int* tile, int* mtile);


/* ------------------------------------------------------------------ inversion */

/* Per-cycle inversion set-up: which parameters and Stokes parameters are
 * inverted, the nodes of each parameter in this cycle, regularisation and
 * solver.  Mirrors INV_ATMPAR, INV_STK, NSLAB_PER_FREEV, VREGULARIZATION,
 * PEN_TYP, PEN_ALP, SOLVER_METHOD, NODE_EXPANSION, RSVDTOL, INU, MPERT and
 * MAXSTEPS (invert_param.f90, code_modes.f90, input_file.f90:421-453,
 * get_dmodel.f90:415-421).  P0 (atmpar[7]) is not supported. */
typedef struct {
  int32_t atmpar[8];     /* T, Pg, rho, Bx, By, Bz, v_z, P0 */
  int32_t invstk[4];     /* I, Q, U, V */
  int32_t nodes[8];      /* nodes of each inverted parameter (1..nz) */
  int32_t pen_typ[8];    /* 0: amplitude, 1: first differences */
  double pen_alp[8];
  int32_t regul;         /* VREGULARIZATION */
  int32_t solver;        /* 1: truncated SVD pseudo-inverse, 2: CG iteration */
  int32_t node_expansion;/* 1: SIR/NICOLE re-projection, 2: direct */
  int32_t maxsteps;      /* pixels with more steps without improvement are done */
  double svdtol;         /* RSVDTOL (0 -> 1e-4) */
  double inu;            /* INU (0 -> NFREQ) */
  double mpert;          /* MPERT (0 -> 0.15) */
} fz_inversion;

/* Node basis COEFS(nz, nj) of GET_INT_COEFS (get_dmodel.f90:563-644):
 * hat functions on nj equidistant height indices; coefs[j*nz + k]. Host only. */
int fz_int_coefs(int nz, int nj, double* coefs);

/* Configure the inversion of this context (after fz_configure); returns the
 * number of free parameters IFREEP through *ifreep and NFREQ through *nfreq.
 * Replaces SET_INVERSION_CONSTANTS / INITIALIZE_CYCLE's per-cycle set-up
 * (pre_post_duties.f90:625-700, invert3d.f90:228-325).  IFREEP up to 2048:
 * up to 64 the LM step keeps its work arrays in LDS, beyond in device memory
 * that the library allocates per launch batch. */
int fz_set_inversion(fz_ctx* ctx, const fz_inversion* inv, int* ifreep, int* nfreq);

/* FORWARD3D in inversion mode on device data: Stokes as fz_forward_dev and
 * the node-contracted response functions EQVDSYN (forward.f90:552-681) in
 * d_eqv[ncol][IFREEP][NFREQ] float (times GAIN1D). */
int fz_forward_nodes_dev(fz_ctx* ctx, int ncol, float* d_atm, float* d_stokes, float* d_eqv,
                         int32_t* d_status, void* stream);

/* One Levenberg-Marquardt step for every pixel with d_done[pix] == 0
 * (GET_DMODEL3DS / GET_PERT_1D, get_dmodel.f90:1327-1566, 1201-1244):
 *   d_obs      [ncol][NFREQ] double   observed profiles (OBS3D)
 *   d_stokes   [ncol][4][NUMW] float  current synthesis (fz_forward_nodes_dev)
 *   d_best_syn [ncol][NFREQ] double   in/out (BEST_SYN)
 *   d_sla      [ncol][9] double       in/out (INV_MAS(1:9): steps, chi2,
 *                                     previous chi2, PRECYC, ..., lambda)
 *   d_eqv      [ncol][IFREEP][NFREQ]  current EQVDSYN; d_eqv_best in/out
 *   d_isigmap  [NFREQ] double         weights WSTK^2/ISIGMA^2
 *   d_atm      [12][nz][ncol]         in: current model; out: next trial
 *   d_atm_best [12][nz][ncol]         in/out: best model
 *   d_done     [ncol] int32           in/out: AM_I_DONE (set when steps > maxsteps)
 *   d_delta    [ncol][IFREEP] double  applied node perturbation, or NULL */
int fz_lm_step_dev(fz_ctx* ctx, int ncol, const double* d_obs, const float* d_stokes, double* d_best_syn,
                   double* d_sla, const float* d_eqv, float* d_eqv_best, const double* d_isigmap,
                   float* d_atm, float* d_atm_best, int32_t* d_done, double* d_delta, void* stream);

/* GET_DMODEL3DS() (get_dmodel.f90:1327-1577) on host arrays in the
 * reference's module layout: one LM step (GET_PERT_1D, :1201-1244) for every
 * pixel with am_i_done == 0, after fz_set_inversion.  All arrays
 * column-major as the reference declares them, NFREQ as fz_set_inversion
 * returns it (the inverted Stokes blocks), IFREEP the free parameters:
 *   obs3d      OBS3D(NFREQ,NY,NX)      double in
 *   syn3d      SYN3D(NFREQ,NY,NX)      double in (current synthesis)
 *   best_syn   BEST_SYN(NFREQ,NY,NX)   double in/out
 *   inv_mas    INV_MAS(ld_inv_mas,NY,NX) double in/out, rows 1..9 (steps,
 *              chi2, previous chi2, PRECYC, ..., lambda)
 *   dsyn       DSYN(NFREQ,IFREEP,NY,NX) float in (current EQVDSYN)
 *   best_dsyn  BEST_DSYN(NFREQ,IFREEP,NY,NX) float in/out
 *   isigmap3d  ISIGMAP3D(NFREQ,NY,NX)  double in
 *   model[12]  TEM3D, PG3D, RHO3D, PEL3D, MW3D, BX3D, BY3D, BZ3D, VX3D,
 *              VY3D, VZ3D, TAU3D5 (NZ,NY,NX) float in/out: current -> next trial
 *   best_model[12] BEST_TEM3D ... BEST_TAU3D5 float in/out
 *   am_i_done  AM_I_DONE(NY,NX) in/out: set where INV_MAS(1) > MAXSTEPS
 * Synchronous.  Not combined with per-pixel windows (fz_lm_step_dev). */
int fz_get_dmodel3ds(fz_ctx* ctx, int nx, int ny, const double* obs3d, const double* syn3d, double* best_syn,
                     double* inv_mas, int ld_inv_mas, const float* dsyn, float* best_dsyn, const double* isigmap3d,
                     float* const* model, float* const* best_model, int32_t* am_i_done);

/* GET_PERTURBATION(NP,NF,IJACOBIAN,DIFF,LAMBDAP,DELTA,REG_PEN,REG_HSS,DTYPE,SVDTOL)
 * (lm.f90:137-204) on host arrays in lm.f90's layout, REAL*8 column-major:
 * ijacobian IJACOBIAN(NF,NP) (ijacobian[p*nf + f]), diff DIFF(NF), reg_pen
 * REG_PEN(NP) or NULL (zero), reg_hss REG_HSS(NP,NP) or NULL (zero), delta
 * DELTA(NP) out.  TERM1 = J^T J + REG_HSS with its diagonal times
 * (1 + lambdap), TERM2 = J^T DIFF - REG_PEN; dtype 1: DELTA = TERM1^+ TERM2
 * through the truncated pseudo-inverse of GET_INVERSE (lm.f90:61-133:
 * singular values below svdtol times the largest dropped; svdtol <= 0 means
 * the reference's default 1e-3), dtype 2: the CG iteration of
 * NEW_CALCULATE_DELTA_UBICGSTAB (lm.f90:375-453).  Any other dtype returns
 * FZ_EARG (the reference stops, lm.f90:194-201).  NP up to 2048.  J^T J and
 * J^T DIFF are formed on the FP64 matrix cores.  Synchronous. */
int fz_get_perturbation(fz_ctx* ctx, int np, int nf, const double* ijacobian, const double* diff, double lambdap,
                        double* delta, const double* reg_pen, const double* reg_hss, int dtype, double svdtol);

/* fz_get_perturbation on device arrays of the same layout, asynchronous on
 * stream (NULL = default stream).  The normal matrix and the solver's work
 * arrays live in device memory that the context allocates on first use and
 * keeps. */
int fz_get_perturbation_dev(fz_ctx* ctx, int np, int nf, const double* d_ijacobian, const double* d_diff,
                            double lambdap, double* d_delta, const double* d_reg_pen, const double* d_reg_hss,
                            int dtype, double svdtol, void* stream);

/* fz_get_perturbation_dev without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, hosts that forbid
 * allocation).  The normal matrix and the solver's work arrays live in a
 * fixed 6.3 MB device array of the library (one per GPU, held until the
 * process ends, its address taken in fz_create); the caller passes inputs
 * and outputs as for fz_get_perturbation_dev.  At most 512 parameters
 * (FZ_EARG beyond).  Calls that overlap on one GPU share that array and
 * race; each call writes every entry it reads, so no state carries over. */
int fz_get_perturbation_dev_noalloc(fz_ctx* ctx, int np, int nf, const double* d_ijacobian, const double* d_diff,
                                    double lambdap, double* d_delta, const double* d_reg_pen,
                                    const double* d_reg_hss, int dtype, double svdtol, void* stream);

/* HESSIAN(NP,NF,JACOBIAN,HSS) (lm.f90:30-39): HSS = JACOBIAN JACOBIAN^T on
 * host arrays, jacobian JACOBIAN(NP,NF) column-major (jacobian[f*np + p]),
 * hss HSS(NP,NP) out (symmetric).  FP64 matrix cores.  Synchronous. */
int fz_hessian(fz_ctx* ctx, int np, int nf, const double* jacobian, double* hss);

/* Control of a whole inversion (the INVERSION SETUP, STOKES SETUP and MISC
 * SETUP sections of the control file, input_file.f90:1064-1135, 1280-1303,
 * 170-190). */
typedef struct {
  int32_t atmpar[8];     /* inverted parameters (T, Pg, rho, Bx, By, Bz, v_z, P0) */
  int32_t nslb_max[8];   /* maximum nodes of each parameter (NSLB_MAX, "TEM 4") */
  int32_t invstk[4];     /* inverted Stokes parameters */
  float wstk[4];         /* Stokes weights before normalisation (STKI 1 ...) */
  double noise;          /* ISIGMA (NOISE), 0 -> 4e-3 */
  int32_t cycpow;        /* POW, 0 -> 2 */
  int32_t maxsteps;      /* MAXITER keyword (MAXSTEPS), 0 -> 5 */
  int32_t solver;        /* LINEAR_SOLVER, 0 -> 1 */
  int32_t node_expansion;/* PERTURBATION_TYPE, 0 -> 1 */
  int32_t regul;
  int32_t pen_typ[8];
  double pen_alp[8];
  double svdtol;         /* 0 -> 1e-4 */
  int32_t max_iterations;/* cap on LM iterations per cycle, 0 -> 1000 */
  int32_t smoothing;     /* SMOOTHING_B at every cycle start (the reference always applies it) */
  int32_t first_cycle;   /* first cycle to run, 0 -> 1 (a start model already on the final node
                            basis can skip the coarse cycles); no counterpart in the reference */
  int32_t assist;        /* SET_TAU assist initialisers of the first cycle (MISC SETUP TEMP, PGAS,
                            BVEC, VLOS: ASSIST_T/P/B/V, input_file.f90:170-176): FZ_ASSIST_* bits */
} fz_invert_control;

enum { FZ_ASSIST_B = 1, FZ_ASSIST_V = 2, FZ_ASSIST_T = 4, FZ_ASSIST_P = 8 };

/* STOKES SETUP: AUTO (input_file.f90:1294, SET_WEIGHTS get_dmodel.f90:2024-2047).
 * With it on, fz_invert3d_dev / fz_invert3d weight every pixel by its own
 * observations (fz_auto_weights_dev) in place of the control's wstk; the
 * coupled inversion and per-pixel windows refuse it (FZ_ECONFIG).  Default off. */
int fz_set_auto_weights(fz_ctx* ctx, int on);

/* SET_WEIGHTS3D (get_dmodel.f90:2051-2170) on device data: for every pixel of
 * d_obs [ncol][NFREQ] (the inverted Stokes blocks, I Q U V order) the weight
 * of block s is 1 / STDDEV of that block (n - 1 variance), capped at 10 x
 * the pixel's smallest, the four normalised to a root mean square of 1 over
 * the inverted parameters (SET_WEIGHTS1D :2191-2224) and rounded to REAL(4);
 * d_isigmap [ncol][NFREQ] receives WSTK^2 / ISIGMA^2 (SET_ISIGMAP :356-373),
 * ISIGMA the REAL(4) noise (0 -> 4e-3).  A block of equal values gives an
 * infinite weight before the cap, as in the reference.  Needs NUMW >= 2 and
 * the dense spectral grid.  Asynchronous on stream. */
int fz_auto_weights_dev(fz_ctx* ctx, int ncol, const double* d_obs, const int32_t* invstk, float noise,
                        double* d_isigmap, void* stream);

/* SET_TAU's assist initialisers (invert3d.f90:528-700) on device data, for
 * hosts that keep their own cycle loop: from the observations d_obs
 * [ncol][NFREQ] (the inverted Stokes blocks, as fz_invert3d_dev takes them;
 * the first region's wavelengths are read) they overwrite rows of the start
 * model d_atm [12][nz][ncol]: FZ_ASSIST_B Bx, By, Bz (needs I, Q, U, V
 * inverted), FZ_ASSIST_V v_z, FZ_ASSIST_T T (shifted to the brightness
 * temperature at height ceiling(nz/3)), FZ_ASSIST_P Pg (scaled to 1.333e5 there);
 * T and P also reset log tau5 as the reference does before its synthesis.
 * After fz_set_inversion (for the inverted Stokes parameters). */
int fz_assist_dev(fz_ctx* ctx, int ncol, const double* d_obs, int flags, float* d_atm, void* stream);

/* INVERT3D (invert3d.f90:57-178) on device data: the cycle schedule (nodes
 * per cycle from GET_NJEVALS, get_dmodel.f90:375-407), SET_TAU, the 1-2-1
 * smoothing of the inverted parameters, LM iterations (forward with
 * node-contracted response functions + fz_lm_step_dev) until the mean step
 * count reaches max(1, MAXSTEPS - (MAXITER - cycle)) or every pixel is done,
 * and END_CYCLE's best -> current + forward synthesis.
 *   d_atm    [12][nz][nx*ny] in: initial model, out: inverted model
 *   d_obs    [nx*ny][NFREQ] double observed profiles (inverted Stokes blocks)
 *   d_stokes [nx*ny][4][NUMW] float out: synthesis of the inverted model
 *   d_chi2   [nx*ny] double out (or NULL): best chi^2 of the last cycle
 *   iterations (or NULL): LM iterations done, summed over cycles
 * Synchronises the stream once per iteration (the stop test). */
int fz_invert3d_dev(fz_ctx* ctx, int nx, int ny, float* d_atm, const double* d_obs, const fz_invert_control* ctl,
                    float* d_stokes, double* d_chi2, int* iterations, void* stream);

/* INVERT3D of a map observed through the spatial PSF of fz_set_psf: the
 * coupled inversion whose state INIT_COUPLED_INVERSION_STATIC_VARS /
 * _DYNAMIC_VARS allocate (coupled_inversion.f90:10-134) and whose
 * mathematics the reference leaves unwired (GET_DMODEL3DC is a stub,
 * get_dmodel.f90:1707-1711).  Arguments as fz_invert3d_dev; d_obs holds the
 * observed (PSF-degraded) profiles.  Each LM iteration convolves the
 * synthesis with the PSF, accepts or rejects on the chi^2 of the whole map
 * (lambda / sqrt(10) or x 10 as UPDATE_LAMBDA, get_dmodel.f90:917-956),
 * and steps every pixel with the exact gradient of the map's chi^2
 * (J_q^T W (PSF^T r)_q, the adjoint convolution of the residual) and the
 * pixel's own block of the coupled Gauss-Newton matrix ((sum PSF^2)
 * J_q^T W J_q).  d_stokes receives the PSF-convolved synthesis of the
 * inverted model, d_chi2 each pixel's chi^2 against d_obs.  A cycle ends
 * when the map's step count (UPDATE_LAMBDA's STEPS, one for the map)
 * reaches MAX(1, MAXSTEPS - (MAXITER - cycle)) (MAKE_DECISION,
 * invert3d.f90:820-859) or at max_iterations.  Dense spectral grid only
 * (no fz_set_windows). */
int fz_coupled_invert3d_dev(fz_ctx* ctx, int nx, int ny, float* d_atm, const double* d_obs,
                            const fz_invert_control* ctl, float* d_stokes, double* d_chi2, int* iterations,
                            void* stream);

/* INVERT3D on host arrays in the reference's layout (invert3d.f90:57-178
 * with its module globals): the twelve cubes (NZ,NY,NX) float are the
 * initial model on entry and the inverted model on return; OBS3D(NFREQ,NY,NX)
 * double holds the observed profiles of the inverted Stokes parameters
 * (Stokes-major blocks of NUMW); SYN3D(4*NUMW,NY,NX) double receives the
 * synthesis of the inverted model, CHI2(NY,NX) (or NULL) its chi^2. */
int fz_invert3d(fz_ctx* ctx, int nx, int ny, float* tem3d, float* pg3d, float* rho3d, float* pel3d, float* mw3d,
                float* bx3d, float* by3d, float* bz3d, const float* vx3d, const float* vy3d, float* vz3d,
                float* tau3d5, const double* obs3d, const fz_invert_control* ctl, double* syn3d, double* chi2,
                int* iterations);

/* The coupled inversion (fz_coupled_invert3d_dev) on host arrays, arguments
 * as fz_invert3d: OBS3D holds the PSF-degraded observations, SYN3D receives
 * the PSF-convolved synthesis of the inverted model.  The PSF is the one of
 * fz_set_psf. */
int fz_coupled_invert3d(fz_ctx* ctx, int nx, int ny, float* tem3d, float* pg3d, float* rho3d, float* pel3d,
                        float* mw3d, float* bx3d, float* by3d, float* bz3d, const float* vx3d, const float* vy3d,
                        float* vz3d, float* tau3d5, const double* obs3d, const fz_invert_control* ctl, double* syn3d,
                        double* chi2, int* iterations);

/* ---- Learned initial model.  No counterpart in the reference, which starts
 * every inversion from the input model file (invert3d.f90:57-178); these
 * entry points propose that initial model from the observed profiles and
 * leave the inversion unchanged.
 *
 * Network: nin -> nhidden (nlayers times, ReLU) -> 5*nodes outputs, the node
 * values (hat basis of GET_INT_COEFS, get_dmodel.f90:563-644, on the
 * configured heights) of T, Bz, B_perp cos 2phi, B_perp sin 2phi and v_z.
 * Weights and activations BF16, FP32 accumulation (MFMA), FP32 master
 * weights and Adam moments.  One network per context. */
int fz_mlp_create(fz_ctx* ctx, int nin, int nhidden, int nlayers, int nodes, uint64_t seed);
/* Length of the parameter vector: weights and biases (W_l [out][in] row-major,
 * then b_l, layer by layer) followed by xmean[nin], xstd[nin], tmean[nout],
 * tstd[nout]; -1 without a network. */
int64_t fz_mlp_size(const fz_ctx* ctx);
int fz_mlp_get_params(fz_ctx* ctx, float* params);
int fz_mlp_set_params(fz_ctx* ctx, const float* params);
/* Input and target standardisation from a training set on the device:
 * d_x [n][nin], d_t [n][5*nodes] float. */
int fz_mlp_fit_norms_dev(fz_ctx* ctx, int n, const float* d_x, const float* d_t, void* stream);
/* Training targets of a model cube d_atm [12][nz][ncol] (least-squares node
 * values): d_t [ncol][5*nodes]. */
int fz_mlp_targets_dev(fz_ctx* ctx, int ncol, const float* d_atm, float* d_t, void* stream);
/* Standardised outputs d_y [n][5*nodes] for inputs d_x (float) or d_xd
 * (double) [n][nin]; the other pointer NULL. */
int fz_mlp_forward_dev(fz_ctx* ctx, int n, const float* d_x, const double* d_xd, float* d_y, void* stream);
/* Mean squared error on standardised targets (d_loss[0]) and its gradient
 * (d_grad, fz_mlp_size floats of which the normalisation entries are left
 * unwritten) over the rows d_idx[0..n) of d_x/d_t (d_idx NULL: rows 0..n). */
int fz_mlp_gradient_dev(fz_ctx* ctx, int n, const float* d_x, const float* d_t, const int32_t* d_idx, float* d_loss,
                        float* d_grad, void* stream);
/* One Adam step (beta 0.9/0.999, eps 1e-8, learning rate lr) on the
 * gradient of the rows d_idx[0..n). */
int fz_mlp_train_step_dev(fz_ctx* ctx, int n, const float* d_x, const float* d_t, const int32_t* d_idx, float lr,
                          float* d_loss, void* stream);
/* Initial model from observations d_obs [ncol][nin] (double, as fz_invert3d_dev
 * takes them): overwrites T (kept in 2500-18000 K), Bx, By, Bz and v_z of
 * d_atm [12][nz][ncol]; the other rows stay as given.  Synchronises. */
int fz_learned_init_dev(fz_ctx* ctx, int ncol, const double* d_obs, float* d_atm, void* stream);

#ifdef __cplusplus
}
#endif
#endif
