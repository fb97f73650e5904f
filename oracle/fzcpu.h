/* FP64 CPU restatement of the FIRTEZ-dz per-column forward + response-function
 * path.  TEST INFRASTRUCTURE ONLY (the oracle): tests, the smoke check and the
 * CPU-baseline leg of the benchmark call it as the checker; the product path
 * never links it. */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FZC_MAXLINES 16
#define FZC_MAXREG 8
#define FZC_MAXBLEND 4
#define FZC_MAXCOMP 64

/* line table row: forward_param.f90:16-20 (LINE_ZN, LINE_ION, LINE_L0, LOGGF,
 * EPLOW, ALPHA, SIGMA, SL..JU) */
typedef struct {
  int32_t zn, ion;
  double lambda0, loggf, eplow, sigma, alpha;
  double sl, ll, jl, su, lu, ju;
} fzc_line;

/* spectral region: one LINES: row (input_file.f90:253-367) */
typedef struct {
  int32_t nblend;            /* 1 + number of blends */
  int32_t line[FZC_MAXBLEND];/* 0-based line-table indices; [0] = main line */
  int32_t nwave;
  double wave_start, wave_step;   /* mA relative to the main line */
} fzc_region;

/* Configure lines/regions; resp != 0 turns the response functions on;
 * atmpar[8] mirrors INV_ATMPAR (T,Pg,rho,Bx,By,Bz,vz,P0). Returns numw. */
int fzc_setup(const fzc_line* lines, int nlines, const fzc_region* regs, int nreg,
              int resp, const int32_t* atmpar);
int fzc_numw(void);
void fzc_wave(double* wave);

/* atomic data and partition functions (atom_database.f90) */
void fzc_atomdata(double* xi, double* xii, double* ab, double* m, double* deno);
void fzc_partition(int elem, double T, double* u, double* du);

/* equation of state at one height; out = NH, NE, RHO, MU, PEL; d = 25 DERIVVAR */
int fzc_eos(float T, float P, double* ne_inout, double* out, double* d);
int fzc_opac_cont(float T, double lam, double ne, double nh, double* kc, double* d);
int fzc_opac_line(int line, float T, double ne, double nh, double* kl, double* d);
int fzc_damping(int line, float T, double nh, double* damp, double* d);
void fzc_voigt(int n, double a, const double* v, double* h, double* f);
int fzc_zeeman(int line, double* shift, double* strength, int32_t* type);

/* evolution operator of one layer (rtesolver.f90:269-1004): x7 = eta_I..rho_V,
 * dx[8][7] directions; evol[16] row-major, dev[8][16] */
int fzc_evol(const double* x7, double dz_km, const double* dx, double* evol, double* dev);

/* GAIN1D from the HSRA continuum (normalization.f90:696-754) */
int fzc_gain(float* gain);
void fzc_set_gain(const float* gain);
/* HYDROSTATIC=yes (and HYDROSTATICDER): the forward re-integrates Pg from
 * the top (RK4_INTEGRATION_LOG) and adds the hydrostatic term to dI/dT. */
void fzc_set_hydrostatic(int on);
/* microturbulence xi (cm/s) of every line's Doppler width; 0 (the default): thermal only */
void fzc_set_microturbulence(double xi);
/* FULL_STOKES: forward1d also writes the Stokes vector after every height,
 * buf[nz][4][numw] (times GAIN1D); NULL switches it off */
void fzc_set_full_stokes(double* buf);
/* LINE OPTICAL DEPTH: forward1d also writes log10 tau at every height and
 * wavelength, buf[nz][numw]; NULL switches it off */
void fzc_set_line_tau(double* buf);
int fzc_set_lsf(const double* lsf_kernel, const int32_t* valid);

/* One column (forward.f90:857-1029, HYDROSTATIC=no):
 * model[13*nz] column-major MODEL(NZ,13), updated in place;
 * syn[4*numw]  = SYN1D*GAIN1D;
 * dsyn[4*8*numw*nz] = propagated DSYN1D*GAIN1D (only if resp);
 * kc5[nz]. Returns 0 or an error code. */
/* NLTE departure coefficients POPL(NUML,NZ), POPU(NUML,NZ) column-major
 * (element (r,k) at r + k*nreg); NULL: LTE.  fzc_setup resets them. */
int fzc_set_departure(int nreg, int nz, const double* popl, const double* popu);
int fzc_forward1d(int nz, float* model, double* syn, double* dsyn, double* kc5);

/* ncol independent columns, OpenMP over columns (nthreads<=0: default). */
int fzc_forward_batch(int ncol, int nz, float* models, double* syn, double* dsyn,
                      int nthreads);

/* hydrostatic re-integration from the top (hydrostatic_tools.f90:150-222) */
int fzc_he(int nz, float* tem, float* pg, float* rho, float* pel, float* mw,
           double* nh, const double* zz);


/* ------------------------------------------------ Levenberg-Marquardt step (fzlm.c) */

/* Per-pixel inversion configuration: INV_ATMPAR, INV_STK, nodes per
 * parameter for the current cycle (NSLAB_PER_FREEV), regularisation
 * (VREGULARIZATION, PEN_TYP, PEN_ALP), SOLVER_METHOD, NODE_EXPANSION,
 * RSVDTOL, INU and MPERT (invert_param.f90, code_modes.f90). */
typedef struct {
  int32_t nz, numw;
  int32_t atmpar[8];
  int32_t invstk[4];
  int32_t nodes[8];
  int32_t pen_typ[8];
  double pen_alp[8];
  int32_t regul, solver, node_expansion, curic;
  double svdtol, inu, mpert;
} fzc_lmcfg;

int fzc_int_coefs(int nz, int nj, double* coefs);
int fzc_njevals(int apar, int curic, int maxiter, int cycpow, int toffset, int nz);
int fzc_eqv_dsyn(const fzc_lmcfg* c, const double* coefs, const double* dsyn, float* eqv);
int fzc_get_perturbation(int np, int nf, const double* jac, const double* diff, double lambda,
                         double* delta, const double* reg_pen, const double* reg_hss, int dtype,
                         double svdtol);
int fzc_lm_step(const fzc_lmcfg* c, const double* coefs, const double* obs, const double* syn,
                double* best_syn, double* sla, const float* jac_cur, float* jac_best,
                const double* isigmap, float* model, double* delta_out);

#ifdef __cplusplus
}
#endif
