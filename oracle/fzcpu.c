/* FP64 CPU restatement of the FIRTEZ-dz column forward + response functions.
 *
 * TEST INFRASTRUCTURE (the oracle).  Only tests/, the smoke check and the
 * CPU-baseline leg of bench.py use it, always as the checker.  Each routine
 * cites the reference routine it restates; arithmetic follows the reference's
 * order and its REAL(4) roundings so the results agree with the compiled
 * reference to rounding.  Thread safety: per-column state lives in struct
 * fzc_work; the configuration set by fzc_setup is read-only afterwards.
 */
#include "fzcpu.h"

#include <float.h>
#include <omp.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../firtez_dz_amd/csrc/fz_atomdata.h"

/* ------------------------------------------------ constants: cons_param.f90:14-31 */
#define DPI 3.141592653589793238462643
#define LIGHT 2.99792458e10
#define KBOL 1.38056503e-16
#define MELE 9.10938215e-28
#define QELE 4.80320425e-10
#define EVOLT 1.60217656e-12
#define MAMU 1.66053892e-24
#define HPLA 6.62606957e-27
#define GRAV 2.7414e4
static double CSAHA1, CSAHA2, RBOHR;
static double DENO_SA, DENO_SAM;
static double ABFAC[FZ_NELEM]; /* 10**(ABUND-12) */

enum { E_OK = 0, E_EOS = 1, E_NAN = 2, E_ION = 3, E_LAMBDA = 4, E_CONFIG = 5 };
int fzc_set_departure(int nreg, int nz, const double* popl, const double* popu);

/* ------------------------------------------------ configuration (read-only) */
static fzc_line LN[FZC_MAXLINES];
static int NLN;
static fzc_region RG[FZC_MAXREG];
static int NRG, NUMW, RESP;
static int HYDRO;   /* HYDROSTATIC (and HYDROSTATICDER) */
/* microturbulence xi^2 (cm^2 s^-2): the Doppler width sqrt(2 k T / m + xi^2)
 * of every line, i.e. the thermal width at T + m xi^2 / (2 k) (0: the
 * reference's thermal width, absorption_matrix.f90:1794-1795) */
static double XI2 = 0.0;
/* FULL_STOKES output [nz][4][NUMW] of the next forward1d calls, or NULL */
static double* FSOUT = NULL;
/* LINE OPTICAL DEPTH output [nz][NUMW] (log10 tau) of the next forward1d
 * calls, or NULL */
static double* TLOUT = NULL;
static double xi_t(int z) { return XI2 * fz_matom[z] * MAMU / (2.0 * KBOL); }
static int ATMPAR[8];
static int PIX0[FZC_MAXREG], PIX1[FZC_MAXREG];
static double *WAVE;
static float *GAIN;
/* spectral LSF (MLSF): LSF_KERNEL(NUMW) and LSF_VALID(NUML) as LSF_INIT leaves
 * them (pre_post_duties.f90:770-898) */
static int LSF_ON;
static double *LSFK;
static int LSFV[FZC_MAXREG];
static double BLDIFF[FZC_MAXREG][FZC_MAXBLEND];
/* Zeeman pattern per line (magsplit.f90:38-187), precomputed once */
static int ZN_N[FZC_MAXLINES];
static double ZN_SHIFT[FZC_MAXLINES][FZC_MAXCOMP], ZN_STR[FZC_MAXLINES][FZC_MAXCOMP];
static int ZN_TYPE[FZC_MAXLINES][FZC_MAXCOMP];

static int is_bad(double x) { /* checknan.f90:3-36 */
  return !(x == x) || x > DBL_MAX || x < -DBL_MAX || x == x + 1.0;
}

static void init_constants(void) {
  static int done = 0;
  if (done) return;
  CSAHA1 = pow(HPLA, 3.0) / pow(2.0 * DPI * MELE * KBOL, 3.0 / 2.0);
  CSAHA2 = EVOLT / KBOL;
  RBOHR = pow(HPLA, 2.0) / (4.0 * pow(DPI, 2.0) * MELE * pow(QELE, 2.0));
  DENO_SA = 0.0;
  DENO_SAM = 0.0;
  for (int i = 0; i < FZ_NELEM; ++i) { /* atom_database.f90:112-118 */
    ABFAC[i] = pow(10.0, fz_abund[i] - 12.0);
    DENO_SA += ABFAC[i];
    DENO_SAM += MAMU * fz_matom[i] * ABFAC[i];
  }
  done = 1;
}

/* ------------------------------------------------ partition functions
 * Interpreter for the predicated stack programs of fz_atomdata.h, which encode
 * PARTITION_FUNCTION (atom_database.f90:121-1036). */
void fzc_partition(int elem, double T, double* u, double* du) {
  double v[10] = {0};
  double st[32];
  int sp = 0;
  v[FZ_PV_T] = T;
  v[FZ_PV_X] = log(5040.0 / T);
  v[FZ_PV_Y] = 1.0e-3 * T;
  v[FZ_PV_DX] = -1.0 / T;
  for (int pc = fz_pf_start[elem - 1]; pc < fz_pf_start[elem]; ++pc) {
    int op = fz_pf_op[pc], a = fz_pf_arg[pc];
    double x, y;
    switch (op) {
      case FZ_PF_END: break;
      case FZ_PF_CONST: st[sp++] = fz_pf_const[a]; break;
      case FZ_PF_VAR: st[sp++] = v[a]; break;
      case FZ_PF_STORE: v[a] = st[--sp]; break;
      case FZ_PF_NEG: st[sp - 1] = -st[sp - 1]; break;
      case FZ_PF_EXP: st[sp - 1] = exp(st[sp - 1]); break;
      case FZ_PF_SEL: { /* new, old, cond */
        double c = st[--sp];
        double old = st[--sp];
        double nw = st[--sp];
        st[sp++] = (c != 0.0) ? nw : old;
        break;
      }
      default:
        y = st[--sp];
        x = st[--sp];
        switch (op) {
          case FZ_PF_ADD: x = x + y; break;
          case FZ_PF_SUB: x = x - y; break;
          case FZ_PF_MUL: x = x * y; break;
          case FZ_PF_DIV: x = x / y; break;
          case FZ_PF_MAX: x = x > y ? x : y; break;
          case FZ_PF_AMAX1: {
            float fx = (float)x, fy = (float)y;
            x = (double)(fx > fy ? fx : fy);
            break;
          }
          case FZ_PF_GT: x = (x > y) ? 1.0 : 0.0; break;
          case FZ_PF_LT: x = (x < y) ? 1.0 : 0.0; break;
          case FZ_PF_LE: x = (x <= y) ? 1.0 : 0.0; break;
          case FZ_PF_GE: x = (x >= y) ? 1.0 : 0.0; break;
          case FZ_PF_NE: x = (x != y) ? 1.0 : 0.0; break;
          case FZ_PF_EQ: x = (x == y) ? 1.0 : 0.0; break;
          case FZ_PF_AND: x = (x != 0.0 && y != 0.0) ? 1.0 : 0.0; break;
        }
        st[sp++] = x;
    }
  }
  u[0] = v[FZ_PV_UI];
  u[1] = v[FZ_PV_UII];
  u[2] = v[FZ_PV_UIII];
  du[0] = v[FZ_PV_DUI];
  du[1] = v[FZ_PV_DUII];
  du[2] = v[FZ_PV_DUIII];
}

void fzc_atomdata(double* xi, double* xii, double* ab, double* m, double* deno) {
  init_constants();
  for (int i = 0; i < FZ_NELEM; ++i) {
    xi[i] = fz_xi[i];
    xii[i] = fz_xii[i];
    ab[i] = fz_abund[i];
    m[i] = fz_matom[i];
  }
  deno[0] = DENO_SA;
  deno[1] = DENO_SAM;
}

/* ------------------------------------------------ Saha factors
 * ALPHAI/ALPHAII and their T-derivatives, as every ION_CALC_* routine forms
 * them (chemical_equilibrium.f90:1601-1652). */
typedef struct {
  double u[FZ_NELEM][3], du[FZ_NELEM][3];
  double a1[FZ_NELEM], a2[FZ_NELEM], da1[FZ_NELEM], da2[FZ_NELEM];
  double T;
} saha_t;

static void saha_init(saha_t* s, double T) {
  s->T = T;
  double lam = CSAHA1 * pow(T, -3.0 / 2.0);
  double dlam = -3.0 / (2.0 * T);
  for (int i = 0; i < FZ_NELEM; ++i) {
    fzc_partition(i + 1, T, s->u[i], s->du[i]);
    double e1 = CSAHA2 * fz_xi[i] / T, e2 = CSAHA2 * fz_xii[i] / T;
    if (e1 > 0.0) e1 = fmin(e1, 650.0);
    if (e1 < 0.0) e1 = fmax(e1, -650.0);
    if (e2 > 0.0) e2 = fmin(e2, 650.0);
    if (e2 < 0.0) e2 = fmax(e2, -650.0);
    e1 = exp(e1);
    e2 = exp(e2);
    s->a1[i] = (s->u[i][0] / (2.0 * s->u[i][1])) * lam * e1;
    s->a2[i] = (s->u[i][1] / (2.0 * s->u[i][2])) * lam * e2;
    double de1 = -CSAHA2 * fz_xi[i] / pow(T, 2.0), de2 = -CSAHA2 * fz_xii[i] / pow(T, 2.0);
    s->da1[i] = s->a1[i] * (s->du[i][0] - s->du[i][1] + dlam + de1);
    s->da2[i] = s->a2[i] * (s->du[i][1] - s->du[i][2] + dlam + de2);
  }
}

/* stage populations of one element (ION_CALC_TEMPRHO, :1708-1787) */
static void stages(double N, double NE, double a1, double a2, double* ni, double* nii, double* niii) {
  *ni = N / (1.0 + (1.0 / (NE * a1)) * (1.0 + (1.0 / (NE * a2))));
  *nii = N / (NE * a1 + 1.0 + (1.0 / (NE * a2)));
  *niii = N / (1.0 + NE * a2 * (NE * a1 + 1.0));
}

/* ------------------------------------------------ per-height EOS state (DERIVVAR) */
typedef struct {
  double dnhyddtemp_pg, dnhyddtemp_rho, dnhyddpg_temp, dnhyddrho_temp;
  double ddamdtemp_pg, ddamdtemp_rho, ddamdpg_temp, ddamdrho_temp;
  double dnelecdpg_temp, dnelecdtemp_rho, dnelecdtemp_pg, dnelecdtemp_pgas, dnelecdrho_temp;
  double dkldtemp_rho, dkldtemp_pg, dkldpg_temp, dkldrho_temp;
  double dkcdtemp_rho, dkcdtemp_pg, dkcdpg_temp, dkcdrho_temp;
  double dnirdpg_temp, dnirdrho_temp;
} derivvar_t;

static void pack_dv(const derivvar_t* d, double* o) {
  const double* p = &d->dnhyddtemp_pg;
  for (int i = 0; i < 23; ++i) o[i] = p[i];
  o[23] = o[24] = 0.0;
}

/* GET_RHO (chemical_equilibrium.f90:165-278): density from T and Pg.
 * pelec is the electron pressure of the starting guess, as in the reference. */
static int get_rho(const saha_t* s, double T, double PG, double* nhyd, double* nelec,
                   double* pelec, double* dens, double* molecw) {
  double ne = *nelec;
  if (ne != ne) ne = 0.0;
  if (ne < 1.0) ne = PG / (101.0 * KBOL * T);
  *pelec = ne * KBOL * T;
  if (*pelec > PG) ne = (PG * 0.5) / (KBOL * T);
  double err = 1.0, nh = 0.0, nnew = ne, nold = ne;
  int iter = 1;
  while (err > 1.0e-4 && iter <= 100 && nh >= 0.0) {
    nold = nnew;
    double nume = PG - nnew * KBOL * T;
    nh = nume / (DENO_SA * KBOL * T);
    double est, s2 = 0.0, s3 = 0.0, ni, nii, niii;
    stages(nh * ABFAC[0], nnew, s->a1[0], s->a2[0], &ni, &nii, &niii);
    est = niii - ni;
    for (int j = 1; j < FZ_NELEM; ++j) {
      stages(nh * ABFAC[j], nnew, s->a1[j], s->a2[j], &ni, &nii, &niii);
      s2 += nii;
      s3 += niii;
    }
    est = est + s2 + s3;
    nnew = ((est > 0.0 ? est : 0.0) + nold) / 2.0;
    err = fabs((nnew - nold) / nnew);
    iter++;
  }
  if (iter >= 100 || nh < 0.0) return E_EOS;
  ne = nnew;
  nh = (PG - ne * KBOL * T) / (DENO_SA * KBOL * T);
  *dens = nh * DENO_SAM + ne * MELE;
  *molecw = ((*dens * KBOL * T) / MAMU) * pow(PG, -1.0);
  if (nh != nh || nnew != nnew) return E_NAN;
  if (ne < 0.0 || nh <= 0.0) return E_EOS;
  *nelec = ne;
  *nhyd = nh;
  return E_OK;
}

/* GET_NE_TEMP_SNE_ALT (:892-957) with ION_CALC_NETEMP_NEW (:1376-1442) */
static void dne_dpg_t(const saha_t* s, double T, double P, double NE, derivvar_t* dv) {
  double NH = (P - NE * KBOL * T) / (KBOL * T * DENO_SA);
  double sn = 0.0, sd = 0.0;
  for (int j = 0; j < FZ_NELEM; ++j) {
    double N = NH * ABFAC[j], a1 = s->a1[j], a2 = s->a2[j], ni, nii, niii;
    stages(N, NE, a1, a2, &ni, &nii, &niii);
    double deno = ((a1 * NE + 1.0) * a2) * NE + 1.0;
    double q = 2.0 * a1 * a2 * NE + a2;
    double n1 = ni / N * ABFAC[j] / DENO_SA / KBOL / T;
    double n2 = nii / N * ABFAC[j] / DENO_SA / KBOL / T;
    double n3 = niii / N * ABFAC[j] / DENO_SA / KBOL / T;
    double d1 = -ni / N * ABFAC[j] / DENO_SA + N * (2.0 * a1 * a2 * NE) / deno - ni * q / deno;
    double d2 = -nii / N * ABFAC[j] / DENO_SA + N * a2 / deno - nii * q / deno;
    double d3 = -niii / N * ABFAC[j] / DENO_SA + N * 0.0 / deno - niii * q / deno;
    if (j > 0) {
      sn = sn + n2 + n3;
      sd = sd + d2 + d3;
    } else {
      sn = sn - n1 + n3;
      sd = sd - d1 + d3;
    }
  }
  dv->dnelecdpg_temp = sn / (1.0 - sd);
  dv->dnhyddpg_temp = 1.0 / DENO_SA * (1.0 / KBOL / T - dv->dnelecdpg_temp);
}

/* GET_NE_RHO_NIT (:509-570) with ION_CALC_TEMPRHO_NIT (:1876-1970) */
static void dne_dt_rho(const saha_t* s, double T, double RHO, double NE, derivvar_t* dv) {
  double NH = (RHO - NE * MELE) / DENO_SAM;
  double sn = 0.0, sd = 0.0;
  for (int j = 0; j < FZ_NELEM; ++j) {
    double N = NH * ABFAC[j], a1 = s->a1[j], a2 = s->a2[j], da1 = s->da1[j], da2 = s->da2[j];
    double ni, nii, niii;
    stages(N, NE, a1, a2, &ni, &nii, &niii);
    double dn = -MELE / DENO_SAM * ABFAC[j];
    double deno = ((a1 * NE + 1.0) * a2) * NE + 1.0;
    double r = (da1 * a2 * NE + (a1 * NE + 1.0) * da2) * NE / deno;
    double q = 2.0 * a1 * a2 * NE + a2;
    double n1 = N * NE * NE * (da1 * a2 + a1 * da2) / deno - ni * r;
    double n2 = N * NE * da2 / deno - nii * r;
    double n3 = N * 0.0 / deno - niii * r;
    double d1 = dn * ni / N + N * (2.0 * a1 * a2 * NE) / deno - ni * q / deno;
    double d2 = dn * nii / N + N * a2 / deno - nii * q / deno;
    double d3 = dn * niii / N + N * 0.0 / deno - niii * q / deno;
    if (j > 0) {
      sn = sn + n2 + n3;
      sd = sd + d2 + d3;
    } else {
      sn = sn - n1 + n3;
      sd = sd - d1 + d3;
    }
  }
  dv->dnelecdtemp_rho = sn / (1.0 - sd);
}

/* GET_NE_DNEDRHO_TEMP_NIT (:1120-1185) with ION_CALC_NETEMP (:1316-1372) */
static void dne_drho_t(const saha_t* s, double T, double RHO, double NE, derivvar_t* dv) {
  (void)T;
  double NH = (RHO - NE * MELE) / DENO_SAM;
  double num = 0.0, den = 0.0;
  for (int j = 0; j < FZ_NELEM; ++j) {
    double NT = NH * ABFAC[j], a1 = s->a1[j], a2 = s->a2[j];
    double deno = pow(NE, 2.0) * a1 * a2 + NE * a2 + 1.0;
    double fi = pow(NE, 2.0) * a1 * a2 / deno, fii = NE * a2 / deno, fiii = 1.0 / deno;
    double q = NE * 2.0 * a1 * a2 + a2;
    double dfi = fi / NE * 2.0 - fi * q / deno;
    double dfii = a2 / deno - fii * q / deno;
    double dfiii = -fiii * q / deno;
    if (j == 0) {
      num = num + NT * (-dfi + dfiii);
      den = den + ABFAC[j] * (-fi + fiii);
    } else {
      num = num + NT * (dfii + dfiii);
      den = den + ABFAC[j] * (fii + fiii);
    }
  }
  double densum = (1.0 - num) / den;
  dv->dnelecdrho_temp = 1.0 / (densum * DENO_SAM + MELE);
  dv->dnhyddrho_temp = (1.0 - dv->dnelecdrho_temp * MELE) / DENO_SAM;
}

/* one pass of the GET_DNE_PG sum (:2062-2158) with ION_CALC_TEMPPGAS (:1791-1872) */
static double dne_pg_sum(const saha_t* s, double T, double P, double NE, double NH, double dne) {
  double est = 0.0;
  for (int j = 0; j < FZ_NELEM; ++j) {
    double N = NH * ABFAC[j], a1 = s->a1[j], a2 = s->a2[j], da1 = s->da1[j], da2 = s->da2[j];
    double ni, nii, niii;
    stages(N, NE, a1, a2, &ni, &nii, &niii);
    double dn = -1.0 / DENO_SA * (P / (KBOL * T * T) + dne) * ABFAC[j];
    double g1 = 1.0 + (1.0 / (NE * a1)) * (1.0 + 1.0 / (NE * a2));
    double dni = dn * ni / N -
                 N * pow(g1, -2.0) * (1.0 / (pow(NE, 2.0) * a1)) *
                     ((-1.0 / a1) * (dne * a1 + NE * da1) * (1.0 + 1.0 / (NE * a2)) -
                      (1.0 / (NE * pow(a2, 2.0))) * (dne * a2 + NE * da2));
    double g2 = NE * a1 + 1.0 + 1.0 / (NE * a2);
    double dnii = dn * nii / N -
                  N * pow(g2, -2.0) *
                      (dne * a1 + NE * da1 - (dne * a2 + NE * da2) / (pow(NE, 2.0) * pow(a2, 2.0)));
    double g3 = 1.0 + NE * a2 * (NE * a1 + 1.0);
    double dniii = dn * niii / N -
                   N * pow(g3, -2.0) *
                       (2.0 * NE * dne * a1 * a2 + pow(NE, 2.0) * da2 * a1 + pow(NE, 2.0) * a2 * da1 +
                        dne * a2 + NE * da2);
    if (j > 0)
      est = est + dnii + dniii;
    else
      est = est - dni + dniii;
  }
  return est;
}

static void dne_dt_pg(const saha_t* s, double T, double P, double NE, derivvar_t* dv) {
  double NH = (P - NE * KBOL * T) / (KBOL * T * DENO_SA);
  double dold = 0.0, dnew = 0.0, err = 1.0e20;
  int iter = 0;
  while (err > 1.0e-6 && iter <= 100) {
    dold = dnew;
    double est = dne_pg_sum(s, T, P, NE, NH, dnew);
    dnew = (est + dold) / 2.0;
    if (fabs(dnew) > 1.0e40) dnew = 0.0;
    err = fabs((dnew - dold) / dnew);
    iter++;
  }
  dv->dnelecdtemp_pg = dne_pg_sum(s, T, P, NE, NH, dnew);
}

/* GET_DNHYDDTEMP (:2008-2052) alone, at the N_H, n_e of a hydrostatic
 * integration (forward.f90:914-922, HYDROSTATIC=yes). */
static void eos_derivs_only(float T, float P, float D, double ne, derivvar_t* dv, saha_t* s) {
  double T8 = (double)T, P8 = (double)P, d8 = (double)D;
  saha_init(s, T8);
  dne_dpg_t(s, T8, P8, ne, dv);
  dne_dt_rho(s, T8, d8, ne, dv);
  dne_drho_t(s, T8, d8, ne, dv);
  dne_dt_pg(s, T8, P8, ne, dv);
  dv->dnhyddtemp_pg = (-1.0 / DENO_SA) * (P8 / (KBOL * pow(T8, 2.0)) + dv->dnelecdtemp_pg);
  dv->dnhyddtemp_rho = (-MELE / DENO_SAM) * dv->dnelecdtemp_rho;
}

/* GET_NHNEMU_G (:2391-2441, INV_ATMPAR(3)=.FALSE. branch) followed by
 * GET_DNHYDDTEMP (:2008-2052).  Updates rho, mu and pel in place. */
static int eos_height(float* T, float* P, float* D, float* MU, float* PEL, double* nh, double* ne,
                      derivvar_t* dv, saha_t* s) {
  double T8 = (double)*T, P8 = (double)*P, D8 = 0.0, MU8 = 0.0, pelec = 0.0;
  saha_init(s, T8);
  int e = get_rho(s, T8, P8, nh, ne, &pelec, &D8, &MU8);
  if (e) return e;
  *D = (float)D8;
  *MU = (float)MU8;
  *PEL = (float)pelec;
  if (*ne < 0.0) return E_EOS;
  if (RESP) {
    double d8 = (double)*D;
    dne_dpg_t(s, T8, P8, *ne, dv);
    dne_dt_rho(s, T8, d8, *ne, dv);
    dne_drho_t(s, T8, d8, *ne, dv);
    dne_dt_pg(s, T8, P8, *ne, dv);
    dv->dnhyddtemp_pg = (-1.0 / DENO_SA) * (P8 / (KBOL * pow(T8, 2.0)) + dv->dnelecdtemp_pg);
    dv->dnhyddtemp_rho = (-MELE / DENO_SAM) * dv->dnelecdtemp_rho;
  }
  return E_OK;
}

/* ------------------------------------------------ continuum opacity (cont_opacity.f90) */
static int opac_cont(double T8, double lam, double NE, double NH, double* kc, derivvar_t* dv) {
  /* H- free-free, John 1988 (OPAC_HMINUS_FF :48-120) */
  static const double A1[6] = {0, 2483.346, -3449.889, 2200.04, -696.271, 88.283};
  static const double B1[6] = {0, 285.827, -1158.382, 2427.719, -1841.4, 444.517};
  static const double C1[6] = {0, -2054.291, 8746.523, -13651.105, 8624.97, -1863.864};
  static const double D1[6] = {0, 2827.776, -11485.632, 16755.524, -10051.53, 2095.288};
  static const double E1[6] = {0, -1341.537, 5303.609, -7510.494, 4400.067, -901.788};
  static const double F1[6] = {0, 208.952, -812.939, 1132.738, -655.02, 132.985};
  static const double A2[4] = {518.1021, 473.2636, -482.2089, 115.5291};
  static const double B2[4] = {-734.8666, 1443.4137, -737.1616, 169.6374};
  static const double C2[4] = {1021.1775, -1977.3395, 1096.8827, -245.649};
  static const double D2[4] = {-479.0721, 922.3575, -521.1341, 114.243};
  static const double E2[4] = {93.1373, -178.9275, 101.7963, -21.9972};
  static const double F2[4] = {-6.4285, 12.36, -7.0571, 1.5097};
  if (lam < 1800.0) return E_LAMBDA;
  double lm = lam / 1.0e4, th = 5040.0 / T8, dth = -th / T8, p1, p2;
  if (lm < 0.3645) {
    double c[4];
    for (int i = 0; i < 4; ++i)
      c[i] = A2[i] * pow(lm, 2.0) + B2[i] + C2[i] / lm + D2[i] / pow(lm, 2.0) + E2[i] / pow(lm, 3.0) +
             F2[i] / pow(lm, 4.0);
    p1 = c[0] * th + c[1] * pow(th, 1.5) + c[2] * pow(th, 2.0) + c[3] * pow(th, 2.5);
    p2 = (c[0] + 1.5 * c[1] * pow(th, 0.5) + 2.0 * c[2] * th + 2.5 * c[3] * pow(th, 1.5)) * dth;
  } else {
    double c[6];
    for (int i = 0; i < 6; ++i)
      c[i] = A1[i] * pow(lm, 2.0) + B1[i] + C1[i] / lm + D1[i] / pow(lm, 2.0) + E1[i] / pow(lm, 3.0) +
             F1[i] / pow(lm, 4.0);
    p1 = c[0] * th + c[1] * pow(th, 1.5) + c[2] * pow(th, 2.0) + c[3] * pow(th, 2.5) + c[4] * pow(th, 3.0) +
         c[5] * pow(th, 3.5);
    p2 = (c[0] + 1.5 * c[1] * pow(th, 0.5) + 2.0 * c[2] * th + 2.5 * c[3] * pow(th, 1.5) +
          3.0 * c[4] * pow(th, 2.0) + 3.5 * c[5] * pow(th, 2.5)) *
         dth;
  }
  double ff = 1.0e-29 * p1 * KBOL * NE * T8;
  double ff_r = ff * (dv->dnelecdtemp_rho / NE + 1.0 / T8 + (p2 / p1));
  double ff_p = ff * (dv->dnelecdtemp_pg / NE + 1.0 / T8 + (p2 / p1));
  double ff_dp = ff / NE * dv->dnelecdpg_temp, ff_dr = ff / NE * dv->dnelecdrho_temp;
  /* H- bound-free (OPAC_HMINUS_BF :124-183) */
  const double LP = 1.6419, CTE = 0.75e-18, ALPHA = (HPLA * LIGHT / KBOL) * 1.0e4;
  static const double CC[6] = {152.519, 49.534, -118.858, 92.536, -34.194, 4.982};
  double bf = 0.0, bf_r = 0.0, bf_p = 0.0, bf_dp = 0.0, bf_dr = 0.0;
  if (lm < LP) {
    double com = 1.0 / lm - 1.0 / LP;
    double sg = CC[0] + CC[1] * pow(com, 0.5) + CC[2] * com + CC[3] * pow(com, 1.5) + CC[4] * pow(com, 2.0) +
                CC[5] * pow(com, 2.5);
    sg = CTE * sg * pow(lm, 3.0) * pow(com, 1.5);
    double part = pow(T8, -2.5) * exp(ALPHA / (T8 * LP)) * (1.0 - exp(-ALPHA / (T8 * lm)));
    double dpart = (part / T8) * (-2.5 + ALPHA * com / T8 -
                                  (ALPHA / (lm * T8)) * (1.0 / (1.0 - exp(-ALPHA / (lm * T8)))));
    bf = part * sg * NE * KBOL * T8;
    bf_r = bf * (dpart / part + 1.0 / T8 + dv->dnelecdtemp_rho / NE);
    bf_p = bf * (dpart / part + 1.0 / T8 + dv->dnelecdtemp_pg / NE);
    bf_dp = bf / NE * dv->dnelecdpg_temp;
    bf_dr = bf / NE * dv->dnelecdrho_temp;
  }
  /* Thomson (OPAC_THOMSON :187-218) */
  const double CT = (8.0 * DPI / 3.0) * (pow(QELE / LIGHT, 4.0)) / pow(MELE, 2.0);
  double th_ = CT * (NE / NH);
  double th_p = th_ * (dv->dnelecdtemp_pg / NE - dv->dnhyddtemp_pg / NH);
  double th_r = th_ * (dv->dnelecdtemp_rho / NE - dv->dnhyddtemp_rho / NH);
  double th_dp = th_ / NE * dv->dnelecdpg_temp - th_ / NH * dv->dnhyddpg_temp;
  double th_dr = th_ / NE * dv->dnelecdrho_temp - th_ / NH * dv->dnhyddrho_temp;
  /* Rayleigh by H (OPAC_RAYLEIGH_H :221-240) */
  double ray = (5.799e-13 + (1.422e-6 + 2.784 / pow(lam, 2.0)) / pow(lam, 2.0)) / pow(lam, 4.0);
  /* neutral hydrogen bound-bound/bound-free (OPAC_HYDROGEN :243-301); IOPOT and
   * the 1E-8 factors are REAL(4) literals in the reference */
  const double IOPOT = (double)13.595f;
  const double HC1 = (double)1e-8f * IOPOT * EVOLT / (HPLA * LIGHT);
  const double HC2 = pow((double)1e-8f, 3.0) * 64.0 * pow(DPI, 4.0) * MELE * pow(QELE, 10.0) /
                     (3.0 * sqrt(3.0) * pow(LIGHT, 4.0) * pow(HPLA, 6.0));
  const double HC3 = HPLA * LIGHT / (KBOL * (double)1e-8f);
  const double HC4 = IOPOT * EVOLT / KBOL;
  int n0 = 1 + (int)floor(sqrt(HC1 * lam));
  double sum, dsum;
  if (n0 <= 8) {
    sum = exp(HC4 / (T8 * pow((double)n0, 2.0))) * pow((double)n0, -3.0);
    dsum = -HC4 * exp(HC4 / (T8 * pow((double)n0, 2.0))) * pow((double)n0, -5.0) * pow(T8, -2.0);
    for (int i = n0 + 1; i <= 8; ++i) {
      sum = sum + exp(HC4 / (T8 * pow((double)i, 2.0))) * pow((double)i, -3.0);
      dsum = dsum - HC4 * exp(HC4 / (T8 * pow((double)i, 2.0))) * pow((double)i, -5.0) * pow(T8, -2.0);
    }
    sum = sum + (0.117 + exp(HC4 / (T8 * pow(9.0, 2.0)))) * (T8 / (2.0 * HC4));
    dsum = dsum + (1.0 / (2.0 * HC4)) * (0.117 + exp(HC4 / (pow(9.0, 2.0) * T8))) -
           (1.0 / (2.0 * pow(9.0, 2.0) * T8)) * exp(HC4 / (T8 * pow(9.0, 2.0)));
  } else {
    sum = (0.117 + exp(HC4 / (T8 * pow((double)n0, 2.0)))) * (T8 / (2.0 * HC4));
    dsum = (1.0 / (2.0 * HC4)) * (0.117 + exp(HC4 / (pow(9.0, 2.0) * T8))) -
           (1.0 / (2.0 * pow(9.0, 2.0) * T8)) * exp(HC4 / (T8 * pow(9.0, 2.0)));
  }
  double hyd = HC2 * sum * (1.0 - exp(-HC3 / (T8 * lam))) * exp(-HC4 / T8) * pow(lam, 3.0);
  double hyd_d = hyd * (dsum / sum + HC4 / pow(T8, 2.0)) -
                 (HC2 * HC3 * pow(lam, 2.0) / pow(T8, 2.0)) * exp(-HC4 / T8) * exp(-HC3 / (lam * T8));
  /* OPAC_CONTINUUM (:305-357) */
  double fac = ff + bf + th_ + ray + hyd;
  *kc = NH * fac;
  if (RESP) {
    dv->dkcdtemp_rho = fac * dv->dnhyddtemp_rho + NH * (ff_r + bf_r + th_r + hyd_d);
    dv->dkcdtemp_pg = fac * dv->dnhyddtemp_pg + NH * (ff_p + bf_p + th_p + hyd_d);
    dv->dkcdpg_temp = fac * dv->dnhyddpg_temp + NH * (ff_dp + bf_dp + th_dp);
    dv->dkcdrho_temp = fac * dv->dnhyddrho_temp + NH * (ff_dr + bf_dr + th_dr);
  }
  return E_OK;
}

/* ------------------------------------------------ line opacity (line_opacity.f90:34-217) */
static int opac_line(double BL, double BU, int li, float Tf, double NE, double NH, double* kl,
                     derivvar_t* dv, const saha_t* s) {
  const fzc_line* L = &LN[li];
  int z = L->zn - 1, ion = L->ion;
  double T8 = (double)Tf;
  if (ion < 1 || ion > 3) return E_ION;
  double N = NH * pow(10.0, fz_abund[z] - 12.0);
  double a1 = s->a1[z], a2 = s->a2[z], da1 = s->da1[z], da2 = s->da2[z];
  /* ION_CALC_TEMPX (chemical_equilibrium.f90:1511-1597) */
  double deno = pow(NE, 2.0) * a1 * a2 + NE * a2 + 1.0;
  double dd_pg = 2.0 * NE * dv->dnelecdtemp_pg * a1 * a2 + pow(NE, 2.0) * da1 * a2 + pow(NE, 2.0) * a1 * da2 +
                 dv->dnelecdtemp_pg * a2 + NE * da2;
  double dd_rho = 2.0 * NE * dv->dnelecdtemp_rho * a1 * a2 + pow(NE, 2.0) * da1 * a2 +
                  pow(NE, 2.0) * a1 * da2 + dv->dnelecdtemp_rho * a2 + NE * da2;
  double nume, dn_pg, dn_rho;
  if (ion == 1) {
    nume = pow(NE, 2.0) * a1 * a2;
    dn_pg = 2.0 * NE * dv->dnelecdtemp_pg * a1 * a2 + pow(NE, 2.0) * da1 * a2 + pow(NE, 2.0) * a1 * da2;
    dn_rho = 2.0 * NE * dv->dnelecdtemp_rho * a1 * a2 + pow(NE, 2.0) * da1 * a2 + pow(NE, 2.0) * a1 * da2;
  } else if (ion == 2) {
    nume = NE * a2;
    dn_pg = dv->dnelecdtemp_pg * a2 + NE * da2;
    dn_rho = dv->dnelecdtemp_rho * a2 + NE * da2;
  } else {
    nume = 1.0;
    dn_pg = dn_rho = 0.0;
  }
  double npi = N * nume / deno;
  double abf = pow(10.0, fz_abund[z] - 12.0);
  double dnir_pg = (nume / deno) * abf * dv->dnhyddtemp_pg + N * (dn_pg * deno - dd_pg * nume) / pow(deno, 2.0);
  double dnir_rho =
      (nume / deno) * abf * dv->dnhyddtemp_rho + N * (dn_rho * deno - dd_rho * nume) / pow(deno, 2.0);
  double upi = s->u[z][ion - 1], dupi = s->du[z][ion - 1];
  double t1 = sqrt(DPI) * pow(QELE, 2.0) / (MELE * LIGHT) * pow(10.0, L->loggf) * L->lambda0 * 1.0e-8;
  const double Td = T8 + xi_t(z);   /* Doppler temperature */
  double t2 = sqrt(fz_matom[z] * MAMU / (2.0 * KBOL * Td));
  double t3 = BL * npi / upi;
  double t4 = exp(-L->eplow * EVOLT / (KBOL * T8));
  double t5 = 1.0 - (BU / BL) * exp(-HPLA * LIGHT / (1.0e-8 * L->lambda0 * KBOL * T8));
  *kl = t1 * t2 * t3 * t4 * t5;
  if (RESP) {
    /* ION_CALC_PGTEMP (:1446-1507) for d/dPg and d/drho at constant T */
    double dnir[2];
    const double dne_in[2] = {dv->dnelecdpg_temp, dv->dnelecdrho_temp};
    const double dnh_in[2] = {dv->dnhyddpg_temp, dv->dnhyddrho_temp};
    for (int w = 0; w < 2; ++w) {
      double DNE = dne_in[w];
      double dno = pow(NE, 2.0) * a1 * a2 + NE * a2 + 1.0;
      double fi = pow(NE, 2.0) * a1 * a2 / dno, fii = NE * a2 / dno, fiii = 1.0 / dno;
      double denoi = ((a1 * NE + 1.0) * a2) * NE + 1.0;
      double q = (2.0 * a1 * a2 * NE + a2) / denoi * DNE;
      double dfi = 2.0 * a1 * a2 * NE * DNE / denoi - fi * q;
      double dfii = a2 * DNE / denoi - fii * q;
      double dfiii = 0.0 * DNE / denoi - fiii * q;
      double fir = ion == 1 ? fi : (ion == 2 ? fii : fiii);
      double dfir = ion == 1 ? dfi : (ion == 2 ? dfii : dfiii);
      dnir[w] = N / NH * fir * dnh_in[w] + N * dfir;
    }
    dv->dnirdpg_temp = dnir[0];
    dv->dnirdrho_temp = dnir[1];
    double d2 = -0.5 / Td * t2;
    double d3 = BL * (dnir_pg - npi * dupi) / upi;
    double d4 = L->eplow * EVOLT / (KBOL * (double)Tf * (double)Tf) * t4;
    double d5 = -HPLA * LIGHT / (1.0e-8 * L->lambda0 * KBOL * T8 * T8) * (1.0 - t5);
    dv->dkldtemp_pg = t1 * (d2 * t3 * t4 * t5 + t2 * d3 * t4 * t5 + t2 * t3 * d4 * t5 + t2 * t3 * t4 * d5);
    d3 = BL * (dnir_rho - npi * dupi) / upi;
    dv->dkldtemp_rho = t1 * (d2 * t3 * t4 * t5 + t2 * d3 * t4 * t5 + t2 * t3 * d4 * t5 + t2 * t3 * t4 * d5);
    dv->dkldpg_temp = *kl / npi * dv->dnirdpg_temp;
    dv->dkldrho_temp = *kl / npi * dv->dnirdrho_temp;
  }
  return E_OK;
}

/* ------------------------------------------------ air refraction (atom_database.f90:1039-1070) */
static double refrax(double w) {
  const double T = 15.0, P = 760.0, H = 0.0;
  double s = 1.0 / pow(w, 2.0);
  double r = 6.4328e-5 + 2.94981e-2 / (146.0 - s) + 2.554e-4 / (41.0 - s);
  double y = 1.0 / (273.15 + T);
  double al = 3.67e-3 + 3.3e-4 * exp((0.19 - w) / 0.12);
  double wvsp = pow(10.0, -55.1132754 +
                              y * (1.19551822e5 +
                                   y * (-9.70108858e7 +
                                        y * (4.12856913e10 + y * (-9.87931835e12 + y * (1.25868645e15 -
                                                                                       6.66885272e16 * y))))));
  double wvp = 1.0e-2 * H * wvsp;
  return 1.0 + 1.31579e-3 * r * P / (1.0 + al * (T - 1.0) / (1.0 + 15.0 * al)) - 5.49e-8 * wvp / (1.0 + al * T);
}

/* ------------------------------------------------ damping (damping.f90:842-1012) */
static double damping(int li, float Tf, double NH, derivvar_t* dv) {
  const fzc_line* L = &LN[li];
  int z = L->zn - 1;
  double T8 = (double)Tf, lam = L->lambda0 * 1e-8;
  const double Td = T8 + xi_t(z);   /* Doppler temperature */
  double vd = sqrt(2.0 * KBOL * Td / (MAMU * fz_matom[z]));
  double grad = 0.2223 / pow(lam, 2.0);
  double muh = MAMU * (fz_matom[z] * fz_matom[0]) / (fz_matom[z] + fz_matom[0]);
  double muhe = MAMU * (fz_matom[z] * fz_matom[1]) / (fz_matom[z] + fz_matom[1]);
  double heab = pow(10.0, fz_abund[1] - 12.0);
  double nhe = NH * heab;
  const double HEPOL = (double)0.3181818f;
  double gcol, dgp = 0.0, dgr = 0.0;
  if (L->sigma > 0 && L->alpha > 0) {
    double al = L->alpha;
    double e1 = 2.0 - al / 2.0, e2 = (1.0 - al) / 2.0;
    double gf = 1.0 + (-0.5748646 + (0.9512363 + (-0.6998588 + (0.4245549 - 0.1010678 * e1) * e1) * e1) * e1) * e1;
    double beta = pow(4.0 / DPI, al / 2.0) * gf * pow(1.0e6, al) * (L->sigma * pow(RBOHR, 2.0)) *
                  pow(8.0 * KBOL / DPI, e2);
    beta = 2.0 * beta;
    gcol = beta * pow(T8, e2) * (NH * pow(muh, -e2) + HEPOL * nhe * pow(muhe, -e2));
    if (RESP) {
      dgp = gcol * e2 / T8 + beta * pow(T8, e2) * dv->dnhyddtemp_pg * (pow(muh, -e2) + HEPOL * pow(muhe, -e2) * heab);
      dgr = gcol * e2 / T8 + beta * pow(T8, e2) * dv->dnhyddtemp_rho * (pow(muh, -e2) + HEPOL * pow(muhe, -e2) * heab);
    }
  } else {
    double ridx = (double)1.0004f;
    if (lam * 1e8 >= 1800.0) ridx = refrax(lam * 1e4);
    double lvac = ridx * lam;
    double eup = L->eplow + (HPLA * LIGHT / lvac) / EVOLT, ed1, ed2, ch;
    if (L->ion == 1) {
      ed1 = fmax(fz_xi[z] - eup - fz_xi[z] * 0.0, 1.0);
      ed2 = fmax(fz_xi[z] - L->eplow - fz_xi[z] * 0.0, 3.0);
      ch = 1.0;
    } else {
      ed1 = fmax(fz_xii[z] - eup - fz_xi[z] * 1.0, 1.0);
      ed2 = fmax(fz_xii[z] - L->eplow - fz_xi[z] * 1.0, 3.0);
      ch = (double)1.741f;
    }
    ch = ch * 17.0 * pow(8.0 * KBOL / DPI, 0.3);
    ch = ch * pow(pow(10.0, -30.5325) * (1.0 / pow(ed1, 2.0) - 1.0 / pow(ed2, 2.0)), 0.4);
    double chp = ch * (NH * pow(muh, -0.3) + nhe * pow(muhe, -0.3));
    gcol = chp * pow(T8, 0.3);
    if (RESP) {
      dgr = ch * (pow(muh, -0.3) + heab * pow(muhe, -0.3)) * (dv->dnhyddtemp_rho * pow(T8, 0.3) + 0.3 * NH * pow(T8, -0.7));
      dgp = ch * (pow(muh, -0.3) + heab * pow(muhe, -0.3)) * (dv->dnhyddtemp_pg * pow(T8, 0.3) + 0.3 * NH * pow(T8, -0.7));
    }
  }
  double damp = (grad + gcol) * lam / (4.0 * DPI * vd);
  if (RESP) {
    dv->ddamdtemp_pg = damp / (grad + gcol) * dgp - damp / vd * (vd / (2.0 * Td));
    dv->ddamdtemp_rho = damp / (grad + gcol) * dgr - damp / vd * (vd / (2.0 * Td));
    dv->ddamdpg_temp = lam / (4.0 * DPI * vd) * gcol / NH * dv->dnhyddpg_temp;
    dv->ddamdrho_temp = lam / (4.0 * DPI * vd) * gcol / NH * dv->dnhyddrho_temp;
  }
  return damp;
}

/* ------------------------------------------------ Voigt / Faraday-Voigt (voigt.f:17-98) */
static const double XDWS[28] = {.1, .2, .3, .4, .5, .6, .7, .8, .9, 1., 1.2, 1.4, 1.6, 1.8,
                                2., 3., 4., 5., 6., 7., 8., 9., 10., 12., 14., 16., 18., 20.};
static const double YDWS[28] = {
    9.9335991E-02, 1.9475104E-01, 2.8263167E-01, 3.5994348E-01, 4.2443639E-01, 4.7476321E-01,
    5.1050407E-01, 5.3210169E-01, 5.4072434E-01, 5.3807950E-01, 5.0727350E-01, 4.5650724E-01,
    3.9993989E-01, 3.4677279E-01, 3.0134040E-01, 1.7827103E-01, 1.2934799E-01, 1.0213407E-01,
    8.4542692E-02, 7.2180972E-02, 6.3000202E-02, 5.5905048E-02, 5.0253846E-02, 4.1812878E-02,
    3.5806101E-02, 3.1311397E-02, 2.7820844E-02, 2.5031367E-02};

static void voigt1(double a, double g, double* hh, double* ff) {
  /* Hui, Armstrong & Wray (1978) rational approximation of w(z) */
  static const double A[7] = {122.607931777104326, 214.382388694706425, 181.928533092181549,
                              93.155580458138441,  30.180142196210589,  5.912626209773153,
                              .564189583562615};
  static const double B[7] = {122.60793177387535, 352.730625110963558, 457.334478783897737,
                              348.703917719495792, 170.354001821091472, 53.992906912940207,
                              10.479857114260399};
  int sg = (g < 0) ? -1 : 1;
  double v = sg * g, h, f;
  if (a == 0.0) {
    double v2 = v * v, d;
    h = exp(-v2);
    if (v <= XDWS[0]) {
      d = v * (1. - .66666667 * v2);
    } else if (v > XDWS[27]) {
      double y = .5 / v;
      d = y * (1. + y / v);
    } else {
      int k = 27;
      for (int i = 2; i <= 27; ++i)
        if (!(XDWS[i - 1] < v)) {
          k = i;
          break;
        }
      int kk = k - 1, kkk = k + 1;
      double d1 = v - XDWS[kk - 1], d2 = v - XDWS[k - 1], d3 = v - XDWS[kkk - 1];
      double d12 = XDWS[kk - 1] - XDWS[k - 1], d13 = XDWS[kk - 1] - XDWS[kkk - 1], d23 = XDWS[k - 1] - XDWS[kkk - 1];
      d = YDWS[kk - 1] * d2 * d3 / (d12 * d13) - YDWS[k - 1] * d1 * d3 / (d12 * d23) +
          YDWS[kkk - 1] * d1 * d2 / (d13 * d23);
    }
    f = sg * 5.641895836E-1 * d;
  } else {
    /* z = a - i v; complex Horner for numerator and denominator */
    double zr = a, zi = -v;
    double nr = A[6], ni = 0.0, dr = 1.0, di = 0.0, t;
    for (int k = 5; k >= 0; --k) {
      t = nr * zr - ni * zi + A[k];
      ni = nr * zi + ni * zr;
      nr = t;
    }
    dr = zr + B[6];
    di = zi;
    for (int k = 5; k >= 0; --k) {
      t = dr * zr - di * zi + B[k];
      di = dr * zi + di * zr;
      dr = t;
    }
    double den = dr * dr + di * di;
    h = (nr * dr + ni * di) / den;
    f = .5 * sg * ((ni * dr - nr * di) / den);
  }
  *hh = h;
  *ff = f * 2.0;
}

void fzc_voigt(int n, double a, const double* v, double* h, double* f) {
  for (int i = 0; i < n; ++i) voigt1(a, v[i], &h[i], &f[i]);
}

/* ------------------------------------------------ Zeeman pattern (magsplit.f90:38-187) */
static int zeeman(const fzc_line* L, double* shift, double* str, int* type) {
  int npi = (int)(2 * fmin(L->jl, L->ju) + 1), nsr = (int)(L->jl + L->ju), nsb = nsr;
  int ntot = npi + nsr + nsb;
  int dj = (int)(L->ju - L->jl);
  if (abs(dj) > 1 || ntot > FZC_MAXCOMP) return -1;
  double gl = 0.0, gu = 0.0;
  if (L->jl != 0) gl = 1.5 + (L->sl * (L->sl + 1.0) - L->ll * (L->ll + 1.0)) / (2.0 * L->jl * (L->jl + 1.0));
  if (L->ju != 0) gu = 1.5 + (L->su * (L->su + 1.0) - L->lu * (L->lu + 1.0)) / (2.0 * L->ju * (L->ju + 1.0));
  int nml = (int)(2 * L->jl + 1), nmu = (int)(2 * L->ju + 1), c = 0;
  for (int j = 0; j < nml; ++j) {
    double ml = (double)j - L->jl;
    for (int k = 0; k < nmu; ++k) {
      double mu = (double)k - L->ju;
      int dm = (int)(mu - ml);
      if (abs(dm) > 1) continue;
      type[c] = dm;
      shift[c] = (double)(float)(4.66685e-10 * (ml * gl - mu * gu) * pow(L->lambda0, 2));
      double s = 0.0, jl = L->jl, ju = L->ju;
      if (dj == -1) {
        if (dm == -1) s = (jl + mu) * (ju + mu + 2.0);
        else if (dm == 0) s = 2.0 * (pow(jl, 2.0) - pow(mu, 2.0));
        else s = (jl - mu) * (ju - mu + 2.0);
      } else if (dj == 0) {
        if (dm == -1) s = (ju - mu) * (ju + mu + 1.0);
        else if (dm == 0) s = 2.0 * pow(mu, 2.0);
        else s = (ju + mu) * (ju - mu + 1.0);
      } else {
        if (dm == -1) s = (ju - mu) * (jl - mu);
        else if (dm == 0) s = 2.0 * (pow(ju, 2.0) - pow(mu, 2.0));
        else s = (ju + mu) * (jl + mu);
      }
      str[c] = s;
      c++;
    }
  }
  double sum[3] = {0, 0, 0};
  for (int j = 0; j < c; ++j) sum[type[j] + 1] += str[j];
  for (int j = 0; j < c; ++j) str[j] = str[j] / sum[type[j] + 1];
  return c;
}

int fzc_zeeman(int line, double* shift, double* strength, int32_t* type) {
  if (line < 0 || line >= NLN) return -1;
  for (int i = 0; i < ZN_N[line]; ++i) {
    shift[i] = ZN_SHIFT[line][i];
    strength[i] = ZN_STR[line][i];
    type[i] = ZN_TYPE[line][i];
  }
  return ZN_N[line];
}

/* ------------------------------------------------ setup */
int fzc_setup(const fzc_line* lines, int nlines, const fzc_region* regs, int nreg, int resp,
              const int32_t* atmpar) {
  fzc_set_departure(0, 0, NULL, NULL);   /* LTE until set (input_file.f90:504-508) */
  init_constants();
  if (nlines > FZC_MAXLINES || nreg > FZC_MAXREG || nlines < 1 || nreg < 1) return -E_CONFIG;
  NLN = nlines;
  memcpy(LN, lines, sizeof(fzc_line) * nlines);
  NRG = nreg;
  memcpy(RG, regs, sizeof(fzc_region) * nreg);
  RESP = resp;
  for (int i = 0; i < 8; ++i) ATMPAR[i] = atmpar ? atmpar[i] : 0;
  NUMW = 0;
  for (int r = 0; r < nreg; ++r) {
    if (RG[r].nblend < 1 || RG[r].nblend > FZC_MAXBLEND) return -E_CONFIG;
    PIX0[r] = NUMW;
    NUMW += RG[r].nwave;
    PIX1[r] = NUMW;
    for (int b = 0; b < RG[r].nblend; ++b)
      BLDIFF[r][b] = b == 0 ? 0.0 : 1.0e3 * (LN[RG[r].line[b]].lambda0 - LN[RG[r].line[0]].lambda0);
  }
  free(WAVE);
  free(GAIN);
  fzc_set_lsf(NULL, NULL);
  WAVE = (double*)malloc(sizeof(double) * NUMW);
  GAIN = (float*)malloc(sizeof(float) * 4 * NUMW);
  for (int r = 0; r < nreg; ++r) /* pre_post_duties.f90:592-621 */
    for (int j = PIX0[r]; j < PIX1[r]; ++j) WAVE[j] = RG[r].wave_start + RG[r].wave_step * (j - PIX0[r]);
  for (int j = 0; j < 4 * NUMW; ++j) GAIN[j] = 1.0f;
  for (int l = 0; l < nlines; ++l) {
    ZN_N[l] = zeeman(&LN[l], ZN_SHIFT[l], ZN_STR[l], ZN_TYPE[l]);
    if (ZN_N[l] < 0) return -E_CONFIG;
  }
  return NUMW;
}

int fzc_numw(void) { return NUMW; }
void fzc_wave(double* w) { memcpy(w, WAVE, sizeof(double) * NUMW); }
void fzc_set_gain(const float* g) { memcpy(GAIN, g, sizeof(float) * 4 * NUMW); }
void fzc_set_hydrostatic(int on) { HYDRO = on ? 1 : 0; }
void fzc_set_microturbulence(double xi) { XI2 = xi * xi; }
void fzc_set_full_stokes(double* buf) { FSOUT = buf; }
void fzc_set_line_tau(double* buf) { TLOUT = buf; }

/* lsf_kernel: NUMW doubles, region r's taps YK/SKI at PIX0[r]..PIX1[r]-1 with
 * tap 0 first (LSF_INIT builds them so: pre_post_duties.f90:855-876); valid:
 * NRG flags; NULL switches the convolution off. */
int fzc_set_lsf(const double* lsf_kernel, const int32_t* valid) {
  free(LSFK);
  LSFK = NULL;
  LSF_ON = 0;
  if (!lsf_kernel) return 0;
  for (int r = 0; r < NRG; ++r)
    if (PIX1[r] - PIX0[r] > 4096) return -E_CONFIG;
  LSFK = (double*)malloc(sizeof(double) * NUMW);
  memcpy(LSFK, lsf_kernel, sizeof(double) * NUMW);
  for (int r = 0; r < NRG; ++r) LSFV[r] = valid ? (valid[r] != 0) : 1;
  LSF_ON = 1;
  return 0;
}

/* One Stokes row x[NUMW] (stride st) through the LSF: per valid region the
 * reference multiplies the r2c DFT of the row by the DFT of LSF_KERNEL and
 * returns with an unnormalised c2r (forward.f90:621-657, 814-844), i.e. the
 * circular convolution y[i] = SKI * sum_j x[j] LSF_KERNEL[(i - j) mod SKI];
 * the result is stored through REAL() (single precision). */
static void lsf_row(double* x, size_t st) {
  double y[4096];
  for (int r = 0; r < NRG; ++r) {
    if (!LSFV[r]) continue;
    const int w0 = PIX0[r], n = PIX1[r] - PIX0[r];
    for (int i = 0; i < n; ++i) {
      double acc = 0.0;
      for (int j = 0; j < n; ++j) {
        int t = i - j;
        if (t < 0) t += n;
        acc += x[(size_t)(w0 + j) * st] * LSFK[w0 + t];
      }
      y[i] = acc * (double)n;
    }
    for (int i = 0; i < n; ++i) x[(size_t)(w0 + i) * st] = (double)(float)y[i];
  }
}

/* ------------------------------------------------ unit entry points */
int fzc_eos(float T, float P, double* ne, double* out, double* d) {
  saha_t s;
  derivvar_t dv = {0};
  float D = 0, MU = 0, PEL = 0;
  double nh = 0.0;
  int e = eos_height(&T, &P, &D, &MU, &PEL, &nh, ne, &dv, &s);
  out[0] = nh;
  out[1] = *ne;
  out[2] = D;
  out[3] = MU;
  out[4] = PEL;
  pack_dv(&dv, d);
  return e;
}

/* these two use DERIVVAR values passed in d (as left by an EOS call) */
static void unpack_dv(const double* o, derivvar_t* d) {
  double* p = &d->dnhyddtemp_pg;
  for (int i = 0; i < 23; ++i) p[i] = o[i];
}

int fzc_opac_cont(float T, double lam, double ne, double nh, double* kc, double* d) {
  derivvar_t dv;
  unpack_dv(d, &dv);
  int e = opac_cont((double)T, lam, ne, nh, kc, &dv);
  pack_dv(&dv, d);
  return e;
}

int fzc_opac_line(int line, float T, double ne, double nh, double* kl, double* d) {
  derivvar_t dv;
  saha_t s;
  unpack_dv(d, &dv);
  saha_init(&s, (double)T);
  int e = opac_line(1.0, 1.0, line, T, ne, nh, kl, &dv, &s);
  pack_dv(&dv, d);
  return e;
}

int fzc_damping(int line, float T, double nh, double* damp, double* d) {
  derivvar_t dv;
  unpack_dv(d, &dv);
  *damp = damping(line, T, nh, &dv);
  pack_dv(&dv, d);
  return 0;
}

/* ------------------------------------------------ absorption matrix
 * Seven propagation-matrix elements (ETAI,ETAQ,ETAU,ETAV,RHOQ,RHOU,RHOV) and
 * their derivatives w.r.t. the eight physical parameters in the order the
 * evolution operator chains them: Bx, By, Bz, vlos, T|Pg, T|rho, Pg|T, rho|T
 * (absorption_matrix.f90:1717-2642 with the blend bookkeeping of
 * ABS_MAT_FIRST_STEP/END_STEP, :293-845). */
enum { P_BX, P_BY, P_BZ, P_VLOS, P_TPG, P_TRHO, P_PGT, P_RHOT, NPD };

typedef struct {
  double* eta;   /* [7][numw] */
  double* deta;  /* [NPD][7][numw] */
  double* teta;  /* blend scratch [7][numw] */
  double* tdeta; /* [NPD][7][numw] */
  double* hh;    /* per-pixel scratch */
  double* ff;
  double* uu;
} absmat_t;

#define ETA(e, c, l) (e)[(c)*NUMW + (l)]
#define DETA(d, p, c, l) (d)[((p)*7 + (c)) * NUMW + (l)]

static int abs_mat_block(int K, int L, int B, float Tf, const float* bx, const float* by, const float* bz,
                         const float* vz, double kc, double klin, double damp, const derivvar_t* dv,
                         absmat_t* A) {
  (void)K;
  int li = RG[L].line[B];
  const fzc_line* LL = &LN[li];
  double wdif = BLDIFF[L][B];
  int p0 = PIX0[L], p1 = PIX1[L];
  double* e = A->teta;
  double* de = A->tdeta;
  for (int c = 0; c < 7; ++c)
    for (int l = p0; l < p1; ++l) ETA(e, c, l) = 0.0;
  for (int p = 0; p < NPD; ++p)
    for (int c = 0; c < 7; ++c)
      for (int l = p0; l < p1; ++l) DETA(de, p, c, l) = 0.0;
  double T8 = (double)Tf;
  const double Td = T8 + xi_t(LL->zn - 1);   /* Doppler temperature */
  double vdop = sqrt(2.0 * KBOL * Td / (fz_matom[LL->zn - 1] * MAMU));
  double dld = 1.0e3 * (vdop * LL->lambda0 / LIGHT);
  double BX = bx[0], BY = by[0], BZ = bz[0];
  double bx2 = pow(BX, 2.0), by2 = pow(BY, 2.0), bz2 = pow(BZ, 2.0);
  double bxm = bx2 - by2, bxp = bx2 + by2;
  double b2 = pow(BX, 2.0) + pow(BY, 2.0) + pow(BZ, 2.0), bb = sqrt(b2);
  double cg = 0.0, cgf = 1.0, sgf = 1.0, c2p = 0.0, s2p = 0.0;
  double dcgf[3] = {0, 0, 0}, dsgf[3] = {0, 0, 0}, dc2p[3] = {0, 0, 0}, ds2p[3] = {0, 0, 0};
  double dcg[3] = {0, 0, 0}, db[3] = {0, 0, 0};
  if (bb > 0) {
    cg = BZ / bb;
    cgf = 1 + bz2 / b2;
    sgf = 1 - bz2 / b2;
    dcgf[0] = -2.0 * BX * bz2 / b2 / b2;
    dcgf[1] = -2.0 * BY * bz2 / b2 / b2;
    dcgf[2] = (2.0 * BZ * (b2 - bz2)) / b2 / b2;
    for (int i = 0; i < 3; ++i) dsgf[i] = -dcgf[i];
    dc2p[0] = 4.0 * BX * by2 / bxp / bxp;
    dc2p[1] = -4.0 * BY * bx2 / bxp / bxp;
    dc2p[2] = 0.0;
    ds2p[0] = -2.0 * BY * bxm / bxp / bxp;
    ds2p[1] = 2.0 * BX * bxm / bxp / bxp;
    ds2p[2] = 0.0;
    dcg[0] = -BZ * BX / bb / b2;
    dcg[1] = -BZ * BY / bb / b2;
    dcg[2] = (1.0 - cg * cg) / bb;
    db[0] = BX / bb;
    db[1] = BY / bb;
    db[2] = BZ / bb;
  }
  if (BX != 0 || BY != 0) {
    c2p = bxm / bxp;
    s2p = 2. * BX * BY / bxp;
  }
  if (is_bad(cg) || is_bad(cgf) || is_bad(sgf) || is_bad(c2p) || is_bad(s2p)) return E_NAN;
  double dlv = -1.0e3 * LL->lambda0 * (double)vz[0] / LIGHT;
  double duudv = (-1.0e3 * LL->lambda0 / LIGHT) / dld;
  int nc = ZN_N[li];
  for (int m = 0; m < nc; ++m) {
    double str = ZN_STR[li][m];
    int ty = ZN_TYPE[li][m];
    /* geometry coefficients of this component type and their B derivatives */
    double ci, cq, cu, cv, dci[3], dcq[3], dcu[3], dcv[3];
    if (ty == 0) {
      ci = 0.5 * sgf;
      cq = 0.5 * sgf * c2p;
      cu = 0.5 * sgf * s2p;
      cv = 0.0;
      for (int i = 0; i < 3; ++i) {
        dci[i] = 0.5 * dsgf[i];
        dcq[i] = 0.5 * (dsgf[i] * c2p + sgf * dc2p[i]);
        dcu[i] = 0.5 * (dsgf[i] * s2p + sgf * ds2p[i]);
        dcv[i] = 0.0;
      }
    } else {
      double sv = (ty == -1) ? 0.5 : -0.5;
      ci = 0.25 * cgf;
      cq = -0.25 * sgf * c2p;
      cu = -0.25 * sgf * s2p;
      cv = sv * cg;
      for (int i = 0; i < 3; ++i) {
        dci[i] = 0.25 * dcgf[i];
        dcq[i] = -0.25 * (dsgf[i] * c2p + sgf * dc2p[i]);
        dcu[i] = -0.25 * (dsgf[i] * s2p + sgf * ds2p[i]);
        dcv[i] = sv * dcg[i];
      }
    }
    const double ce[7] = {ci, cq, cu, cv, cq, cu, cv};
    const double* dce[7] = {dci, dcq, dcu, dcv, dcq, dcu, dcv};
    double duudb = ZN_SHIFT[li][m] / dld;
    for (int l = p0; l < p1; ++l) {
      double dlb = ZN_SHIFT[li][m] * bb;
      double u = (WAVE[l] - wdif + dlb + dlv) / dld;
      if (is_bad(u)) return E_NAN;
      double hv, fv;
      voigt1(damp, u, &hv, &fv);
      if (is_bad(hv) || is_bad(fv)) return E_NAN;
      double dh_dv = 2.0 * (-u * hv + damp * fv) * str;
      double dl_dd = dh_dv;
      double dh_dd = 2.0 * (-1.0 / sqrt(DPI) + damp * hv + u * fv) * str;
      double dl_dv = -dh_dd;
      double duudt = -0.5 / Td * u;
      double shh = str * hv, sff = str * fv;
      for (int c = 0; c < 7; ++c) {
        int isf = c >= 4;
        double sv = isf ? sff : shh;
        double du = ce[c] * (isf ? dl_dv : dh_dv);
        double da = ce[c] * (isf ? dl_dd : dh_dd);
        ETA(e, c, l) += ce[c] * sv;
        for (int i = 0; i < 3; ++i) DETA(de, P_BX + i, c, l) += sv * dce[c][i];
        DETA(de, P_TPG, c, l) += du * duudt + da * dv->ddamdtemp_pg;
        DETA(de, P_TRHO, c, l) += du * duudt + da * dv->ddamdtemp_rho;
        for (int i = 0; i < 3; ++i) DETA(de, P_BX + i, c, l) += du * duudb * db[i];
        DETA(de, P_VLOS, c, l) += du * duudv;
        DETA(de, P_PGT, c, l) += da;
        DETA(de, P_RHOT, c, l) += da;
      }
    }
  }
  /* line strength, continuum and the T/Pg/rho chain (:2467-2629) */
  for (int l = p0; l < p1; ++l) {
    for (int c = 0; c < 7; ++c) {
      double x = ETA(e, c, l);
      double cc = (c == 0);
      DETA(de, P_TPG, c, l) = cc * dv->dkcdtemp_pg + dv->dkldtemp_pg * x + klin * DETA(de, P_TPG, c, l);
      DETA(de, P_TRHO, c, l) = cc * dv->dkcdtemp_rho + dv->dkldtemp_rho * x + klin * DETA(de, P_TRHO, c, l);
      DETA(de, P_PGT, c, l) = cc * dv->dkcdpg_temp + dv->dkldpg_temp * x + klin * DETA(de, P_PGT, c, l) * dv->ddamdpg_temp;
      DETA(de, P_RHOT, c, l) = cc * dv->dkcdrho_temp + dv->dkldrho_temp * x + klin * DETA(de, P_RHOT, c, l) * dv->ddamdrho_temp;
      ETA(e, c, l) = (c == 0 ? kc : 0.0) + klin * x;
      for (int p = P_BX; p <= P_VLOS; ++p) DETA(de, p, c, l) = klin * DETA(de, p, c, l);
    }
  }
  /* blend bookkeeping (ABS_MAT_END_STEP): continuum counted once */
  for (int l = p0; l < p1; ++l) {
    for (int c = 0; c < 7; ++c) {
      if (B == 0) {
        ETA(A->eta, c, l) = ETA(e, c, l);
        for (int p = 0; p < NPD; ++p) DETA(A->deta, p, c, l) = DETA(de, p, c, l);
      } else {
        if (c == 0)
          ETA(A->eta, c, l) = (ETA(e, c, l) - kc) + ETA(A->eta, c, l);
        else
          ETA(A->eta, c, l) = ETA(e, c, l) + ETA(A->eta, c, l);
        for (int p = 0; p < NPD; ++p) {
          double sub = 0.0;
          if (c == 0) {
            if (p == P_TPG) sub = dv->dkcdtemp_pg;
            if (p == P_TRHO) sub = dv->dkcdtemp_rho;
            if (p == P_PGT) sub = dv->dkcdpg_temp;
            if (p == P_RHOT) sub = dv->dkcdrho_temp;
          }
          DETA(A->deta, p, c, l) = (DETA(de, p, c, l) - sub) + DETA(A->deta, p, c, l);
        }
      }
    }
  }
  return E_OK;
}

/* ------------------------------------------------ evolution operator (rtesolver.f90:269-1004) */
typedef double m44[4][4];

static void mzero(m44 m) { memset(m, 0, sizeof(m44)); }

/* M2 (or M3 with swap=1) of Landi Degl'Innocenti & Landi Degl'Innocenti 1985 from
 * scalars l1,l2,sig and the vectors eq=(eta_Q,U,V), rq=(rho_Q,U,V); not divided by theta */
static void build_m23(m44 M, int third, double l1, double l2, double sig, const double* e, const double* r) {
  mzero(M);
  if (!third) {
    M[1][0] = l2 * e[0] - sig * l1 * r[0];
    M[2][0] = l2 * e[1] - sig * l1 * r[1];
    M[3][0] = l2 * e[2] - sig * l1 * r[2];
    M[2][1] = sig * l1 * e[2] + l2 * r[2];
    M[3][1] = -sig * l1 * e[1] - l2 * r[1];
    M[3][2] = sig * l1 * e[0] + l2 * r[0];
  } else {
    M[1][0] = l1 * e[0] + sig * l2 * r[0];
    M[2][0] = l1 * e[1] + sig * l2 * r[1];
    M[3][0] = l1 * e[2] + sig * l2 * r[2];
    M[2][1] = -sig * l2 * e[2] + l1 * r[2];
    M[3][1] = sig * l2 * e[1] - l1 * r[1];
    M[3][2] = -sig * l2 * e[0] + l1 * r[0];
  }
  M[0][1] = M[1][0];
  M[0][2] = M[2][0];
  M[1][2] = -M[2][1];
  M[0][3] = M[3][0];
  M[1][3] = -M[3][1];
  M[2][3] = -M[3][2];
}

typedef struct {
  m44 evol;
  m44 dev[NPD]; /* derivatives w.r.t. Bx,By,Bz,vlos,T|Pg,T|rho,Pg|T,rho|T */
} evol_t;

static int evol_operator(int K, int nz, const double* zz, const double* x7, const double* dx /*[NPD][7]*/,
                         evol_t* out) {
  double ei = x7[0], eq = x7[1], eu = x7[2], ev = x7[3], rq = x7[4], ru = x7[5], rv = x7[6];
  int imat = 0;
  for (int i = 0; i < 7; ++i)
    if (x7[i] == x7[i] + 1.0) imat = 1;
  if (imat) ei = eq = eu = ev = rq = ru = rv = 1.0e15;
  if (is_bad(rv)) return E_NAN; /* CHECKNAN keeps only the last test (:339-352) */
  double sig = 0.0;
  double eta2 = pow(eq, 2.) + pow(eu, 2.) + pow(ev, 2.);
  double rho2 = pow(rq, 2.) + pow(ru, 2.) + pow(rv, 2.);
  double etarho = eq * rq + eu * ru + ev * rv;
  double rp = sqrt((pow(eta2 - rho2, 2.)) / 4. + pow(etarho, 2.)) + (eta2 - rho2) / 2.;
  double l1 = sqrt(rp);
  rp = sqrt((pow(eta2 - rho2, 2.)) / 4. + pow(etarho, 2.)) - (eta2 - rho2) / 2.;
  double l2 = sqrt(rp);
  if (etarho > 0) sig = 1.;
  if (etarho < 0) sig = -1.;
  double theta = 2. * sqrt((pow(eta2 - rho2, 2.)) / 4. + pow(etarho, 2.));
  const double e3[3] = {eq, eu, ev}, r3[3] = {rq, ru, rv};
  m44 M2, M3, M4;
  build_m23(M2, 0, l1, l2, sig, e3, r3);
  build_m23(M3, 1, l1, l2, sig, e3, r3);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      M2[i][j] /= theta;
      M3[i][j] /= theta;
    }
  mzero(M4);
  M4[0][0] = (eta2 + rho2) / 2.;
  M4[1][0] = ev * ru - eu * rv;
  M4[2][0] = eq * rv - ev * rq;
  M4[3][0] = eu * rq - eq * ru;
  M4[0][1] = -M4[1][0];
  M4[1][1] = pow(eq, 2.) + pow(rq, 2.) - M4[0][0];
  M4[2][1] = eq * eu + rq * ru;
  M4[3][1] = ev * eq + rv * rq;
  M4[0][2] = -M4[2][0];
  M4[1][2] = M4[2][1];
  M4[2][2] = pow(eu, 2.) + pow(ru, 2.) - M4[0][0];
  M4[3][2] = eu * ev + ru * rv;
  M4[0][3] = -M4[3][0];
  M4[1][3] = M4[3][1];
  M4[2][3] = M4[3][2];
  M4[3][3] = pow(ev, 2.) + pow(rv, 2.) - M4[0][0];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) M4[i][j] = 2. * M4[i][j] / theta;
  /* layer thickness uses the spacing to the next height (:439-443) */
  double dz = (K == nz - 1) ? 1.0e5 * fabs(zz[K] - zz[K - 1]) : 1.0e5 * fabs(zz[K + 1] - zz[K]);
  if (theta == 0) {
    mzero(M2);
    mzero(M3);
    mzero(M4);
  }
  double ch = cosh(l1 * dz), sh = sinh(l1 * dz);
  if (ch == ch + 1.0) ch = 1.0e15;
  if (sh == sh + 1.0) sh = 1.0e15;
  double c2 = cos(l2 * dz), s2 = sin(l2 * dz), ex = exp(-ei * dz);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      out->evol[i][j] = ex * (0.5 * (ch + c2) * (i == j) - s2 * M2[i][j] - sh * M3[i][j] + 0.5 * (ch - c2) * M4[i][j]);
  if (imat)
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) out->evol[i][j] = (i == j);
  if (!RESP) return E_OK;
  /* derivatives (:476-1002) */
  double dif = eta2 - rho2;
  double a1 = sqrt(0.25 * pow(dif, 2) + pow(etarho, 2));
  double a2 = 0.5 / sqrt(a1 + 0.5 * dif);
  if (a2 == a2 + 1.0) a2 = 0.0;
  double a3 = 0.5 / sqrt(a1 - 0.5 * dif);
  if (is_bad(a3)) a3 = 0.0;
  if (a1 == 0.0) {
    a1 = 0.0;
    a3 = 0.0;
  } else {
    a1 = 1.0 / a1;
  }
  double er2 = 2. * etarho;
  /* x = (eta_Q, eta_U, eta_V, rho_Q, rho_U, rho_V) */
  const double xv[6] = {eq, eu, ev, rq, ru, rv};
  m44 dE[7];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) dE[0][i][j] = -dz * out->evol[i][j];
  for (int q = 0; q < 6; ++q) {
    int isr = q >= 3, k = q % 3;
    double dth = isr ? a1 * (-dif * xv[q] + er2 * xv[k]) : a1 * (dif * xv[q] + er2 * xv[3 + k]);
    double dl1 = isr ? a2 * (0.5 * dth - xv[q]) : a2 * (0.5 * dth + xv[q]);
    double dl2 = isr ? a3 * (0.5 * dth + xv[q]) : a3 * (0.5 * dth - xv[q]);
    m44 dM2, dM3, dM4;
    mzero(dM2);
    mzero(dM3);
    mzero(dM4);
    if (theta != 0.0) {
      /* bilinear parts: M(dl1,dl2; eta,rho) + M(l1,l2; d eta, d rho) */
      double de3[3] = {0, 0, 0}, dr3[3] = {0, 0, 0};
      if (isr) dr3[k] = 1.0;
      else de3[k] = 1.0;
      m44 t1, t2;
      build_m23(t1, 0, dl1, dl2, sig, e3, r3);
      build_m23(t2, 0, l1, l2, sig, de3, dr3);
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) dM2[i][j] = (-dth * M2[i][j] + (t1[i][j] + t2[i][j])) / theta;
      build_m23(t1, 1, dl1, dl2, sig, e3, r3);
      build_m23(t2, 1, l1, l2, sig, de3, dr3);
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) dM3[i][j] = (-dth * M3[i][j] + (t1[i][j] + t2[i][j])) / theta;
      /* M4 is quadratic: derivative of the raw entries (before 2/theta) */
      m44 r4;
      mzero(r4);
      double dq[6] = {0, 0, 0, 0, 0, 0};
      dq[q] = 1.0;
      double deq = dq[0], deu = dq[1], dev = dq[2], drq = dq[3], dru = dq[4], drv = dq[5];
      double d00 = eq * deq + eu * deu + ev * dev + rq * drq + ru * dru + rv * drv;
      r4[0][0] = d00;
      r4[1][0] = dev * ru + ev * dru - deu * rv - eu * drv;
      r4[2][0] = deq * rv + eq * drv - dev * rq - ev * drq;
      r4[3][0] = deu * rq + eu * drq - deq * ru - eq * dru;
      r4[0][1] = -r4[1][0];
      r4[1][1] = 2. * eq * deq + 2. * rq * drq - d00;
      r4[2][1] = deq * eu + eq * deu + drq * ru + rq * dru;
      r4[3][1] = dev * eq + ev * deq + drv * rq + rv * drq;
      r4[0][2] = -r4[2][0];
      r4[1][2] = r4[2][1];
      r4[2][2] = 2. * eu * deu + 2. * ru * dru - d00;
      r4[3][2] = deu * ev + eu * dev + dru * rv + ru * drv;
      r4[0][3] = -r4[3][0];
      r4[1][3] = r4[3][1];
      r4[2][3] = r4[3][2];
      r4[3][3] = 2. * ev * dev + 2. * rv * drv - d00;
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) dM4[i][j] = (-dth * M4[i][j] + 2. * r4[i][j]) / theta;
    }
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        double t = 0.5 * (sh * dl1 - s2 * dl2) * dz * (i == j) - c2 * dz * dl2 * M2[i][j] - s2 * dM2[i][j] -
                   ch * dz * dl1 * M3[i][j] - sh * dM3[i][j] + 0.5 * (sh * dl1 + s2 * dl2) * dz * M4[i][j] +
                   0.5 * (ch - c2) * dM4[i][j];
        dE[1 + q][i][j] = ex * t;
      }
  }
  for (int p = 0; p < NPD; ++p) {
    mzero(out->dev[p]);
    if (imat) continue;
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        double s = 0.0;
        for (int c = 0; c < 7; ++c) s += dE[c][i][j] * dx[p * 7 + c];
        out->dev[p][i][j] = s;
      }
  }
  return E_OK;
}

/* unit entry for tests: evolution operator of one layer of thickness dz_km
 * (a 2-height grid) and its derivatives along 8 directions dx[8][7] */
int fzc_evol(const double* x7, double dz_km, const double* dx, double* evol, double* dev) {
  double zz[2] = {0.0, dz_km};
  evol_t E;
  int saved = RESP;
  RESP = 1;
  int e = evol_operator(0, 2, zz, x7, dx, &E);
  RESP = saved;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      evol[i * 4 + j] = E.evol[i][j];
      for (int p = 0; p < NPD; ++p) dev[(p * 4 + i) * 4 + j] = E.dev[p][i][j];
    }
  return e;
}

/* ------------------------------------------------ NLTE departure coefficients
 * POPL(NUML,NZ), POPU(NUML,NZ) (forward_param.f90:47), 1 unless set
 * (input_file.f90:504-508; READ_DEPARTURE_COEFFICIENTS :2146-2166): per
 * spectral region and height, used by OPAC_LINE for every line of the region
 * (forward.f90:949) and by SOURCE_FUNCTION(_DER) (rtesolver.f90:1030-1064). */
static double* POPL_A = NULL;
static double* POPU_A = NULL;
static int POP_NZ = 0, POP_NREG = 0;

int fzc_set_departure(int nreg, int nz, const double* popl, const double* popu) {
  free(POPL_A);
  free(POPU_A);
  POPL_A = POPU_A = NULL;
  POP_NZ = POP_NREG = 0;
  if (!popl || !popu) return 0;
  POPL_A = (double*)malloc(sizeof(double) * nreg * nz);
  POPU_A = (double*)malloc(sizeof(double) * nreg * nz);
  memcpy(POPL_A, popl, sizeof(double) * nreg * nz);
  memcpy(POPU_A, popu, sizeof(double) * nreg * nz);
  POP_NZ = nz;
  POP_NREG = nreg;
  return 0;
}

static double popl_at(int r, int k) { return POPL_A ? POPL_A[r + (size_t)k * POP_NREG] : 1.0; }
static double popu_at(int r, int k) { return POPU_A ? POPU_A[r + (size_t)k * POP_NREG] : 1.0; }

/* ------------------------------------------------ source function (rtesolver.f90:1008-1066) */
static int region_of(int l) {
  for (int m = 0; m < NRG; ++m)
    if (l >= PIX0[m] && l < PIX1[m]) return m;
  return 0;
}

static double region_lambda_cm(int l) {
  for (int m = 0; m < NRG; ++m)
    if (l >= PIX0[m] && l < PIX1[m]) return (LN[RG[m].line[0]].lambda0 + WAVE[l] / 1.0e3) * 1.0e-8;
  return 0.0;
}

static double source_fn(double lam, float Tf, double bl, double bu) {
  return 2.0 * HPLA * pow(LIGHT, 2.0) / pow(lam, 5.0) *
         (1.0 / ((bl / bu) * exp(HPLA * LIGHT / (lam * KBOL * (double)Tf)) - 1.0));
}

static double source_fn_der(double lam, float Tf, double bl, double bu) {
  const double FACTOR = 8.5685641e-6;
  double T = (double)Tf, x = exp(HPLA * LIGHT / (lam * KBOL * T));
  return (double)(float)((2.0 * FACTOR / (pow(lam, 6.0) * pow(T, 2.0))) *
                         (((bl / bu) * x) / pow((bl / bu) * x - 1.0, 2.0)));
}

/* ------------------------------------------------ GAIN1D from HSRA (normalization.f90:696-754) */
int fzc_gain(float* gain) {
  init_constants();
  int saved = RESP;
  RESP = 0;
  double* kcs = (double*)calloc((size_t)FZ_HSRA_LEN * NRG, sizeof(double));
  double* syn = (double*)calloc(NUMW, sizeof(double));
  double ne = 0.0;
  saha_t* s = (saha_t*)malloc(sizeof(saha_t));
  derivvar_t dv = {0};
  /* SOURCE_FUNCTION_HSRA's region test never fails (:679-682): all pixels use
   * the last region's central wavelength */
  double l0last = LN[RG[NRG - 1].line[0]].lambda0;
  for (int k = 0; k < FZ_HSRA_LEN; ++k) {
    float T = fz_hsra_t[k], P = fz_hsra_pg[k], D = fz_hsra_rho[k], MU = 0, PEL = 0;
    double nh;
    int e = eos_height(&T, &P, &D, &MU, &PEL, &nh, &ne, &dv, s);
    if (e) {
      RESP = saved;
      free(kcs); free(syn); free(s);
      return e;
    }
    for (int r = 0; r < NRG; ++r) opac_cont((double)T, LN[RG[r].line[0]].lambda0, ne, nh, &kcs[k * NRG + r], &dv);
    for (int r = 0; r < NRG; ++r)
      for (int l = PIX0[r]; l < PIX1[r]; ++l) {
        double lam = (l0last + WAVE[l] / 1.0e3) * 1.0e-8;
        double sf = (double)(float)(2.0 * HPLA * pow(LIGHT, 2.0) / pow(lam, 5.0) *
                                    (1.0 / (exp(HPLA * LIGHT / (lam * KBOL * (double)fz_hsra_t[k])) - 1.0)));
        if (k == 0) {
          syn[l] = sf;
        } else {
          double eta = kcs[k * NRG + r];
          int imat = (eta == eta + 1.0);
          if (imat) eta = 1.0e15;
          double dz = (k == FZ_HSRA_LEN - 1) ? 1.0e5 * fabs((double)fz_hsra_z[k] - (double)fz_hsra_z[k - 1])
                                             : 1.0e5 * fabs((double)fz_hsra_z[k + 1] - (double)fz_hsra_z[k]);
          double ev = imat ? 1.0 : exp(-eta * dz) * (0.5 * (1.0 + 1.0));
          syn[l] = (double)(float)((1.0 - ev) * sf + ev * syn[l]);
        }
      }
  }
  for (int l = 0; l < NUMW; ++l) {
    float g = (float)(1.0 / syn[l]);
    gain[4 * l + 0] = g;
    gain[4 * l + 1] = g;
    gain[4 * l + 2] = g;
    gain[4 * l + 3] = -g;
  }
  RESP = saved;
  free(kcs);
  free(syn);
  free(s);
  return E_OK;
}

/* ------------------------------------------------ one column (forward.f90:857-1029) */
int fzc_forward1d(int nz, float* model, double* syn_out, double* dsyn_out, double* kc5_out) {
  if (POPL_A && (POP_NZ != nz || POP_NREG != NRG)) return E_CONFIG;
  float *TEM = model, *PG = model + nz, *RHO = model + 2 * nz, *PEL = model + 3 * nz, *MW = model + 4 * nz;
  float *BX = model + 5 * nz, *BY = model + 6 * nz, *BZ = model + 7 * nz, *VZ = model + 10 * nz;
  float* TAU5 = model + 11 * nz;
  int err = E_OK;
  double* zz = (double*)malloc(sizeof(double) * nz);
  for (int k = 0; k < nz; ++k) zz[k] = (double)model[12 * nz + k];
  double* kc5 = (double*)malloc(sizeof(double) * nz);
  double* syn = (double*)calloc(4 * NUMW, sizeof(double));
  double* dsyn = RESP ? (double*)calloc((size_t)4 * 8 * NUMW * nz, sizeof(double)) : NULL;
  double* evg = RESP ? (double*)calloc((size_t)16 * NUMW * nz, sizeof(double)) : NULL;
  absmat_t A;
  A.eta = (double*)calloc(7 * NUMW, sizeof(double));
  A.deta = (double*)calloc((size_t)NPD * 7 * NUMW, sizeof(double));
  A.teta = (double*)calloc(7 * NUMW, sizeof(double));
  A.tdeta = (double*)calloc((size_t)NPD * 7 * NUMW, sizeof(double));
  saha_t* s = (saha_t*)malloc(sizeof(saha_t));
  const double THR = 2.0e10;
  double ne = 0.0;
  /* HYDROSTATIC=yes: RK4 from the top first (forward.f90:891-906) */
  double* vnh = HYDRO ? (double*)malloc(sizeof(double) * nz) : NULL;
  double* vne = HYDRO ? (double*)malloc(sizeof(double) * nz) : NULL;
  double* hek = (HYDRO && RESP) ? (double*)calloc(nz, sizeof(double)) : NULL;
  if (HYDRO) {
    err = fzc_he(nz, TEM, PG, RHO, PEL, MW, vnh, zz);
    for (int k = 0; k < nz; ++k) vne[k] = (double)PEL[k] / KBOL / (double)TEM[k];
  }
  for (int k = 0; k < nz && !err; ++k) {
    double nh;
    derivvar_t dv;
    memset(&dv, 0, sizeof(dv));
    if (!HYDRO) {
      err = eos_height(&TEM[k], &PG[k], &RHO[k], &MW[k], &PEL[k], &nh, &ne, &dv, s);
    } else {
      nh = vnh[k];
      ne = vne[k];
      if (RESP) eos_derivs_only(TEM[k], PG[k], RHO[k], ne, &dv, s);
    }
    if (err) break;
    err = opac_cont((double)TEM[k], 5000.0, ne, nh, &kc5[k], &dv);
    kc5[k] = kc5[k] / nh * (1.0 / MAMU);
    if (kc5[k] < THR) {
      for (int r = 0; r < NRG && !err; ++r) {
        double kc;
        derivvar_t dvr = dv;
        err = opac_cont((double)TEM[k], LN[RG[r].line[0]].lambda0, ne, nh, &kc, &dvr);
        for (int b = 0; b < RG[r].nblend && !err; ++b) {
          double kl;
          derivvar_t dvb = dvr;
          err = opac_line(popl_at(r, k), popu_at(r, k), RG[r].line[b], TEM[k], ne, nh, &kl, &dvb, s);
          if (err) break;
          double damp = damping(RG[r].line[b], TEM[k], nh, &dvb);
          err = abs_mat_block(k, r, b, TEM[k], &BX[k], &BY[k], &BZ[k], &VZ[k], kc, kl, damp, &dvb, &A);
        }
      }
      if (err) break;
    }
    /* KLINTAU (forward.f90:966-968): ETAI / N_H / m_u / MW where the
     * absorption matrix was formed at this height, 0 where it is opaque */
    if (TLOUT)
      for (int l = 0; l < NUMW; ++l)
        TLOUT[(size_t)k * NUMW + l] = (kc5[k] < THR) ? ETA(A.eta, 0, l) / nh * (1.0 / MAMU) / (double)MW[k] : 0.0;
    for (int l = 0; l < NUMW; ++l) {
      double lam = region_lambda_cm(l);
      const int rg = region_of(l);
      double sf = source_fn(lam, TEM[k], popl_at(rg, k), popu_at(rg, k));
      double* I = &syn[4 * l];
      if (k == 0 || kc5[k] >= THR) { /* SET_BOUNDARY (rtesolver.f90:1134-1147) */
        I[0] = sf;
        I[1] = I[2] = I[3] = 0.0;
        if (RESP)
          for (int p = 0; p < 8; ++p)
            for (int q = 0; q < 4; ++q) dsyn[(((size_t)k * NUMW + l) * 8 + p) * 4 + q] = 0.0;
        continue;
      }
      /* OLD_ANALYTICAL_SOLVER (rtesolver.f90:163-265) */
      double x7[7], dx[NPD * 7];
      for (int c = 0; c < 7; ++c) x7[c] = ETA(A.eta, c, l);
      for (int p = 0; p < NPD; ++p)
        for (int c = 0; c < 7; ++c) dx[p * 7 + c] = DETA(A.deta, p, c, l);
      evol_t E;
      err = evol_operator(k, nz, zz, x7, dx, &E);
      if (err) break;
      double in[4] = {I[0], I[1], I[2], I[3]};
      for (int i = 0; i < 4; ++i) {
        double a = 0.0, b = 0.0;
        for (int j = 0; j < 4; ++j) {
          a += ((i == j) - E.evol[i][j]) * (j == 0 ? sf : 0.0);
          b += E.evol[i][j] * in[j];
        }
        I[i] = a + b;
      }
      if (!RESP) continue;
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) evg[(((size_t)k * NUMW + l) * 4 + i) * 4 + j] = E.evol[i][j];
      double dsf = source_fn_der(lam, TEM[k], popl_at(rg, k), popu_at(rg, k));
      /* map evolution-operator derivatives to the DSYN1D parameter slots:
       * 1 T|Pg, 2 Pg|T, 3 rho|T, 4 Bx, 5 By, 6 Bz, 7 vlos, 8 P0 */
      const int slot_src[7] = {P_TPG, P_PGT, P_RHOT, P_BX, P_BY, P_BZ, P_VLOS};
      for (int p = 0; p < 7; ++p) {
        m44* D = &E.dev[slot_src[p]];
        double* o = &dsyn[(((size_t)k * NUMW + l) * 8 + p) * 4];
        /* With HYDROSTATIC(DER) the reference computes rho|T, Bx, By, Bz and
         * v after its hydrostatic block, which leaves STKIN = SYN1D(:,L), the
         * layer's outgoing vector (rtesolver.f90:218-258): reproduced. */
        const double* iv = (HYDRO && p >= 2) ? I : in;
        for (int i = 0; i < 4; ++i) {
          double a = 0.0, b = 0.0, c = 0.0;
          for (int j = 0; j < 4; ++j) {
            a += -(*D)[i][j] * (j == 0 ? sf : 0.0);
            b += (*D)[i][j] * iv[j];
          }
          if (p == 0) c = ((i == 0) - E.evol[i][0]) * dsf;
          o[i] = (p == 0) ? (a + c) + b : a + b;
        }
      }
      double* o8 = &dsyn[(((size_t)k * NUMW + l) * 8 + 7) * 4];
      for (int i = 0; i < 4; ++i) o8[i] = 1.0;
      if (HYDRO) {
        /* GET_HYD_DERIVATIVES (forward.f90:1033-1061): dPg_i/dT_k for i < k,
         * with the REAL(4) ratio PG(I)/PG(K); then the global term of
         * rtesolver.f90:218-231: the local Pg|T responses of the layers
         * below, carried up through their evolution operators (a layer with
         * O_11 <= 1e-5 is skipped altogether). */
        const double cte = MAMU / KBOL * GRAV * (zz[k] - zz[k - 1]) * 1.0e3 * 1.0e2;
        for (int i = 0; i < k; ++i) {
          float ratio = PG[i] / PG[k];
          float pm = -PG[k] * MW[k];
          hek[i] = (double)ratio * ((double)pm * cte / (double)TEM[k] / (double)TEM[k]);
        }
        hek[k] = 0.0;
        double st[4] = {0.0, 0.0, 0.0, 0.0};
        for (int kb = 0; kb <= k; ++kb) {
          const double* ev = (kb == k) ? &E.evol[0][0] : &evg[(((size_t)kb * NUMW + l) * 4) * 4];
          if (ev[0] > 1.0e-5) {
            const double* d2 = &dsyn[(((size_t)kb * NUMW + l) * 8 + 1) * 4];
            double r[4];
            for (int i = 0; i < 4; ++i) {
              double a = 0.0;
              for (int q = 0; q < 4; ++q) a += ev[i * 4 + q] * st[q];
              r[i] = d2[i] * hek[kb] + a;
            }
            memcpy(st, r, sizeof(st));
          }
        }
        double* o1 = &dsyn[(((size_t)k * NUMW + l) * 8 + 0) * 4];
        for (int i = 0; i < 4; ++i) o1[i] = o1[i] + st[i];
      }
    }
    /* FULL_STOKES (forward.f90:980-983): the Stokes vector after height k,
     * REAL(SYN1D) * GAIN1D */
    if (FSOUT)
      for (int l = 0; l < NUMW; ++l)
        for (int q = 0; q < 4; ++q)
          FSOUT[((size_t)k * 4 + q) * NUMW + l] = (double)((float)syn[4 * l + q] * GAIN[4 * l + q]);
  }
  if (!err && TLOUT) {
    /* TAULIN (forward.f90:1008-1015): REAL(4) trapezoid from the top, log10 */
    for (int l = 0; l < NUMW; ++l) {
      float tau = 1.0e-9f;
      float kln = (float)TLOUT[(size_t)(nz - 1) * NUMW + l];
      TLOUT[(size_t)(nz - 1) * NUMW + l] = (double)log10f(tau);
      for (int k = nz - 2; k >= 0; --k) {
        const float klk = (float)TLOUT[(size_t)k * NUMW + l];
        const float x = (klk * RHO[k] + kln * RHO[k + 1]) / 2.0f;
        tau = (float)((double)tau + (double)x * ((zz[k + 1] - zz[k]) * 1000.0 * 100.0));
        TLOUT[(size_t)k * NUMW + l] = (double)log10f(tau);
        kln = klk;
      }
    }
  }
  if (!err) {
    for (int k = 0; k < nz; ++k) kc5_out[k] = kc5[k];
    if (RESP) { /* OLD_DSYN_PROPAGATION (forward.f90:511-548) */
      for (int l = 0; l < NUMW; ++l) {
        m44 acc;
        mzero(acc);
        for (int i = 0; i < 4; ++i) acc[i][i] = 1.0;
        for (int k = nz - 2; k >= 0; --k) {
          const double* ev = &evg[(((size_t)(k + 1) * NUMW + l) * 4) * 4];
          m44 t;
          for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
              double s_ = 0.0;
              for (int q = 0; q < 4; ++q) s_ += acc[i][q] * ev[q * 4 + j];
              t[i][j] = s_;
            }
          memcpy(acc, t, sizeof(m44));
          for (int p = 0; p < 8; ++p) {
            int on = (p == 1) ? (ATMPAR[1] || ATMPAR[7]) : ATMPAR[p];
            if (!on) continue;
            double* d = &dsyn[(((size_t)k * NUMW + l) * 8 + p) * 4];
            double r[4];
            for (int i = 0; i < 4; ++i) {
              double s_ = 0.0;
              for (int q = 0; q < 4; ++q) s_ += acc[i][q] * d[q];
              r[i] = s_;
            }
            memcpy(d, r, sizeof(r));
          }
          if (ATMPAR[7]) { /* P0 (forward.f90:543-545) */
            double* d8 = &dsyn[(((size_t)k * NUMW + l) * 8 + 7) * 4];
            const double* d2 = &dsyn[(((size_t)k * NUMW + l) * 8 + 1) * 4];
            for (int i = 0; i < 4; ++i) d8[i] = d8[i] * (double)PG[k] / (double)PG[nz - 1] * d2[i];
          }
        }
      }
      for (int k = 0; k < nz; ++k)
        for (int l = 0; l < NUMW; ++l)
          for (int p = 0; p < 8; ++p)
            for (int q = 0; q < 4; ++q) {
              size_t i = (((size_t)k * NUMW + l) * 8 + p) * 4 + q;
              dsyn_out[i] = dsyn[i] * (double)GAIN[4 * l + q];
            }
      if (LSF_ON)
        for (int k = 0; k < nz; ++k)
          for (int p = 0; p < 8; ++p)
            for (int q = 0; q < 4; ++q) lsf_row(&dsyn_out[((size_t)k * NUMW * 8 + p) * 4 + q], 32);
    }
    for (int l = 0; l < NUMW; ++l)
      for (int q = 0; q < 4; ++q) syn_out[4 * l + q] = syn[4 * l + q] * (double)GAIN[4 * l + q];
    if (LSF_ON)
      for (int q = 0; q < 4; ++q) lsf_row(&syn_out[q], 4);
    /* log tau_5000, REAL(4) trapezoid (forward.f90:1000-1006) */
    TAU5[nz - 1] = 1e-9f;
    for (int k = nz - 2; k >= 0; --k) {
      float a = ((float)kc5[k] * RHO[k] + (float)kc5[k + 1] * RHO[k + 1]) / 2.0f;
      TAU5[k] = (float)((double)TAU5[k + 1] + (double)a * ((zz[k + 1] - zz[k]) * 1000.0 * 100.0));
    }
    for (int k = 0; k < nz; ++k) TAU5[k] = log10f(TAU5[k]);
  }
  free(vnh); free(vne); free(hek);
  free(zz); free(kc5); free(syn); free(dsyn); free(evg);
  free(A.eta); free(A.deta); free(A.teta); free(A.tdeta); free(s);
  return err;
}

int fzc_forward_batch(int ncol, int nz, float* models, double* syn, double* dsyn, int nthreads) {
  int err = 0;
  size_t sz_syn = (size_t)4 * NUMW, sz_dsyn = (size_t)4 * 8 * NUMW * nz;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
  for (int c = 0; c < ncol; ++c) {
    double* kc5 = (double*)malloc(sizeof(double) * nz);
    int e = fzc_forward1d(nz, models + (size_t)c * 13 * nz, syn + c * sz_syn, dsyn ? dsyn + c * sz_dsyn : NULL, kc5);
    free(kc5);
    if (e) {
#pragma omp atomic write
      err = e;
    }
  }
  return err;
}

/* ------------------------------------------------ hydrostatic RK4 (hydrostatic_tools.f90:53-71,150-222,350-396) */
static double rk4_rhs(double Pln, double tem, double* nel, int* err, saha_t* s) {
  double nh, pel, dens, mw;
  saha_init(s, tem);
  int e = get_rho(s, tem, exp(Pln), &nh, nel, &pel, &dens, &mw);
  if (e) *err = e;
  return -mw * MAMU * GRAV / (KBOL * tem);
}

int fzc_he(int nz, float* tem, float* pg, float* rho, float* pel, float* mw, double* vnh, const double* zz) {
  init_constants();
  saha_t* s = (saha_t*)malloc(sizeof(saha_t));
  double *pg8 = (double*)malloc(sizeof(double) * nz), *rho8 = (double*)malloc(sizeof(double) * nz);
  double *pel8 = (double*)malloc(sizeof(double) * nz), *mw8 = (double*)malloc(sizeof(double) * nz);
  int err = 0;
  int top = nz - 1;
  pg8[top] = log((double)pg[top]);
  double nel = (double)pel[top] / KBOL / (double)tem[top];
  double t0 = zz[top] * 1000.0 * 100.0, u0 = pg8[top];
  double nh, pe, dens, mol;
  saha_init(s, (double)tem[top]);
  err = get_rho(s, (double)tem[top], exp(u0), &nh, &nel, &pe, &dens, &mol);
  rho8[top] = dens;
  pel8[top] = pe;
  mw8[top] = mol;
  vnh[top] = nh;
  for (int k = nz - 2; k >= 0 && !err; --k) {
    double dz = (zz[k] - zz[k + 1]) * 1000.0 * 100.0;
    double tm0 = (double)tem[k + 1], tm1 = (double)tem[k];
    /* RK4_LOG (:350-396) */
    double k1 = rk4_rhs(u0, tm0, &nel, &err, s) * dz;
    double it1 = t0 + dz / 2.0;
    double tmid = tm0 / dz * (t0 + dz - it1) + tm1 / dz * (it1 - t0);
    double k2 = rk4_rhs(u0 + k1 / 2.0, tmid, &nel, &err, s) * dz;
    double k3 = rk4_rhs(u0 + k2 / 2.0, tmid, &nel, &err, s) * dz;
    double k4 = rk4_rhs(u0 + k3, tm1, &nel, &err, s) * dz;
    double u1 = u0 + 1.0 / 6.0 * (k1 + 2.0 * k2 + 2.0 * k3 + k4);
    saha_init(s, tm1);
    int e = get_rho(s, tm1, exp(u1), &nh, &nel, &pe, &dens, &mol);
    if (e) err = e;
    rho8[k] = dens;
    pel8[k] = pe;
    mw8[k] = mol;
    pg8[k] = u1;
    vnh[k] = nh;
    u0 = u1;
    t0 = t0 + dz;
  }
  for (int k = 0; k < nz; ++k) {
    pg[k] = (float)exp(pg8[k]);
    rho[k] = (float)rho8[k];
    pel[k] = (float)pel8[k];
    mw[k] = (float)mw8[k];
  }
  free(s); free(pg8); free(rho8); free(pel8); free(mw8);
  return err;
}
