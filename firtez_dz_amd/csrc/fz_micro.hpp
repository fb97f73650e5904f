// FP64 device microphysics for one (column, height): equation of state with
// its implicit derivatives, continuum and line opacity, collisional damping.
//
// Reference routines (file:line) are cited per function.  The arithmetic is
// FP64 throughout: the electron-density fixed point stops at a relative change
// of 1e-4 (chemical_equilibrium.f90:214), so its result depends on the
// iteration path and is reproduced iteration by iteration, with the same warm
// start (the previous height's n_e) as the reference column loop.
#pragma once
#include <hip/hip_runtime.h>

#define FZ_DATA_QUAL static __constant__
#include "fz_atomdata.h"

namespace fz {

constexpr double DPI = 3.141592653589793238462643;
constexpr double LIGHT = 2.99792458e10;
constexpr double KBOL = 1.38056503e-16;
constexpr double MELE = 9.10938215e-28;
constexpr double QELE = 4.80320425e-10;
constexpr double EVOLT = 1.60217656e-12;
constexpr double MAMU = 1.66053892e-24;
constexpr double HPLA = 6.62606957e-27;
constexpr double GRAV = 2.7414e4;

// derived constants, filled once by fz_init_constants_kernel
struct Consts {
  double csaha1, csaha2, rbohr, deno_sa, deno_sam;
  double abfac[FZ_NELEM];   // 10**(ABUND-12)
};
__device__ Consts g_c;

__global__ void fz_init_constants_kernel() {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  g_c.csaha1 = pow(HPLA, 3.0) / pow(2.0 * DPI * MELE * KBOL, 1.5);   // cons_param.f90:25
  g_c.csaha2 = EVOLT / KBOL;
  g_c.rbohr = pow(HPLA, 2.0) / (4.0 * pow(DPI, 2.0) * MELE * pow(QELE, 2.0));
  double sa = 0.0, sam = 0.0;
  for (int i = 0; i < FZ_NELEM; ++i) {          // atom_database.f90:112-118
    double f = pow(10.0, fz_abund[i] - 12.0);
    g_c.abfac[i] = f;
    sa += f;
    sam += MAMU * fz_matom[i] * f;
  }
  g_c.deno_sa = sa;
  g_c.deno_sam = sam;
}

// Partition functions U_I,II,III(T) and dlnU/dT: runs the element's predicated
// program from fz_atomdata.h (data taken from atom_database.f90:121-1036).
// All lanes of a wave evaluate the same element, so the opcode stream is
// wave-uniform and the dispatch is a scalar branch.
__device__ __forceinline__ void partition(int elem, double T, double u[3], double du[3]) {
  double v[10];
  double st[16];
  int sp = 0;
  v[FZ_PV_T] = T;
  v[FZ_PV_X] = log(5040.0 / T);
  v[FZ_PV_Y] = 1.0e-3 * T;
  v[FZ_PV_DX] = -1.0 / T;
#pragma unroll
  for (int i = 4; i < 10; ++i) v[i] = 0.0;
  const int pc0 = fz_pf_start[elem], pc1 = fz_pf_start[elem + 1];
  for (int pc = pc0; pc < pc1; ++pc) {
    const int op = fz_pf_op[pc], a = fz_pf_arg[pc];
    if (op == FZ_PF_CONST) { st[sp++] = fz_pf_const[a]; continue; }
    if (op == FZ_PF_VAR) { st[sp++] = v[a]; continue; }
    if (op == FZ_PF_STORE) { v[a] = st[--sp]; continue; }
    if (op == FZ_PF_NEG) { st[sp - 1] = -st[sp - 1]; continue; }
    if (op == FZ_PF_EXP) { st[sp - 1] = exp(st[sp - 1]); continue; }
    if (op == FZ_PF_END) continue;
    if (op == FZ_PF_SEL) {
      const double c = st[sp - 1], old = st[sp - 2], nw = st[sp - 3];
      sp -= 2;
      st[sp - 1] = (c != 0.0) ? nw : old;
      continue;
    }
    const double y = st[--sp];
    double x = st[sp - 1];
    switch (op) {
      case FZ_PF_ADD: x = x + y; break;
      case FZ_PF_SUB: x = x - y; break;
      case FZ_PF_MUL: x = x * y; break;
      case FZ_PF_DIV: x = x / y; break;
      case FZ_PF_MAX: x = x > y ? x : y; break;
      case FZ_PF_AMAX1: { float fx = (float)x, fy = (float)y; x = (double)(fx > fy ? fx : fy); break; }
      case FZ_PF_GT: x = (x > y) ? 1.0 : 0.0; break;
      case FZ_PF_LT: x = (x < y) ? 1.0 : 0.0; break;
      case FZ_PF_LE: x = (x <= y) ? 1.0 : 0.0; break;
      case FZ_PF_GE: x = (x >= y) ? 1.0 : 0.0; break;
      case FZ_PF_NE: x = (x != y) ? 1.0 : 0.0; break;
      case FZ_PF_EQ: x = (x == y) ? 1.0 : 0.0; break;
      case FZ_PF_AND: x = (x != 0.0 && y != 0.0) ? 1.0 : 0.0; break;
      default: break;
    }
    st[sp - 1] = x;
  }
  u[0] = v[FZ_PV_UI]; u[1] = v[FZ_PV_UII]; u[2] = v[FZ_PV_UIII];
  du[0] = v[FZ_PV_DUI]; du[1] = v[FZ_PV_DUII]; du[2] = v[FZ_PV_DUIII];
}

// Saha factors of every element at one temperature (the quantities each
// ION_CALC_* routine forms, chemical_equilibrium.f90:1601-1652).  Held in the
// thread's private segment: lanes are interleaved there, so the per-element
// loops below read them coalesced.
struct Saha {
  double a1[FZ_NELEM], a2[FZ_NELEM], da1[FZ_NELEM], da2[FZ_NELEM];
};

// The same four per-element rows read from a device buffer laid out
// [height][element][4][column] (columns fastest), as the parallel Saha
// kernel writes them: row q of element j at p[(j*4 + q) * stride].
struct SahaRow {
  const double* p;
  long st;
  __device__ __forceinline__ double operator[](int j) const { return p[(long)j * st]; }
};
struct SahaView {
  SahaRow a1, a2, da1, da2;
  __device__ __forceinline__ SahaView(const double* base, long stride)
      : a1{base, 4 * stride}, a2{base + stride, 4 * stride}, da1{base + 2 * stride, 4 * stride},
        da2{base + 3 * stride, 4 * stride} {}
};

// Saha factors of one element at temperature T (the loop body of saha_init)
__device__ __forceinline__ void saha_element(int i, double T, double lam, double dlam, bool der, double& a1,
                                             double& a2, double& da1, double& da2) {
  double u[3], du[3];
  partition(i, T, u, du);
  double e1 = g_c.csaha2 * fz_xi[i] / T, e2 = g_c.csaha2 * fz_xii[i] / T;
  e1 = fmax(fmin(e1, 650.0), -650.0);
  e2 = fmax(fmin(e2, 650.0), -650.0);
  a1 = (u[0] / (2.0 * u[1])) * lam * exp(e1);
  a2 = (u[1] / (2.0 * u[2])) * lam * exp(e2);
  da1 = da2 = 0.0;
  if (der) {
    const double T2 = T * T;
    da1 = a1 * (du[0] - du[1] + dlam + (-g_c.csaha2 * fz_xi[i] / T2));
    da2 = a2 * (du[1] - du[2] + dlam + (-g_c.csaha2 * fz_xii[i] / T2));
  }
}

__device__ inline void saha_init(Saha& s, double T, bool der) {
  const double lam = g_c.csaha1 * pow(T, -1.5);
  const double dlam = -3.0 / (2.0 * T);
  for (int i = 0; i < FZ_NELEM; ++i) {
    double a1, a2, da1, da2;
    saha_element(i, T, lam, dlam, der, a1, a2, da1, da2);
    s.a1[i] = a1;
    s.a2[i] = a2;
    if (der) {
      s.da1[i] = da1;
      s.da2[i] = da2;
    }
  }
}

__device__ __forceinline__ void stages(double N, double ne, double a1, double a2,
                                       double& ni, double& nii, double& niii) {
  ni = N / (1.0 + (1.0 / (ne * a1)) * (1.0 + (1.0 / (ne * a2))));
  nii = N / (ne * a1 + 1.0 + (1.0 / (ne * a2)));
  niii = N / (1.0 + ne * a2 * (ne * a1 + 1.0));
}

// DERIVVAR scalars (derivvar.f90:16-24) needed downstream
struct Deriv {
  double nh_T_pg, nh_T_rho, nh_pg_T, nh_rho_T;
  double ne_pg_T, ne_T_rho, ne_T_pg, ne_rho_T;
};

enum : int { ERR_OK = 0, ERR_EOS = 1, ERR_NAN = 2, ERR_ION = 3, ERR_LAMBDA = 4 };

// GET_RHO (chemical_equilibrium.f90:165-278).  ne: warm start in, result out.
// pelec is the starting guess's electron pressure, as the reference returns.
template <class S>
__device__ inline int get_rho(const S& s, double T, double PG, double& ne, double& nh,
                              double& pelec, double& dens, double& mu) {
  const double kT = KBOL * T;
  double n = ne;
  if (n != n) n = 0.0;
  if (n < 1.0) n = PG / (101.0 * kT);
  pelec = n * kT;
  if (pelec > PG) n = (PG * 0.5) / kT;
  double err = 1.0, nnew = n, h = 0.0;
  int iter = 1;
  while (err > 1.0e-4 && iter <= 100 && h >= 0.0) {
    const double nold = nnew;
    h = (PG - nnew * kT) / (g_c.deno_sa * kT);
    double ni, nii, niii;
    stages(h * g_c.abfac[0], nnew, s.a1[0], s.a2[0], ni, nii, niii);
    double est = niii - ni, s2 = 0.0, s3 = 0.0;
    for (int j = 1; j < FZ_NELEM; ++j) {
      stages(h * g_c.abfac[j], nnew, s.a1[j], s.a2[j], ni, nii, niii);
      s2 += nii;
      s3 += niii;
    }
    est = est + s2 + s3;
    nnew = (fmax(est, 0.0) + nold) / 2.0;
    err = fabs((nnew - nold) / nnew);
    ++iter;
  }
  if (iter >= 100 || h < 0.0) return ERR_EOS;
  n = nnew;
  h = (PG - n * kT) / (g_c.deno_sa * kT);
  dens = h * g_c.deno_sam + n * MELE;
  mu = ((dens * kT) / MAMU) * (1.0 / PG);
  if (h != h || n != n) return ERR_NAN;
  if (n < 0.0 || h <= 0.0) return ERR_EOS;
  ne = n;
  nh = h;
  return ERR_OK;
}

// The four implicit-derivative solves of GET_DNHYDDTEMP
// (chemical_equilibrium.f90:2008-2052).
template <class S>
__device__ inline void eos_derivatives(const S& s, double T, double P, double RHO, double NE,
                                       Deriv& d) {
  const double kT = KBOL * T;
  const double SA = g_c.deno_sa, SAM = g_c.deno_sam;
  // d n_e / d Pg at constant T: GET_NE_TEMP_SNE_ALT (:892-957) + ION_CALC_NETEMP_NEW (:1376-1442)
  {
    const double NH = (P - NE * kT) / (kT * SA);
    double sn = 0.0, sd = 0.0;
    for (int j = 0; j < FZ_NELEM; ++j) {
      const double N = NH * g_c.abfac[j], a1 = s.a1[j], a2 = s.a2[j];
      double ni, nii, niii;
      stages(N, NE, a1, a2, ni, nii, niii);
      const double den = ((a1 * NE + 1.0) * a2) * NE + 1.0;
      const double q = 2.0 * a1 * a2 * NE + a2;
      const double f = g_c.abfac[j] / SA;
      const double n1 = ni / N * f / KBOL / T, n2 = nii / N * f / KBOL / T, n3 = niii / N * f / KBOL / T;
      const double d1 = -ni / N * f + N * (2.0 * a1 * a2 * NE) / den - ni * q / den;
      const double d2 = -nii / N * f + N * a2 / den - nii * q / den;
      const double d3 = -niii / N * f - niii * q / den;
      if (j > 0) { sn += n2 + n3; sd += d2 + d3; }
      else { sn += -n1 + n3; sd += -d1 + d3; }
    }
    d.ne_pg_T = sn / (1.0 - sd);
    d.nh_pg_T = 1.0 / SA * (1.0 / KBOL / T - d.ne_pg_T);
  }
  // d n_e / d T at constant rho: GET_NE_RHO_NIT (:509-570) + ION_CALC_TEMPRHO_NIT (:1876-1970)
  {
    const double NH = (RHO - NE * MELE) / SAM;
    double sn = 0.0, sd = 0.0;
    for (int j = 0; j < FZ_NELEM; ++j) {
      const double N = NH * g_c.abfac[j], a1 = s.a1[j], a2 = s.a2[j], da1 = s.da1[j], da2 = s.da2[j];
      double ni, nii, niii;
      stages(N, NE, a1, a2, ni, nii, niii);
      const double dn = -MELE / SAM * g_c.abfac[j];
      const double den = ((a1 * NE + 1.0) * a2) * NE + 1.0;
      const double r = (da1 * a2 * NE + (a1 * NE + 1.0) * da2) * NE / den;
      const double q = 2.0 * a1 * a2 * NE + a2;
      const double n1 = N * NE * NE * (da1 * a2 + a1 * da2) / den - ni * r;
      const double n2 = N * NE * da2 / den - nii * r;
      const double n3 = -niii * r;
      const double d1 = dn * ni / N + N * (2.0 * a1 * a2 * NE) / den - ni * q / den;
      const double d2 = dn * nii / N + N * a2 / den - nii * q / den;
      const double d3 = dn * niii / N - niii * q / den;
      if (j > 0) { sn += n2 + n3; sd += d2 + d3; }
      else { sn += -n1 + n3; sd += -d1 + d3; }
    }
    d.ne_T_rho = sn / (1.0 - sd);
  }
  // d n_e / d rho at constant T: GET_NE_DNEDRHO_TEMP_NIT (:1120-1185) + ION_CALC_NETEMP (:1316-1372)
  {
    const double NH = (RHO - NE * MELE) / SAM;
    double num = 0.0, dsum = 0.0;
    for (int j = 0; j < FZ_NELEM; ++j) {
      const double NT = NH * g_c.abfac[j], a1 = s.a1[j], a2 = s.a2[j];
      const double den = NE * NE * a1 * a2 + NE * a2 + 1.0;
      const double fi = NE * NE * a1 * a2 / den, fii = NE * a2 / den, fiii = 1.0 / den;
      const double q = NE * 2.0 * a1 * a2 + a2;
      const double dfi = fi / NE * 2.0 - fi * q / den, dfii = a2 / den - fii * q / den, dfiii = -fiii * q / den;
      if (j == 0) { num += NT * (-dfi + dfiii); dsum += g_c.abfac[j] * (-fi + fiii); }
      else { num += NT * (dfii + dfiii); dsum += g_c.abfac[j] * (fii + fiii); }
    }
    const double dsm = (1.0 - num) / dsum;
    d.ne_rho_T = 1.0 / (dsm * SAM + MELE);
    d.nh_rho_T = (1.0 - d.ne_rho_T * MELE) / SAM;
  }
  // d n_e / d T at constant Pg: GET_DNE_PG (:2062-2158) with ION_CALC_TEMPPGAS (:1791-1872).
  // The summand is affine in the trial derivative x, est(x) = A + B x, so the
  // damped iteration x <- (est(x) + x)/2 is run on the scalars A, B (one pass
  // over the elements instead of one per iteration); the stopping rule and the
  // final re-evaluation are the reference's.
  {
    const double NH = (P - NE * kT) / (kT * SA);
    double A = 0.0, B = 0.0;
    for (int j = 0; j < FZ_NELEM; ++j) {
      const double N = NH * g_c.abfac[j], a1 = s.a1[j], a2 = s.a2[j], da1 = s.da1[j], da2 = s.da2[j];
      double ni, nii, niii;
      stages(N, NE, a1, a2, ni, nii, niii);
      const double f = g_c.abfac[j] / SA;
      // dn = -f*(P/(kT^2) + x) : constant part and x-coefficient
      const double dn0 = -f * (P / (KBOL * T * T)), dn1 = -f;
      const double g1 = 1.0 + (1.0 / (NE * a1)) * (1.0 + 1.0 / (NE * a2));
      const double h1 = N / (g1 * g1) * (1.0 / (NE * NE * a1));
      const double c1 = 1.0 + 1.0 / (NE * a2);
      // dni = dn*ni/N - h1*((-1/a1)*(x*a1 + NE*da1)*c1 - (1/(NE*a2^2))*(x*a2 + NE*da2))
      const double ni0 = dn0 * ni / N - h1 * ((-1.0 / a1) * (NE * da1) * c1 - (1.0 / (NE * a2 * a2)) * (NE * da2));
      const double ni1 = dn1 * ni / N - h1 * ((-1.0 / a1) * a1 * c1 - (1.0 / (NE * a2 * a2)) * a2);
      const double g2 = NE * a1 + 1.0 + 1.0 / (NE * a2);
      const double h2 = N / (g2 * g2);
      const double nii0 = dn0 * nii / N - h2 * (NE * da1 - (NE * da2) / (NE * NE * a2 * a2));
      const double nii1 = dn1 * nii / N - h2 * (a1 - a2 / (NE * NE * a2 * a2));
      const double g3 = 1.0 + NE * a2 * (NE * a1 + 1.0);
      const double h3 = N / (g3 * g3);
      const double niii0 = dn0 * niii / N - h3 * (NE * NE * da2 * a1 + NE * NE * a2 * da1 + NE * da2);
      const double niii1 = dn1 * niii / N - h3 * (2.0 * NE * a1 * a2 + a2);
      if (j > 0) { A += nii0 + niii0; B += nii1 + niii1; }
      else { A += -ni0 + niii0; B += -ni1 + niii1; }
    }
    double x = 0.0, err = 1.0e20;
    int iter = 0;
    while (err > 1.0e-6 && iter <= 100) {
      const double xo = x;
      x = ((A + B * x) + xo) / 2.0;
      if (fabs(x) > 1.0e40) x = 0.0;
      err = fabs((x - xo) / x);
      ++iter;
    }
    d.ne_T_pg = A + B * x;
  }
  d.nh_T_pg = (-1.0 / SA) * (P / (KBOL * T * T) + d.ne_T_pg);
  d.nh_T_rho = (-MELE / SAM) * d.ne_T_rho;
}

// Continuum opacity and its derivatives (cont_opacity.f90:48-357).
// kc[0] = kappa_c, kc[1] = d/dT|Pg, kc[2] = d/dPg|T, kc[3] = d/drho|T.
__device__ inline int opac_cont(double T, double lam, double NE, double NH, const Deriv& d, bool der,
                                double kc[4]) {
  if (lam < 1800.0) return ERR_LAMBDA;
  const double lm = lam / 1.0e4, th = 5040.0 / T, dth = -th / T;
  double p1, p2;
  // H- free-free, John 1988
  if (lm < 0.3645) {
    const double A[4] = {518.1021, 473.2636, -482.2089, 115.5291}, B[4] = {-734.8666, 1443.4137, -737.1616, 169.6374};
    const double C[4] = {1021.1775, -1977.3395, 1096.8827, -245.649}, D[4] = {-479.0721, 922.3575, -521.1341, 114.243};
    const double E[4] = {93.1373, -178.9275, 101.7963, -21.9972}, F[4] = {-6.4285, 12.36, -7.0571, 1.5097};
    double c[4];
    for (int i = 0; i < 4; ++i)
      c[i] = A[i] * lm * lm + B[i] + C[i] / lm + D[i] / (lm * lm) + E[i] / (lm * lm * lm) + F[i] / (lm * lm * lm * lm);
    const double sq = sqrt(th);
    p1 = c[0] * th + c[1] * th * sq + c[2] * th * th + c[3] * th * th * sq;
    p2 = (c[0] + 1.5 * c[1] * sq + 2.0 * c[2] * th + 2.5 * c[3] * th * sq) * dth;
  } else {
    const double A[6] = {0, 2483.346, -3449.889, 2200.04, -696.271, 88.283};
    const double B[6] = {0, 285.827, -1158.382, 2427.719, -1841.4, 444.517};
    const double C[6] = {0, -2054.291, 8746.523, -13651.105, 8624.97, -1863.864};
    const double D[6] = {0, 2827.776, -11485.632, 16755.524, -10051.53, 2095.288};
    const double E[6] = {0, -1341.537, 5303.609, -7510.494, 4400.067, -901.788};
    const double F[6] = {0, 208.952, -812.939, 1132.738, -655.02, 132.985};
    double c[6];
    for (int i = 0; i < 6; ++i)
      c[i] = A[i] * lm * lm + B[i] + C[i] / lm + D[i] / (lm * lm) + E[i] / (lm * lm * lm) + F[i] / (lm * lm * lm * lm);
    const double sq = sqrt(th), t2 = th * th;
    p1 = c[0] * th + c[1] * th * sq + c[2] * t2 + c[3] * t2 * sq + c[4] * t2 * th + c[5] * t2 * th * sq;
    p2 = (c[0] + 1.5 * c[1] * sq + 2.0 * c[2] * th + 2.5 * c[3] * th * sq + 3.0 * c[4] * t2 + 3.5 * c[5] * t2 * sq) * dth;
  }
  const double ff = 1.0e-29 * p1 * KBOL * NE * T;
  // H- bound-free
  const double LP = 1.6419, CTE = 0.75e-18, AL = (HPLA * LIGHT / KBOL) * 1.0e4;
  double bf = 0.0, dpart_over = 0.0;
  if (lm < LP) {
    const double com = 1.0 / lm - 1.0 / LP, sc = sqrt(com);
    const double CC[6] = {152.519, 49.534, -118.858, 92.536, -34.194, 4.982};
    double sg = CC[0] + CC[1] * sc + CC[2] * com + CC[3] * com * sc + CC[4] * com * com + CC[5] * com * com * sc;
    sg = CTE * sg * lm * lm * lm * com * sc;
    const double ex = exp(-AL / (T * lm));
    const double part = pow(T, -2.5) * exp(AL / (T * LP)) * (1.0 - ex);
    dpart_over = (1.0 / T) * (-2.5 + AL * com / T - (AL / (lm * T)) * (1.0 / (1.0 - ex)));
    bf = part * sg * NE * KBOL * T;
  }
  // Thomson
  const double CT = (8.0 * DPI / 3.0) * pow(QELE / LIGHT, 4.0) / (MELE * MELE);
  const double thm = CT * (NE / NH);
  // Rayleigh by hydrogen
  const double l2 = lam * lam;
  const double ray = (5.799e-13 + (1.422e-6 + 2.784 / l2) / l2) / (l2 * l2);
  // bound-free/bound-bound of neutral H; IOPOT and the 1E-8 factors are REAL(4) literals
  const double IOPOT = (double)13.595f, F8 = (double)1e-8f;
  const double HC1 = F8 * IOPOT * EVOLT / (HPLA * LIGHT);
  const double HC2 = F8 * F8 * F8 * 64.0 * pow(DPI, 4.0) * MELE * pow(QELE, 10.0) /
                     (3.0 * sqrt(3.0) * pow(LIGHT, 4.0) * pow(HPLA, 6.0));
  const double HC3 = HPLA * LIGHT / (KBOL * F8);
  const double HC4 = IOPOT * EVOLT / KBOL;
  const int n0 = 1 + (int)floor(sqrt(HC1 * lam));
  double sum, dsum;
  const double e81 = exp(HC4 / (T * 81.0));
  if (n0 <= 8) {
    sum = 0.0;
    dsum = 0.0;
    for (int i = n0; i <= 8; ++i) {
      const double fi = (double)i, e = exp(HC4 / (T * fi * fi));
      sum += e / (fi * fi * fi);
      dsum -= HC4 * e / (fi * fi * fi * fi * fi) / (T * T);
    }
    sum += (0.117 + e81) * (T / (2.0 * HC4));
  } else {
    const double fn = (double)n0;
    sum = (0.117 + exp(HC4 / (T * fn * fn))) * (T / (2.0 * HC4));
    dsum = 0.0;
  }
  dsum += (1.0 / (2.0 * HC4)) * (0.117 + e81) - (1.0 / (2.0 * 81.0 * T)) * e81;
  const double exh = exp(-HC4 / T), exl = exp(-HC3 / (T * lam));
  const double hyd = HC2 * sum * (1.0 - exl) * exh * lam * l2;
  const double fac = ff + bf + thm + ray + hyd;
  kc[0] = NH * fac;
  if (der) {
    const double hyd_d = hyd * (dsum / sum + HC4 / (T * T)) - (HC2 * HC3 * l2 / (T * T)) * exh * exl;
    const double ff_p = ff * (d.ne_T_pg / NE + 1.0 / T + p2 / p1);
    const double bf_p = bf * (dpart_over + 1.0 / T + d.ne_T_pg / NE);
    const double th_p = thm * (d.ne_T_pg / NE - d.nh_T_pg / NH);
    kc[1] = fac * d.nh_T_pg + NH * (ff_p + bf_p + th_p + hyd_d);
    const double ffbf_ne = (ff + bf) / NE;
    kc[2] = fac * d.nh_pg_T + NH * (ffbf_ne * d.ne_pg_T + (thm / NE * d.ne_pg_T - thm / NH * d.nh_pg_T));
    kc[3] = fac * d.nh_rho_T + NH * (ffbf_ne * d.ne_rho_T + (thm / NE * d.ne_rho_T - thm / NH * d.nh_rho_T));
  }
  return ERR_OK;
}

struct LineConst {   // per-line constants prepared on the host
  int zn, ion;       // element index (0-based), ionisation stage (1..3)
  double lambda0, loggf, eplow, sigma, alpha;
};

// Line opacity (line_opacity.f90:34-217) with ION_CALC_TEMPX (:1511-1597)
// and ION_CALC_PGTEMP (:1446-1507).  out: kl, dkl/dT|Pg, dkl/dPg|T, dkl/drho|T.
template <class S>
__device__ inline int opac_line(const LineConst& L, double T, double NE, double NH, const S& s,
                                const Deriv& d, bool der, double out[4]) {
  const int z = L.zn, ion = L.ion;
  if (ion < 1 || ion > 3) return ERR_ION;
  const double ab = g_c.abfac[z], N = NH * ab;
  const double a1 = s.a1[z], a2 = s.a2[z];
  const double den = NE * NE * a1 * a2 + NE * a2 + 1.0;
  const double nume = ion == 1 ? NE * NE * a1 * a2 : (ion == 2 ? NE * a2 : 1.0);
  const double npi = N * nume / den;
  double u[3], du[3];
  partition(z, T, u, du);
  const double upi = u[ion - 1], dupi = du[ion - 1];
  const double t1 = sqrt(DPI) * QELE * QELE / (MELE * LIGHT) * pow(10.0, L.loggf) * L.lambda0 * 1.0e-8;
  const double t2 = sqrt(fz_matom[z] * MAMU / (2.0 * KBOL * T));
  const double t3 = npi / upi;
  const double t4 = exp(-L.eplow * EVOLT / (KBOL * T));
  const double t5 = 1.0 - exp(-HPLA * LIGHT / (1.0e-8 * L.lambda0 * KBOL * T));
  out[0] = t1 * t2 * t3 * t4 * t5;
  if (!der) return ERR_OK;
  const double da1 = s.da1[z], da2 = s.da2[z];
  // d N_ion / dT at constant Pg
  const double dne = d.ne_T_pg;
  const double dden = 2.0 * NE * dne * a1 * a2 + NE * NE * da1 * a2 + NE * NE * a1 * da2 + dne * a2 + NE * da2;
  const double dnum = ion == 1 ? 2.0 * NE * dne * a1 * a2 + NE * NE * da1 * a2 + NE * NE * a1 * da2
                               : (ion == 2 ? dne * a2 + NE * da2 : 0.0);
  const double dnir_T = (nume / den) * ab * d.nh_T_pg + N * (dnum * den - dden * nume) / (den * den);
  // d N_ion / dPg and d/drho at constant T
  const double denoi = ((a1 * NE + 1.0) * a2) * NE + 1.0;
  const double fi = NE * NE * a1 * a2 / den, fii = NE * a2 / den, fiii = 1.0 / den;
  const double q = (2.0 * a1 * a2 * NE + a2) / denoi;
  const double fir = ion == 1 ? fi : (ion == 2 ? fii : fiii);
  const double gfir = ion == 1 ? 2.0 * a1 * a2 * NE / denoi : (ion == 2 ? a2 / denoi : 0.0);
  const double dnir_P = N / NH * fir * d.nh_pg_T + N * (gfir - fir * q) * d.ne_pg_T;
  const double dnir_R = N / NH * fir * d.nh_rho_T + N * (gfir - fir * q) * d.ne_rho_T;
  const double d2 = -0.5 / T * t2;
  const double d3 = (dnir_T - npi * dupi) / upi;
  const double d4 = L.eplow * EVOLT / (KBOL * T * T) * t4;
  const double d5 = -HPLA * LIGHT / (1.0e-8 * L.lambda0 * KBOL * T * T) * (1.0 - t5);
  out[1] = t1 * (d2 * t3 * t4 * t5 + t2 * d3 * t4 * t5 + t2 * t3 * d4 * t5 + t2 * t3 * t4 * d5);
  out[2] = out[0] / npi * dnir_P;
  out[3] = out[0] / npi * dnir_R;
  return ERR_OK;
}

// Air refraction index (atom_database.f90:1039-1070)
__device__ inline double refrax(double w) {
  const double Tc = 15.0, P = 760.0;
  const double s = 1.0 / (w * w);
  const double r = 6.4328e-5 + 2.94981e-2 / (146.0 - s) + 2.554e-4 / (41.0 - s);
  const double al = 3.67e-3 + 3.3e-4 * exp((0.19 - w) / 0.12);
  return 1.0 + 1.31579e-3 * r * P / (1.0 + al * (Tc - 1.0) / (1.0 + 15.0 * al));
}

// Reduced damping and derivatives (damping.f90:842-1012).
// out: a, da/dT|Pg, da/dPg|T, da/drho|T.
__device__ inline void damping(const LineConst& L, double T, double NH, const Deriv& d, bool der,
                               double out[4]) {
  const int z = L.zn;
  const double lam = L.lambda0 * 1e-8;
  const double vd = sqrt(2.0 * KBOL * T / (MAMU * fz_matom[z]));
  const double grad = 0.2223 / (lam * lam);
  const double muh = MAMU * (fz_matom[z] * fz_matom[0]) / (fz_matom[z] + fz_matom[0]);
  const double muhe = MAMU * (fz_matom[z] * fz_matom[1]) / (fz_matom[z] + fz_matom[1]);
  const double heab = g_c.abfac[1];
  const double HEPOL = (double)0.3181818f;
  double gcol, dgp = 0.0;
  if (L.sigma > 0 && L.alpha > 0) {
    const double al = L.alpha, e1 = 2.0 - al / 2.0, e2 = (1.0 - al) / 2.0;
    const double gf = 1.0 + (-0.5748646 + (0.9512363 + (-0.6998588 + (0.4245549 - 0.1010678 * e1) * e1) * e1) * e1) * e1;
    const double beta = 2.0 * pow(4.0 / DPI, al / 2.0) * gf * pow(1.0e6, al) * (L.sigma * g_c.rbohr * g_c.rbohr) *
                        pow(8.0 * KBOL / DPI, e2);
    const double tp = pow(T, e2), mh = pow(muh, -e2), mhe = pow(muhe, -e2);
    gcol = beta * tp * (NH * mh + HEPOL * NH * heab * mhe);
    if (der) dgp = gcol * e2 / T + beta * tp * d.nh_T_pg * (mh + HEPOL * mhe * heab);
  } else {
    double ridx = (double)1.0004f;
    if (lam * 1e8 >= 1800.0) ridx = refrax(lam * 1e4);
    const double eup = L.eplow + (HPLA * LIGHT / (ridx * lam)) / EVOLT;
    double ed1, ed2, ch;
    if (L.ion == 1) {
      ed1 = fmax(fz_xi[z] - eup, 1.0);
      ed2 = fmax(fz_xi[z] - L.eplow, 3.0);
      ch = 1.0;
    } else {
      ed1 = fmax(fz_xii[z] - eup - fz_xi[z], 1.0);
      ed2 = fmax(fz_xii[z] - L.eplow - fz_xi[z], 3.0);
      ch = (double)1.741f;
    }
    ch = ch * 17.0 * pow(8.0 * KBOL / DPI, 0.3);
    ch = ch * pow(pow(10.0, -30.5325) * (1.0 / (ed1 * ed1) - 1.0 / (ed2 * ed2)), 0.4);
    const double mh = pow(muh, -0.3), mhe = pow(muhe, -0.3);
    gcol = ch * (NH * mh + NH * heab * mhe) * pow(T, 0.3);
    if (der) dgp = ch * (mh + heab * mhe) * (d.nh_T_pg * pow(T, 0.3) + 0.3 * NH * pow(T, -0.7));
  }
  const double a = (grad + gcol) * lam / (4.0 * DPI * vd);
  out[0] = a;
  if (der) {
    out[1] = a / (grad + gcol) * dgp - a / (2.0 * T);
    out[2] = lam / (4.0 * DPI * vd) * gcol / NH * d.nh_pg_T;
    out[3] = lam / (4.0 * DPI * vd) * gcol / NH * d.nh_rho_T;
  }
}

}  // namespace fz
