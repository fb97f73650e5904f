// MI355X (gfx950) implementation of the FIRTEZ-dz per-column forward synthesis
// and analytic response functions, behind the C ABI of include/firtez_gpu.h.
//
// Kernels
//   saha/rho/opac/tau kernels : FP64 microphysics per (column, height): Saha
//                  factors, the electron-density fixed point (sequential in
//                  height per column, warm start from the height below as in
//                  FORWARD1D, forward.f90:911-960), implicit EOS derivatives,
//                  continuum/line opacity, damping.  Writes a per-height
//                  record [field][z][column] (columns fastest).
//   rt_kernel<RF> : one thread per (column, wavelength).  Pass 1 bottom->top:
//                  propagation matrix, evolution operator, formal solution
//                  I_k = (1-O_k) S_k + O_k I_{k-1} (rtesolver.f90:163-265);
//                  the incoming I of every height goes to a scratch slab.
//                  Pass 2 top->bottom (response functions): recomputes the
//                  propagation matrix with derivatives, the local response
//                  dO(I_in - S) + (1-O) dS, and propagates it to the top with
//                  the running product O_nz...O_{k+1} (forward.f90:511-548).
//   hsra_saha/rho/opac_kernel, gain_kernel : continuum normalisation (normalization.f90).
//   saha_mid/he_rk4 kernels : hydrostatic RK4 from the top
//                  (hydrostatic_tools.f90:150-396), standalone or inside the
//                  forward with HYDROSTATIC: yes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/firtez_gpu.h"
#include "fz_micro.hpp"
#include "fz_rt.hpp"

namespace fz {

constexpr int MAXL = 16, MAXR = 8, MAXS = 32, MAXC = 1024, MAXNZ = 512, MAXBLK = 128;
// internal response set of the inversion forward when neither Pg, rho nor P0
// is inverted and HYDROSTATIC is off: slots T|Pg, Bx, By, Bz, v_z only (the
// Pg|T and rho|T directional derivatives are not formed)
constexpr int RF_CART_TBV = 4;

constexpr int RT_WAVE = 64;   // wavelengths per radiative-transfer workgroup (one wave)
constexpr int NATM = 12;
enum { A_T = 0, A_PG, A_RHO, A_PEL, A_MU, A_BX, A_BY, A_BZ, A_VX, A_VY, A_VZ, A_TAU };
enum { F_FLAG = 0, F_KC5 = 1, F_REG = 2 };

// per-height record fields after the opacities: the field geometry (Geo,
// seven floats) and 1/sqrt(T)
constexpr int NGEO = 8;

struct Cfg {
  int nlines, nreg, nslot, numw, nz, nfield;
  int fgeo0;            // first geometry field of the per-height record
  LineConst line[MAXL];
  float dopfac[MAXL];   // Doppler width (mA) = dopfac * sqrt(T)
  float idopfac[MAXL];  // 1 / dopfac
  float vfac[MAXL];     // Doppler shift (mA) = vfac * v_z   (absorption_matrix.f90:1896)
  int reg_nblend[MAXR], reg_slot0[MAXR], reg_pix0[MAXR], reg_nwave[MAXR];
  double reg_wstart[MAXR], reg_wstep[MAXR], reg_l0main[MAXR];
  int slot_line[MAXS];
  float slot_bldiff[MAXS];
  int comp0[MAXL], ntype[MAXL][3];   // components sorted sigma_r, pi, sigma_b
  float comp_shift[MAXC], comp_str[MAXC];
  float dz_cm[MAXNZ];
  double zz[MAXNZ];    // heights in km, rounded through REAL(4) like MODEL(:,13)
  // radiative-transfer workgroups of one column: each covers up to RT_WAVE
  // wavelengths of a single region, so the column and the region (and every
  // per-height quantity read) are uniform across the wave
  int nbpc;
  int blk_reg[MAXBLK], blk_l0[MAXBLK], blk_n[MAXBLK];
  // hydrostatic RK4 steps k+1 -> k (hydrostatic_tools.f90:166-212): step dz
  // and the two weights of the mid-step temperature, with the running height
  // t0 accumulated exactly as the reference does
  double he_dz[MAXNZ], he_a[MAXNZ], he_b[MAXNZ];
  // m_u g dz / k per height (GET_HYD_DERIVATIVES, forward.f90:1033-1061)
  float he_cte[MAXNZ];
  // spectral LSF (MLSF): on/off and LSF_VALID per region
  int lsf_on;
  int reg_lsf[MAXR];
  // microturbulence xi: the Doppler width of line li is that of the
  // temperature T + xi_t[li], xi_t = m xi^2 / (2 k) (xi_on 0: thermal only)
  int xi_on;
  double xi_t[MAXL];
  float xi_tf[MAXL];
};
__constant__ Cfg c_cfg;

__device__ __forceinline__ int fslot(int s, int q) { return F_REG + 4 * c_cfg.nreg + 8 * s + q; }
__device__ __forceinline__ int freg(int r, int q) { return F_REG + 4 * r + q; }
__device__ __forceinline__ int fgeo(int q) { return c_cfg.fgeo0 + q; }
// 1 / (N_H m_u MW) of a height whose opacities are formed, 0 where the height
// is opaque (LINE OPTICAL DEPTH: KLINTAU = ETAI / N_H / m_u / MW)
__device__ __forceinline__ int ftlf() { return c_cfg.fgeo0 + NGEO; }

// ---------------------------------------------------------------- microphysics
// FORWARD1D's per-height physics (forward.f90:911-1006) over a tile of
// columns.  A column loop with one thread per column would give a few ten
// thousand threads (~1 wave per SIMD); only the electron-density fixed point
// is sequential in height (warm start from the height below), so the work is
// split into:
//   saha_kernel  one thread per (height, column): partition functions and the
//                Saha factors of all 92 elements -> sbuf[k][e][4][column]
//   rho_group_kernel  eight lanes per column, heights bottom->top: only the
//                warm-started electron-density fixed point (GET_RHO), reading
//                the factors from sbuf -> rho, mu, Pe and (n_e, N_H) per height
//   opac_kernel  one thread per (height, column): implicit EOS derivatives,
//                continuum and line opacity, damping -> the per-height record
//   tau_kernel   one thread per column: the REAL(4) log tau5 trapezoid
// FP64 throughout; the fixed point keeps the reference's iteration path
// (stopping rule 1e-4, warm start), so the EOS outputs are those of the
// sequential column loop.
struct MicroTile {
  int ncol, col0, ntile, der;
  float* atm;
  float* micro;
  double* sbuf;    // [nz][92][4][ntile]
  double* eos;     // [2][nz][ntile]: n_e, N_H
  double* deriv;   // [8][nz][ntile]: the DERIVVAR scalars (with response functions)
  int32_t* status; // [ncol]
  // NLTE departure coefficients [2][nz][nreg] (POPL, POPU), or null (LTE)
  const double* pop;
};

__global__ void __launch_bounds__(256) saha_kernel(MicroTile a) {
  const int nz = c_cfg.nz;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nz * a.ntile) return;
  const int k = (int)(t / a.ntile), ct = (int)(t - (long)k * a.ntile);
  const long cs = a.ncol, ps = (long)nz * a.ncol;
  const double T = (double)a.atm[A_T * ps + k * cs + a.col0 + ct];
  const double lam = g_c.csaha1 * pow(T, -1.5);
  const double dlam = -3.0 / (2.0 * T);
  double* o = a.sbuf + (long)k * FZ_NELEM * 4 * a.ntile + ct;
#pragma unroll 1
  for (int i = 0; i < FZ_NELEM; ++i) {
    double a1, a2, da1, da2;
    saha_element(i, T, lam, dlam, a.der != 0, a1, a2, da1, da2);
    o[(long)(i * 4 + 0) * a.ntile] = a1;
    o[(long)(i * 4 + 1) * a.ntile] = a2;
    o[(long)(i * 4 + 2) * a.ntile] = da1;
    o[(long)(i * 4 + 3) * a.ntile] = da2;
  }
}

// GET_RHO for one column by a group of RHO_G lanes: lane q holds the Saha
// factors of elements q, q + RHO_G, ... of the current height in registers
// (loaded once per height, not once per iteration), evaluates their stage
// populations, and every lane of the group folds the 91 terms in the
// reference's order j = 1 .. 91 through cross-lane reads, so the fixed point
// and its iteration path are those of get_rho bit for bit.  Eight lanes per
// column give eight times the waves of one lane per column for the same
// cube: the height recurrence is sequential, so the column count alone
// cannot fill the SIMDs.
constexpr int RHO_G = 8;
constexpr int RHO_NS = (FZ_NELEM + RHO_G - 1) / RHO_G;
// 32 columns per workgroup: the LDS rows of the groups (48.5 KB) fit three
// workgroups per CU
constexpr int RHO_BLOCK = 32 * RHO_G;

// sh: this group's LDS rows, nii of element j at sh[j], niii at sh[RHO_LD + j],
// ni of element 0 at sh[FZ_NELEM].  The lanes of a group belong to one wave,
// whose LDS operations complete in program order, so the fold reads what the
// stage pass wrote without a workgroup barrier.
constexpr int RHO_LD = 96;
// Row stride of the groups: the 64 / RHO_G groups of a wave read the same
// element index of their own rows at once (every lane of a group the same
// address, a broadcast); a stride of 2 RHO_LD + RHO_G / 4 doubles puts those
// rows 32 / (64 / RHO_G) banks apart, so the reads of a wave hit distinct
// banks instead of all one.
constexpr int RHO_ROW = 2 * RHO_LD + RHO_G / 4;

__device__ inline int get_rho_group(const double (&a1)[RHO_NS], const double (&a2)[RHO_NS], int q, double* sh,
                                    double T, double PG, double& ne, double& nh, double& pelec, double& dens,
                                    double& mu) {
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
#pragma unroll
    for (int m = 0; m < RHO_NS; ++m) {
      const int j = q + RHO_G * m;
      if (j < FZ_NELEM) {
        double ni, nii, niii;
        stages(h * g_c.abfac[j], nnew, a1[m], a2[m], ni, nii, niii);
        sh[j] = nii;
        sh[RHO_LD + j] = niii;
        if (j == 0) sh[FZ_NELEM] = ni;
      }
    }
    __builtin_amdgcn_wave_barrier();
    // the reference's fold: est = (niii_0 - ni_0) + (sum_{j>=1} nii_j) + (sum_{j>=1} niii_j)
    double s2 = 0.0, s3 = 0.0;
#pragma unroll 13
    for (int j = 1; j < FZ_NELEM; ++j) {
      s2 += sh[j];
      s3 += sh[RHO_LD + j];
    }
    double est = sh[RHO_LD] - sh[FZ_NELEM];
    __builtin_amdgcn_wave_barrier();
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

// get_rho_group on the factors of one height (base: [element][4][column]),
// out of line so that a caller with several fixed points (the RK4 steps)
// does not carry one unrolled copy of the element loop per call
struct RhoOut {
  double ne, nh, pel, dens, mu;
  int err;
};

__device__ __noinline__ RhoOut rho_group_at(const double* __restrict__ base, long stride, int q, double* sh, double T,
                                            double PG, double ne) {
  double a1[RHO_NS], a2[RHO_NS];
#pragma unroll
  for (int m = 0; m < RHO_NS; ++m) {
    const int j = q + RHO_G * m;
    a1[m] = j < FZ_NELEM ? base[(long)(j * 4 + 0) * stride] : 1.0;
    a2[m] = j < FZ_NELEM ? base[(long)(j * 4 + 1) * stride] : 1.0;
  }
  RhoOut o;
  o.ne = ne;
  o.err = get_rho_group(a1, a2, q, sh, T, PG, o.ne, o.nh, o.pel, o.dens, o.mu);
  return o;
}

__global__ void __launch_bounds__(RHO_BLOCK) rho_group_kernel(MicroTile a) {
  __shared__ double shall[RHO_BLOCK / RHO_G][RHO_ROW];
  const int lane = threadIdx.x & 63;
  const int q = lane & (RHO_G - 1);
  double* sh = shall[threadIdx.x / RHO_G];
  const int ct = (int)(((long)blockIdx.x * blockDim.x + threadIdx.x) / RHO_G);
  const bool active = ct < a.ntile;
  // a group past the tile repeats the last column without writing, so that
  // every group of a wave runs the same code
  const int cc = active ? ct : a.ntile - 1;
  const int nz = c_cfg.nz, col = a.col0 + cc;
  const long cs = a.ncol, ps = (long)nz * a.ncol;
  double ne = 0.0;
  int err = ERR_OK;
  for (int k = 0; k < nz; ++k) {
    const double T = (double)a.atm[A_T * ps + k * cs + col], P = (double)a.atm[A_PG * ps + k * cs + col];
    const double* sb = a.sbuf + (long)k * FZ_NELEM * 4 * a.ntile + cc;
    double a1[RHO_NS], a2[RHO_NS];
#pragma unroll
    for (int m = 0; m < RHO_NS; ++m) {
      const int j = q + RHO_G * m;
      a1[m] = j < FZ_NELEM ? sb[(long)(j * 4 + 0) * a.ntile] : 1.0;
      a2[m] = j < FZ_NELEM ? sb[(long)(j * 4 + 1) * a.ntile] : 1.0;
    }
    double nh, pel, dens, mu;
    err = get_rho_group(a1, a2, q, sh, T, P, ne, nh, pel, dens, mu);
    if (err) break;
    if (active && q == 0) {
      a.atm[A_RHO * ps + k * cs + col] = (float)dens;
      a.atm[A_MU * ps + k * cs + col] = (float)mu;
      a.atm[A_PEL * ps + k * cs + col] = (float)pel;
      a.eos[(long)k * a.ntile + cc] = ne;
      a.eos[((long)nz + k) * a.ntile + cc] = nh;
    }
  }
  if (active && q == 0) a.status[col] = err;
}

// The implicit EOS derivatives (GET_DNHYDDTEMP and its four solves,
// chemical_equilibrium.f90:2008-2052) per (height, column), kept apart from
// the opacities so that neither kernel carries the other's registers.
__global__ void __launch_bounds__(256) deriv_kernel(MicroTile a) {
  const int nz = c_cfg.nz;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nz * a.ntile) return;
  const int k = (int)(t / a.ntile), ct = (int)(t - (long)k * a.ntile), col = a.col0 + ct;
  if (a.status[col] != 0) return;
  const long cs = a.ncol, ps = (long)nz * a.ncol;
  const double T = (double)a.atm[A_T * ps + k * cs + col], P = (double)a.atm[A_PG * ps + k * cs + col];
  const float rhof = a.atm[A_RHO * ps + k * cs + col];
  const double ne = a.eos[(long)k * a.ntile + ct];
  const SahaView s(a.sbuf + (long)k * FZ_NELEM * 4 * a.ntile + ct, a.ntile);
  Deriv d = {};
  eos_derivatives(s, T, P, (double)rhof, ne, d);
  const long ds = (long)nz * a.ntile, o = (long)k * a.ntile + ct;
  a.deriv[0 * ds + o] = d.nh_T_pg;
  a.deriv[1 * ds + o] = d.nh_T_rho;
  a.deriv[2 * ds + o] = d.nh_pg_T;
  a.deriv[3 * ds + o] = d.nh_rho_T;
  a.deriv[4 * ds + o] = d.ne_pg_T;
  a.deriv[5 * ds + o] = d.ne_T_rho;
  a.deriv[6 * ds + o] = d.ne_T_pg;
  a.deriv[7 * ds + o] = d.ne_rho_T;
}

__global__ void __launch_bounds__(256) opac_kernel(MicroTile a) {
  const int nz = c_cfg.nz;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nz * a.ntile) return;
  const int k = (int)(t / a.ntile), ct = (int)(t - (long)k * a.ntile), col = a.col0 + ct;
  if (a.status[col] != 0) return;
  const long cs = a.ncol, ps = (long)nz * a.ncol;
  auto M = [&](int f) -> float& { return a.micro[f * ps + k * cs + col]; };
  {
    // FP32 field geometry and 1/sqrt(T) for the radiative transfer
    const float* at = a.atm + k * cs + col;
    const Geo g = geometry(at[A_BX * ps], at[A_BY * ps], at[A_BZ * ps]);
    M(fgeo(0)) = g.B;
    M(fgeo(1)) = g.cg;
    M(fgeo(2)) = g.sg;
    M(fgeo(3)) = g.sgf;
    M(fgeo(4)) = g.cgf;
    M(fgeo(5)) = g.c2p;
    M(fgeo(6)) = g.s2p;
    M(fgeo(7)) = rsqrtf(at[A_T * ps]);
  }
  const double T = (double)a.atm[A_T * ps + k * cs + col], P = (double)a.atm[A_PG * ps + k * cs + col];
  const float rhof = a.atm[A_RHO * ps + k * cs + col];
  const double ne = a.eos[(long)k * a.ntile + ct], nh = a.eos[((long)nz + k) * a.ntile + ct];
  const SahaView s(a.sbuf + (long)k * FZ_NELEM * 4 * a.ntile + ct, a.ntile);
  Deriv d = {};
  if (a.der) {
    const long ds = (long)nz * a.ntile, o = (long)k * a.ntile + ct;
    d.nh_T_pg = a.deriv[0 * ds + o];
    d.nh_T_rho = a.deriv[1 * ds + o];
    d.nh_pg_T = a.deriv[2 * ds + o];
    d.nh_rho_T = a.deriv[3 * ds + o];
    d.ne_pg_T = a.deriv[4 * ds + o];
    d.ne_T_rho = a.deriv[5 * ds + o];
    d.ne_T_pg = a.deriv[6 * ds + o];
    d.ne_rho_T = a.deriv[7 * ds + o];
  }
  (void)P;
  (void)rhof;
  double kc[4];
  int err = opac_cont(T, 5000.0, ne, nh, d, false, kc);
  if (!err) {
    const double kc5 = kc[0] / nh * (1.0 / MAMU);
    const bool bnd = (k == 0) || (kc5 >= 2.0e10);     // forward.f90:884,931,971
    M(F_FLAG) = bnd ? 1.0f : 0.0f;
    M(F_KC5) = (float)kc5;
    // the opacities are formed wherever kc5 is below the threshold, the
    // bottom height included (its absorption matrix feeds LINE OPTICAL DEPTH,
    // forward.f90:931-968, while the transfer takes it as the boundary)
    const bool opaque = kc5 >= 2.0e10;
    M(ftlf()) = opaque ? 0.f : (float)(1.0 / (nh * MAMU * (double)a.atm[A_MU * ps + k * cs + col]));
    if (opaque) return;
    for (int r = 0; r < c_cfg.nreg && !err; ++r) {
      err = opac_cont(T, c_cfg.reg_l0main[r], ne, nh, d, a.der != 0, kc);
      for (int q = 0; q < 4; ++q) M(freg(r, q)) = (float)kc[q];
      for (int b = 0; b < c_cfg.reg_nblend[r] && !err; ++b) {
        const int sl = c_cfg.reg_slot0[r] + b;
        const LineConst& L = c_cfg.line[c_cfg.slot_line[sl]];
        double kl[4] = {0, 0, 0, 0}, dm[4] = {0, 0, 0, 0};
        // the region's departure coefficients for every line of the region (forward.f90:949)
        const double bl = a.pop ? a.pop[((long)0 * nz + k) * c_cfg.nreg + r] : 1.0;
        const double bu = a.pop ? a.pop[((long)nz + k) * c_cfg.nreg + r] : 1.0;
        const double Td = T + c_cfg.xi_t[c_cfg.slot_line[sl]];
        err = opac_line(L, T, ne, nh, s, d, a.der != 0, kl, bl, bu, Td);
        damping(L, T, nh, d, a.der != 0, dm, Td);
        M(fslot(sl, 0)) = (float)kl[0];
        M(fslot(sl, 1)) = (float)dm[0];
        M(fslot(sl, 2)) = (float)kl[1];
        M(fslot(sl, 3)) = (float)kl[2];
        M(fslot(sl, 4)) = (float)kl[3];
        M(fslot(sl, 5)) = (float)dm[1];
        M(fslot(sl, 6)) = (float)dm[2];
        M(fslot(sl, 7)) = (float)dm[3];
      }
    }
  }
  if (err) atomicMax(&a.status[col], err);
}

__global__ void __launch_bounds__(64) tau_kernel(MicroTile a) {
  const int ct = blockIdx.x * blockDim.x + threadIdx.x;
  if (ct >= a.ntile) return;
  const int nz = c_cfg.nz, col = a.col0 + ct;
  if (a.status[col] != 0) return;
  const long cs = a.ncol, ps = (long)nz * a.ncol;
  // log tau_5000 by the REAL(4) trapezoid of forward.f90:1000-1006 (z in km)
  float tau = 1e-9f;
  a.atm[A_TAU * ps + (nz - 1) * cs + col] = log10f(tau);
  for (int k = nz - 2; k >= 0; --k) {
    const float x = (a.micro[F_KC5 * ps + k * cs + col] * a.atm[A_RHO * ps + k * cs + col] +
                     a.micro[F_KC5 * ps + (k + 1) * cs + col] * a.atm[A_RHO * ps + (k + 1) * cs + col]) /
                    2.0f;
    tau = (float)((double)tau + (double)x * ((c_cfg.zz[k + 1] - c_cfg.zz[k]) * 1000.0 * 100.0));
    a.atm[A_TAU * ps + k * cs + col] = log10f(tau);
  }
}

// ---------------------------------------------------------------- hydrostatic equilibrium
// RK4_INTEGRATION_LOG (hydrostatic_tools.f90:150-222, RK4_LOG :350-396) over
// a tile: the Saha factors at the heights come from saha_kernel, those at the
// RK4 mid-step temperatures from saha_mid_kernel, and he_rk4_group_kernel
// (eight lanes per column, top -> bottom) runs only the fixed points.
__device__ __forceinline__ double he_tmid(int k, double tm0, double tm1) {
  const double dz = c_cfg.he_dz[k];
  return tm0 / dz * c_cfg.he_a[k] + tm1 / dz * c_cfg.he_b[k];
}

__global__ void __launch_bounds__(256) saha_mid_kernel(MicroTile a, double* __restrict__ smid) {
  const int nz = c_cfg.nz;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)(nz - 1) * a.ntile) return;
  const int k = (int)(t / a.ntile), ct = (int)(t - (long)k * a.ntile);
  const long cs = a.ncol, ps = (long)nz * a.ncol;
  const double tm0 = (double)a.atm[A_T * ps + (k + 1) * cs + a.col0 + ct];
  const double tm1 = (double)a.atm[A_T * ps + k * cs + a.col0 + ct];
  const double T = he_tmid(k, tm0, tm1);
  const double lam = g_c.csaha1 * pow(T, -1.5);
  double* o = smid + (long)k * FZ_NELEM * 4 * a.ntile + ct;
#pragma unroll 1
  for (int i = 0; i < FZ_NELEM; ++i) {
    double a1, a2, da1, da2;
    saha_element(i, T, lam, 0.0, false, a1, a2, da1, da2);
    o[(long)(i * 4 + 0) * a.ntile] = a1;
    o[(long)(i * 4 + 1) * a.ntile] = a2;
  }
}

__device__ __forceinline__ double he_step_rhs_group(const double* base, long stride, int q, double* sh, double lnp,
                                                    double tem, double& nel, int& err) {
  const RhoOut o = rho_group_at(base, stride, q, sh, tem, exp(lnp), nel);
  nel = o.ne;
  if (o.err) err = o.err;
  return -o.mu * MAMU * GRAV / (KBOL * tem);        // RK4_SOLVE_LOG (hydrostatic_tools.f90:53-71)
}

// RK4_INTEGRATION_LOG by groups of RHO_G lanes per column, each fixed point as
// get_rho_group (the values of the one-thread get_rho, bit for bit); the
// integration scalars are identical in the lanes of a group.  Writes Pg, rho,
// Pe, mu of every height and, for the forward, (n_e, N_H) with n_e = Pe/(k T)
// from the REAL(4) Pe and T (forward.f90:904).
__global__ void __launch_bounds__(RHO_BLOCK) he_rk4_group_kernel(MicroTile a, const double* __restrict__ smid) {
  __shared__ double shall[RHO_BLOCK / RHO_G][RHO_ROW];
  const int lane = threadIdx.x & 63;
  const int q = lane & (RHO_G - 1);
  double* sh = shall[threadIdx.x / RHO_G];
  const int ct = (int)(((long)blockIdx.x * blockDim.x + threadIdx.x) / RHO_G);
  const bool active = ct < a.ntile;
  const int cc = active ? ct : a.ntile - 1;
  const bool wr = active && q == 0;
  const int nz = c_cfg.nz, col = a.col0 + cc;
  const long cs = a.ncol, ps = (long)nz * a.ncol;
  const long hs = (long)FZ_NELEM * 4 * a.ntile;
  const double* SH = a.sbuf + cc;
  const double* SM = smid + cc;
  auto A = [&](int p, int k) -> float& { return a.atm[p * ps + k * cs + col]; };
  int err = 0;
  const int top = nz - 1;
  double u0 = log((double)A(A_PG, top));
  double nel = (double)A(A_PEL, top) / KBOL / (double)A(A_T, top);
  double nh, pel, dens, mw;
  {
    const RhoOut o = rho_group_at(SH + top * hs, a.ntile, q, sh, (double)A(A_T, top), exp(u0), nel);
    nel = o.ne, nh = o.nh, pel = o.pel, dens = o.dens, mw = o.mu, err = o.err;
  }
  if (wr) {
    A(A_PG, top) = (float)exp(u0);
    A(A_RHO, top) = (float)dens;
    A(A_PEL, top) = (float)pel;
    A(A_MU, top) = (float)mw;
    if (a.eos) a.eos[((long)nz + top) * a.ntile + cc] = nh;
  }
  for (int k = nz - 2; k >= 0 && !err; --k) {
    const double dz = c_cfg.he_dz[k];
    const double tm0 = (double)A(A_T, k + 1), tm1 = (double)A(A_T, k);
    const double k1 = he_step_rhs_group(SH + (k + 1) * hs, a.ntile, q, sh, u0, tm0, nel, err) * dz;
    const double tmid = he_tmid(k, tm0, tm1);
    const double k2 = he_step_rhs_group(SM + k * hs, a.ntile, q, sh, u0 + k1 / 2.0, tmid, nel, err) * dz;
    const double k3 = he_step_rhs_group(SM + k * hs, a.ntile, q, sh, u0 + k2 / 2.0, tmid, nel, err) * dz;
    const double k4 = he_step_rhs_group(SH + k * hs, a.ntile, q, sh, u0 + k3, tm1, nel, err) * dz;
    const double u1 = u0 + 1.0 / 6.0 * (k1 + 2.0 * k2 + 2.0 * k3 + k4);
    const RhoOut o = rho_group_at(SH + k * hs, a.ntile, q, sh, tm1, exp(u1), nel);
    nel = o.ne, nh = o.nh, pel = o.pel, dens = o.dens, mw = o.mu;
    if (o.err) err = o.err;
    if (wr) {
      A(A_PG, k) = (float)exp(u1);
      A(A_RHO, k) = (float)dens;
      A(A_PEL, k) = (float)pel;
      A(A_MU, k) = (float)mw;
      if (a.eos) a.eos[((long)nz + k) * a.ntile + cc] = nh;
    }
    u0 = u1;
  }
  if (a.eos && wr)
    for (int k = 0; k < nz; ++k)
      a.eos[(long)k * a.ntile + cc] = (double)A(A_PEL, k) / KBOL / (double)A(A_T, k);
  if (wr) a.status[col] = err;
}

// ---------------------------------------------------------------- radiative transfer
// Launch description of rt_kernel (host side; the kernel receives the
// pointers as separate restrict-qualified arguments so that loads whose
// address is uniform across the wave become scalar loads).
// One radiative-transfer workgroup of a cube with per-pixel spectral
// windows: column, region, first wavelength (index in the configured grid),
// wavelengths in this workgroup (<= RT_WAVE) and their position in the
// pixel's own spectrum.
struct RtItem {
  int col;
  short r, n;
  int l0, lout;
};

struct RtArgs {
  int ncol;          // columns in the whole cube (strides of atm/micro)
  int col0;          // first column of this launch
  int ntile;         // columns in this launch
  const float* atm;
  const float* micro;
  const float* gain; // [numw][4]
  int32_t* status;
  float* stokes;     // [ncol][4][numw]
  float* jac;        // [ncol - jcol0][np][nz][4][numw]
  int jcol0;         // column of jac[0]
  float* scratch;    // [nz][4][ntile*numw]
  // per-pixel windows (items null: every column uses every wavelength)
  const RtItem* items;   // this launch's workgroups
  const int64_t* poff;   // [ncol+1] prefix of the pixels' wavelength counts: Stokes at 4*poff[col]
  const double* pop;     // NLTE departure coefficients [2][nz][nreg], or null (LTE)
  float* fs;             // FULL_STOKES [ncol][nz][4][numw] (Stokes after every height), or null
};

enum { D_T = 0, D_P, D_R, D_V, D_B, D_G, D_F, NDIR };

// profile combination of the three component types (sigma_r, pi, sigma_b)
// into (I,Q,U,V) (absorption_matrix.f90:1946-2230)
struct Lin {
  float I, Q, U, V;
};
__device__ __forceinline__ Lin lin(const Geo& g, float xr, float xp, float xb) {
  const float qb = (0.5f * xp - 0.25f * (xr + xb)) * g.sgf;
  return {0.5f * g.sgf * xp + 0.25f * g.cgf * (xr + xb), qb * g.c2p, qb * g.s2p, 0.5f * g.cg * (xr - xb)};
}
__device__ __forceinline__ Lin lin_dgamma(const Geo& g, float xr, float xp, float xb) {
  const float dsgf = 2.f * g.sg * g.cg;   // d sin^2(gamma) / d gamma
  const float qb = (0.5f * xp - 0.25f * (xr + xb)) * dsgf;
  return {0.5f * dsgf * xp - 0.25f * dsgf * (xr + xb), qb * g.c2p, qb * g.s2p, -0.5f * g.sg * (xr - xb)};
}
__device__ __forceinline__ Lin lin_dphi(const Geo& g, float xr, float xp, float xb) {
  const float qb = (0.5f * xp - 0.25f * (xr + xb)) * g.sgf;
  return {0.f, -2.f * qb * g.s2p, 2.f * qb * g.c2p, 0.f};
}

// T, 1/T and 1/sqrt(T) (the inverse Doppler width of line li is
// idopfac[li] / sqrt(T)), v_z and the field geometry of one height
struct Layer {
  float T, iT, rsT, vz;
  Geo g;
};

// propagation matrix without derivatives
__device__ __forceinline__ void absorb(const RtArgs& a, int col, int k, int r, float wave, const Layer& ly,
                                       float x7[7]) {
  const long ps = (long)c_cfg.nz * a.ncol, cs = a.ncol;
  const float* mic = a.micro + k * cs + col;
  x7[0] = mic[freg(r, 0) * ps];
  for (int c = 1; c < 7; ++c) x7[c] = 0.f;
  for (int b = 0; b < c_cfg.reg_nblend[r]; ++b) {
    const int sl = c_cfg.reg_slot0[r] + b, li = c_cfg.slot_line[sl];
    const float kl = mic[fslot(sl, 0) * ps], dmp = mic[fslot(sl, 1) * ps];
    const float rsd = c_cfg.xi_on ? rsqrtf(ly.T + c_cfg.xi_tf[li]) : ly.rsT;   // 1/sqrt(Doppler temperature)
    const float idld = c_cfg.idopfac[li] * rsd;
    const float base = (wave - c_cfg.slot_bldiff[sl] + c_cfg.vfac[li] * ly.vz) * idld;
    const float bsc = ly.g.B * idld;
    float hs[3], fs[3];
    int m = c_cfg.comp0[li];
    for (int t = 0; t < 3; ++t) {
      float h = 0.f, f = 0.f;
      const int n = c_cfg.ntype[li][t];
      for (int j = 0; j < n; ++j, ++m) {
        float H, F;
        voigt(dmp, base + c_cfg.comp_shift[m] * bsc, H, F);
        h += c_cfg.comp_str[m] * H;
        f += c_cfg.comp_str[m] * F;
      }
      hs[t] = h;
      fs[t] = f;
    }
    const Lin p = lin(ly.g, hs[0], hs[1], hs[2]), q = lin(ly.g, fs[0], fs[1], fs[2]);
    x7[0] += kl * p.I;
    x7[1] += kl * p.Q;
    x7[2] += kl * p.U;
    x7[3] += kl * p.V;
    x7[4] += kl * q.Q;
    x7[5] += kl * q.U;
    x7[6] += kl * q.V;
  }
}

// propagation matrix and its derivatives along the 7 internal directions
// T|Pg, Pg|T, rho|T, v_z, |B|, inclination, azimuth (absorption_matrix.f90:1717-2642)
__device__ __forceinline__ void absorb_der(const RtArgs& a, int col, int k, int r, float wave, const Layer& ly,
                                           float x7[7], float dx[NDIR][7]) {
  const long ps = (long)c_cfg.nz * a.ncol, cs = a.ncol;
  const float* mic = a.micro + k * cs + col;
  x7[0] = mic[freg(r, 0) * ps];
  for (int c = 1; c < 7; ++c) x7[c] = 0.f;
  for (int i = 0; i < NDIR; ++i)
    for (int c = 0; c < 7; ++c) dx[i][c] = 0.f;
  dx[D_T][0] = mic[freg(r, 1) * ps];
  dx[D_P][0] = mic[freg(r, 2) * ps];
  dx[D_R][0] = mic[freg(r, 3) * ps];
  const float isp = 0.564189583547756f;   // 1/sqrt(pi)
  for (int b = 0; b < c_cfg.reg_nblend[r]; ++b) {
    const int sl = c_cfg.reg_slot0[r] + b, li = c_cfg.slot_line[sl];
    const float kl = mic[fslot(sl, 0) * ps], dmp = mic[fslot(sl, 1) * ps];
    const float dkl[3] = {mic[fslot(sl, 2) * ps], mic[fslot(sl, 3) * ps], mic[fslot(sl, 4) * ps]};
    const float dda[3] = {mic[fslot(sl, 5) * ps], mic[fslot(sl, 6) * ps], mic[fslot(sl, 7) * ps]};
    const float rsd = c_cfg.xi_on ? rsqrtf(ly.T + c_cfg.xi_tf[li]) : ly.rsT;   // 1/sqrt(Doppler temperature)
    const float idld = c_cfg.idopfac[li] * rsd;
    const float base = (wave - c_cfg.slot_bldiff[sl] + c_cfg.vfac[li] * ly.vz) * idld;
    const float bsc = ly.g.B * idld;
    // per component type: H, F, sum dH/du, dH/da, u*dH/du, u*dH/da, shift*dH/du, shift*dH/da
    float S[8][3];
    int m = c_cfg.comp0[li];
    for (int t = 0; t < 3; ++t) {
      float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      const int n = c_cfg.ntype[li][t];
      for (int j = 0; j < n; ++j, ++m) {
        const float sh = c_cfg.comp_shift[m], st = c_cfg.comp_str[m];
        const float u = base + sh * bsc;
        float H, F;
        voigt(dmp, u, H, F);
        const float hu = 2.f * (dmp * F - u * H);            // dH/du
        const float ha = 2.f * (dmp * H + u * F - isp);      // dH/da
        acc[0] += st * H;
        acc[1] += st * F;
        acc[2] += st * hu;
        acc[3] += st * ha;
        acc[4] += st * hu * u;
        acc[5] += st * ha * u;
        acc[6] += st * hu * sh;
        acc[7] += st * ha * sh;
      }
      for (int q = 0; q < 8; ++q) S[q][t] = acc[q];
    }
    const Geo& g = ly.g;
    const Lin pH = lin(g, S[0][0], S[0][1], S[0][2]);
    const Lin pF = lin(g, S[1][0], S[1][1], S[1][2]);
    const Lin pHu = lin(g, S[2][0], S[2][1], S[2][2]);
    const Lin pHa = lin(g, S[3][0], S[3][1], S[3][2]);
    const Lin pHuu = lin(g, S[4][0], S[4][1], S[4][2]);
    const Lin pHau = lin(g, S[5][0], S[5][1], S[5][2]);
    const Lin pHb = lin(g, S[6][0], S[6][1], S[6][2]);
    const Lin pHab = lin(g, S[7][0], S[7][1], S[7][2]);
    const Lin gH = lin_dgamma(g, S[0][0], S[0][1], S[0][2]);
    const Lin gF = lin_dgamma(g, S[1][0], S[1][1], S[1][2]);
    const Lin fH = lin_dphi(g, S[0][0], S[0][1], S[0][2]);
    const Lin fF = lin_dphi(g, S[1][0], S[1][1], S[1][2]);
    // eta components use the H sums; rho components use the F sums, whose
    // derivatives are dF/du = -dH/da and dF/da = dH/du for w = H + iF
    const float e7[7] = {pH.I, pH.Q, pH.U, pH.V, pF.Q, pF.U, pF.V};
    const float a7[7] = {pHa.I, pHa.Q, pHa.U, pHa.V, pHu.Q, pHu.U, pHu.V};      // d/da
    const float u7[7] = {pHu.I, pHu.Q, pHu.U, pHu.V, -pHa.Q, -pHa.U, -pHa.V};  // d/du
    const float uu7[7] = {pHuu.I, pHuu.Q, pHuu.U, pHuu.V, -pHau.Q, -pHau.U, -pHau.V};
    const float b7[7] = {pHb.I, pHb.Q, pHb.U, pHb.V, -pHab.Q, -pHab.U, -pHab.V};
    const float g7[7] = {gH.I, gH.Q, gH.U, gH.V, gF.Q, gF.U, gF.V};
    const float f7[7] = {fH.I, fH.Q, fH.U, fH.V, fF.Q, fF.U, fF.V};
    const float dudT = c_cfg.xi_on ? -0.5f * rsd * rsd : -0.5f * ly.iT;   // du/dT = -u/(2 T_Doppler)
    const float dudv = c_cfg.vfac[li] * idld;
#pragma unroll
    for (int c = 0; c < 7; ++c) {
      x7[c] += kl * e7[c];
      dx[D_T][c] += dkl[0] * e7[c] + kl * (dudT * uu7[c] + dda[0] * a7[c]);
      dx[D_P][c] += dkl[1] * e7[c] + kl * dda[1] * a7[c];
      dx[D_R][c] += dkl[2] * e7[c] + kl * dda[2] * a7[c];
      dx[D_V][c] += kl * dudv * u7[c];
      dx[D_B][c] += kl * idld * b7[c];
      dx[D_G][c] += kl * g7[c];
      dx[D_F][c] += kl * f7[c];
    }
  }
}

__device__ __forceinline__ Layer load_layer(const RtArgs& a, int col, int k) {
  const long ps = (long)c_cfg.nz * a.ncol, cs = a.ncol;
  const float* at = a.atm + k * cs + col;
  const float* mg = a.micro + k * cs + col;
  Layer ly;
  ly.T = at[A_T * ps];
  ly.iT = 1.0f / ly.T;
  ly.vz = at[A_VZ * ps];
  ly.g.B = mg[fgeo(0) * ps];
  ly.g.cg = mg[fgeo(1) * ps];
  ly.g.sg = mg[fgeo(2) * ps];
  ly.g.sgf = mg[fgeo(3) * ps];
  ly.g.cgf = mg[fgeo(4) * ps];
  ly.g.c2p = mg[fgeo(5) * ps];
  ly.g.s2p = mg[fgeo(6) * ps];
  ly.rsT = mg[fgeo(7) * ps];
  return ly;
}

// LINE OPTICAL DEPTH (forward.f90:966-968, 1008-1015): per (column,
// wavelength) KLINTAU_k = ETAI_k / N_H / m_u / MW_k from the propagation
// matrix of every height whose opacities are formed (0 where the height is
// opaque; the reference keeps the last computed ETAI there), then the REAL(4)
// trapezoid from the top (tau = 1e-9) down, and log10.
// out[((col * nz + k) * numw + l]; one lane per (column, wavelength).
__global__ void __launch_bounds__(256) taulin_kernel(RtArgs a, float* __restrict__ out) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int numw = c_cfg.numw, nz = c_cfg.nz;
  if (g >= (long)a.ntile * numw) return;
  const int ct = (int)(g / numw), l = (int)(g - (long)ct * numw), col = a.col0 + ct;
  if (a.status && a.status[col] != 0) return;
  int r = 0;
  while (r + 1 < c_cfg.nreg && l >= c_cfg.reg_pix0[r + 1]) ++r;
  const float wave = (float)(c_cfg.reg_wstart[r] + c_cfg.reg_wstep[r] * (double)(l - c_cfg.reg_pix0[r]));
  const long ps = (long)nz * a.ncol, cs = a.ncol;
  float* o = out + (long)col * nz * numw + l;
  for (int k = 0; k < nz; ++k) {
    const float fac = a.micro[ftlf() * ps + k * cs + col];
    float kl = 0.f;
    if (fac != 0.f) {
      const Layer ly = load_layer(a, col, k);
      float x7[7];
      absorb(a, col, k, r, wave, ly, x7);
      kl = x7[0] * fac;
    }
    o[(long)k * numw] = kl;
  }
  float tau = 1.0e-9f, kln = o[(long)(nz - 1) * numw];
  o[(long)(nz - 1) * numw] = log10f(tau);
  for (int k = nz - 2; k >= 0; --k) {
    const float klk = o[(long)k * numw];
    const float x = (klk * a.atm[A_RHO * ps + k * cs + col] + kln * a.atm[A_RHO * ps + (k + 1) * cs + col]) / 2.0f;
    tau = (float)((double)tau + (double)x * ((c_cfg.zz[k + 1] - c_cfg.zz[k]) * 1000.0 * 100.0));
    o[(long)k * numw] = log10f(tau);
    kln = klk;
  }
}

// HE (HYDROSTATIC: yes, response functions on): the T response of height K
// also carries the change of Pg below K through hydrostatic equilibrium,
// dPg_i/dT_K = Pg_i (-mu_K m_u g dz_K / (k T_K^2)) (GET_HYD_DERIVATIVES), i.e.
// g_K * W_K with W_K = sum_{i<K} O_K..O_{i+1} (local Pg|T response of i) Pg_i,
// skipping layers with O_11 <= 1e-5 (rtesolver.f90:218-231).  W is built in
// pass 1 with one extra directional derivative per layer, so the term costs
// O(nz) instead of the reference's O(nz^2).  In that mode the reference
// evaluates the rho, B and v responses with the layer's outgoing Stokes
// vector (STKIN is reset to SYN1D after the hydrostatic block); reproduced.
//
// PACK (one region, no per-pixel windows): the lanes of a launch run over
// the flattened (column, wavelength) pairs of the tile, 64 to a wave, so no
// lane idles when the wavelength count is not a multiple of 64 (112 = 64 +
// 48 would leave 16 of 128 lanes idle per column); a wave then spans up to
// two columns and reads their per-height quantities with vector loads.
template <int RF, bool HE, bool PACK>
__global__ void __launch_bounds__(RT_WAVE) rt_kernel(const float* __restrict__ atm, const float* __restrict__ micro,
                                                     const float* __restrict__ gain, int32_t* __restrict__ status,
                                                     float* __restrict__ stokes, float* __restrict__ jac,
                                                     float* __restrict__ scratch, int ncol, int col0, int ntile,
                                                     int jcol0, const RtItem* __restrict__ items,
                                                     const int64_t* __restrict__ poff,
                                                     const double* __restrict__ pop, float* __restrict__ fs) {
  RtArgs a;
  a.ncol = ncol;
  a.col0 = col0;
  a.ntile = ntile;
  a.atm = atm;
  a.micro = micro;
  a.gain = gain;
  a.status = status;
  a.stokes = stokes;
  a.jac = jac;
  a.jcol0 = jcol0;
  a.scratch = scratch;
  a.items = items;
  a.poff = poff;
  a.pop = pop;
  a.fs = fs;
  // parameters: 5 spherical, 7 Cartesian, 8 with P0 (the top pressure)
  constexpr int NP = (RF == FZ_RF_SPHERICAL || RF == RF_CART_TBV)
                         ? 5
                         : (RF == FZ_RF_CARTESIAN ? 7 : (RF == FZ_RF_CARTESIAN_P0 ? 8 : 0));
  constexpr int NPL = RF == FZ_RF_CARTESIAN_P0 ? 7 : NP;    // parameters with a local response
  const int numw = c_cfg.numw, nz = c_cfg.nz;
  // workgroup -> (column, region, first wavelength); lane -> wavelength
  int ct, r, l;
  long sbase;        // the column's Stokes: stokes[sbase + s * snumw + lo]
  int snumw, lo;
  long t, nthr;      // lane slot and lane count of the tile (scratch rows)
  long jbase;        // the column's response functions: jac[jbase + ((p*nz + k)*4 + s)*snumw]
  if constexpr (PACK) {
    const int g = (int)blockIdx.x * RT_WAVE + (int)threadIdx.x;
    if (g >= a.ntile * numw) return;
    ct = g / numw;
    l = g - ct * numw;
    r = 0;
    sbase = (long)(col0 + ct) * 4 * numw;
    snumw = numw;
    lo = l;
    t = g;
    nthr = (long)a.ntile * numw;
    jbase = (long)(col0 + ct - a.jcol0) * NP * nz * 4 * numw + l;
  } else if (items) {
    // per-pixel windows: scratch and response functions are ragged too
    const RtItem it = items[blockIdx.x];
    if ((int)threadIdx.x >= it.n) return;
    ct = it.col - col0;
    r = it.r;
    l = it.l0 + (int)threadIdx.x;
    const int64_t p0 = poff[it.col], pt = poff[col0];
    sbase = 4 * p0;
    snumw = (int)(poff[it.col + 1] - p0);
    lo = it.lout + (int)threadIdx.x;
    t = (p0 - pt) + lo;
    nthr = poff[col0 + a.ntile] - pt;
    jbase = (long)NP * nz * 4 * (p0 - poff[a.jcol0]) + lo;
  } else {
    ct = (int)(blockIdx.x / (unsigned)c_cfg.nbpc);
    const int bc = (int)(blockIdx.x - (unsigned)ct * (unsigned)c_cfg.nbpc);
    if ((int)threadIdx.x >= c_cfg.blk_n[bc]) return;
    r = c_cfg.blk_reg[bc];
    l = c_cfg.blk_l0[bc] + (int)threadIdx.x;
    sbase = (long)(col0 + ct) * 4 * numw;
    snumw = numw;
    lo = l;
    t = (long)ct * numw + l;
    nthr = (long)a.ntile * numw;
    jbase = (long)(col0 + ct - a.jcol0) * NP * nz * 4 * numw + l;
  }
  const int col = a.col0 + ct;
  if (a.status && a.status[col] != 0) {
    for (int s = 0; s < 4; ++s) a.stokes[sbase + (long)s * snumw + lo] = __builtin_nanf("");
    return;
  }
  const double wd = c_cfg.reg_wstart[r] + c_cfg.reg_wstep[r] * (double)(l - c_cfg.reg_pix0[r]);
  const float wave = (float)wd;
  // Planck function constants (rtesolver.f90:1008-1035): S = c1 / (exp(c2/T) - 1)
  const double lam = (c_cfg.reg_l0main[r] + wd / 1.0e3) * 1.0e-8;
  const float c1 = (float)(2.0 * HPLA * LIGHT * LIGHT / (lam * lam * lam * lam * lam));
  const float c2 = (float)(HPLA * LIGHT / (lam * KBOL));
  const long ps = (long)nz * a.ncol, cs = a.ncol;
  constexpr bool HEX = HE && RF != 0;
  float I[4] = {0.f, 0.f, 0.f, 0.f};
  float Ahe[4] = {0.f, 0.f, 0.f, 0.f};
  float* scr = a.scratch + t;
  const long sstr = nthr;
  // POPL/POPU of the wavelength's region at height k (SOURCE_FUNCTION,
  // rtesolver.f90:1030-1033): S = c1 / ((POPL/POPU) exp(c2/T) - 1)
  auto depart = [&](int k) -> float {
    return a.pop ? (float)(a.pop[(long)k * c_cfg.nreg + r] / a.pop[((long)nz + k) * c_cfg.nreg + r]) : 1.0f;
  };
  // FULL_STOKES (forward.f90:980-983): the Stokes vector after every
  // height, times GAIN1D, at fs[((col * nz + k) * 4 + s) * numw + l]
  float* fsp = a.fs ? a.fs + (long)(col0 + ct) * nz * 4 * snumw + lo : nullptr;
  auto store_fs = [&](int k) {
    if (!fsp) return;
#pragma unroll
    for (int s = 0; s < 4; ++s) fsp[(long)(k * 4 + s) * snumw] = I[s] * a.gain[l * 4 + s];
  };
  for (int k = 0; k < nz; ++k) {
    const Layer ly = load_layer(a, col, k);
    const float S = c1 * (1.0f / (depart(k) * expf(c2 * ly.iT) - 1.0f));
    if (RF) {
#pragma unroll
      for (int s = 0; s < 4; ++s) scr[(long)(k * 4 + s) * sstr] = I[s];
    }
    if (a.micro[F_FLAG * ps + k * cs + col] != 0.f) {
      I[0] = S; I[1] = 0.f; I[2] = 0.f; I[3] = 0.f;
      store_fs(k);
      continue;
    }
    float x7[7];
    Evol E;
    if constexpr (HEX) {
      float dx[NDIR][7];
      absorb_der(a, col, k, r, wave, ly, x7, dx);
      const float dz = c_cfg.dz_cm[k];
      evol_build<true>(x7, dz, E);
      float v[4] = {I[0] - S, I[1], I[2], I[3]}, m4v[4], d2[4];
      m4mul(E.e, E.r, E.s0, v, m4v);
      evol_dir(E, dx[D_P], dz, v, m4v, d2);
      float W[4];
      const bool thin = E.O[0][0] > 1.0e-5f;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        W[i] = thin ? E.O[i][0] * Ahe[0] + E.O[i][1] * Ahe[1] + E.O[i][2] * Ahe[2] + E.O[i][3] * Ahe[3] : Ahe[i];
      const float pg = a.atm[A_PG * ps + k * cs + col];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        scr[(long)((nz + k) * 4 + i) * sstr] = W[i];
        if (thin) Ahe[i] = W[i] + d2[i] * pg;
      }
    } else {
      absorb(a, col, k, r, wave, ly, x7);
      evol_build<false>(x7, c_cfg.dz_cm[k], E);
    }
    float In[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
      In[i] = ((i == 0 ? 1.f : 0.f) - E.O[i][0]) * S + E.O[i][0] * I[0] + E.O[i][1] * I[1] + E.O[i][2] * I[2] +
              E.O[i][3] * I[3];
#pragma unroll
    for (int i = 0; i < 4; ++i) I[i] = In[i];
    store_fs(k);
  }
  float gn[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    gn[s] = a.gain[l * 4 + s];
    a.stokes[sbase + (long)s * snumw + lo] = I[s] * gn[s];
  }
  // a non-finite Stokes vector comes from a non-finite profile, propagation
  // matrix or evolution operator, where the reference stops
  // (voigt.f:91-95, absorption_matrix.f90:1882-1889, rtesolver.f90:346-352)
  if (a.status && !(__builtin_isfinite(I[0]) && __builtin_isfinite(I[1]) && __builtin_isfinite(I[2]) &&
                    __builtin_isfinite(I[3])))
    atomicMax(&a.status[col], (int32_t)ERR_NAN);
  if constexpr (RF != 0) {
    float acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = (i == j) ? 1.f : 0.f;
    float* jb = a.jac + jbase;
    for (int k = nz - 1; k >= 0; --k) {
      float outv[NP][4];
      const Layer ly = load_layer(a, col, k);
      if (a.micro[F_FLAG * ps + k * cs + col] != 0.f) {
        // boundary layer: no local response; layers below cannot reach the top
#pragma unroll
        for (int p = 0; p < NP; ++p)
#pragma unroll
          for (int s = 0; s < 4; ++s) outv[p][s] = 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = 0.f;
      } else {
        // S = c1 / (ex - 1), dS/dT = S (ex / (ex - 1)) (c2 / T) / T
        const float x = c2 * ly.iT;
        const float ex = depart(k) * expf(x);
        const float iem1 = 1.0f / (ex - 1.0f);
        const float S = c1 * iem1;
        const float dS = S * (ex * iem1) * x * ly.iT;
        float x7[7], dx[NDIR][7];
        absorb_der(a, col, k, r, wave, ly, x7, dx);
        Evol E;
        const float dz = c_cfg.dz_cm[k];
        evol_build<true>(x7, dz, E);
        float v[4], m4v[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) v[s] = scr[(long)(k * 4 + s) * sstr] - (s == 0 ? S : 0.f);
        m4mul(E.e, E.r, E.s0, v, m4v);
        float vo[4], m4vo[4];
        if constexpr (HEX) {
#pragma unroll
          for (int s = 0; s < 4; ++s) vo[s] = (k == nz - 1 ? I[s] : scr[(long)((k + 1) * 4 + s) * sstr]) - (s == 0 ? S : 0.f);
          m4mul(E.e, E.r, E.s0, vo, m4vo);
        }
        float loc[NDIR][4];
        constexpr int dirs_sph[5] = {D_T, D_B, D_G, D_F, D_V};
        constexpr int dirs_car[7] = {D_T, D_P, D_R, D_B, D_G, D_F, D_V};
#pragma unroll
        for (int q = 0; q < NPL; ++q) {
          const int dd = (RF == FZ_RF_SPHERICAL || RF == RF_CART_TBV) ? dirs_sph[q] : dirs_car[q];
          if (HEX && dd != D_T && dd != D_P) evol_dir(E, dx[dd], dz, vo, m4vo, loc[dd]);
          else evol_dir(E, dx[dd], dz, v, m4v, loc[dd]);
        }
        // source-function term (1 - O) e0 dS/dT for T
#pragma unroll
        for (int i = 0; i < 4; ++i) loc[D_T][i] += ((i == 0 ? 1.f : 0.f) - E.O[i][0]) * dS;
        if constexpr (HEX) {
          const float mu = a.atm[A_MU * ps + k * cs + col];
          const float gk = -mu * c_cfg.he_cte[k] * ly.iT * ly.iT;
#pragma unroll
          for (int i = 0; i < 4; ++i) loc[D_T][i] += gk * scr[(long)((nz + k) * 4 + i) * sstr];
        }
        float lp[NPL][4];
        if constexpr (RF == FZ_RF_SPHERICAL) {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            lp[0][i] = loc[D_T][i];
            lp[1][i] = loc[D_B][i];
            lp[2][i] = loc[D_G][i];
            lp[3][i] = loc[D_F][i];
            lp[4][i] = loc[D_V][i];
          }
        } else {
          // chain rule (|B|, gamma, phi) -> (Bx, By, Bz).  With B = 0 every
          // field derivative is zero, as in the reference (absorption_matrix.f90:1842-1866);
          // with a vertical field the azimuth terms are dropped (the reference's
          // d cos2phi / dBx is 0/0 there).
          const float* at = a.atm + k * cs + col;
          const float bx = at[A_BX * ps], by = at[A_BY * ps], bz = at[A_BZ * ps];
          const float bt2 = bx * bx + by * by, bt = sqrtf(bt2), B = ly.g.B, B2 = B * B;
          const bool hasB = B > 0.f, hasT = bt2 > 0.f;
          const float iB = hasB ? 1.f / B : 0.f, iBt = hasT ? 1.f / (B2 * bt) : 0.f, ibt2 = hasT ? 1.f / bt2 : 0.f;
          const float dB[3] = {bx * iB, by * iB, bz * iB};
          const float dG[3] = {bz * bx * iBt, bz * by * iBt, hasB ? -bt / B2 : 0.f};
          const float dF[3] = {-by * ibt2, bx * ibt2, 0.f};
          // slots T, (Pg, rho,) Bx, By, Bz, v_z
          constexpr int o = RF == RF_CART_TBV ? 0 : 2;
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            lp[0][i] = loc[D_T][i];
            if constexpr (RF != RF_CART_TBV) {
              lp[1][i] = loc[D_P][i];
              lp[2][i] = loc[D_R][i];
            }
#pragma unroll
            for (int j = 0; j < 3; ++j)
              lp[1 + o + j][i] = dB[j] * loc[D_B][i] + dG[j] * loc[D_G][i] + dF[j] * loc[D_F][i];
            lp[4 + o][i] = loc[D_V][i];
          }
        }
#pragma unroll
        for (int p = 0; p < NPL; ++p)
#pragma unroll
          for (int i = 0; i < 4; ++i)
            outv[p][i] = acc[i][0] * lp[p][0] + acc[i][1] * lp[p][1] + acc[i][2] * lp[p][2] + acc[i][3] * lp[p][3];
        if constexpr (RF == FZ_RF_CARTESIAN_P0) {
          // P0 (forward.f90:543-545): the local response is 1 in every Stokes
          // parameter; below the top it is propagated and then multiplied,
          // component by component, by PG(K)/PG(NZ) and the propagated Pg|T
          // response of the same height; the top layer keeps 1
          if (k == nz - 1) {
#pragma unroll
            for (int i = 0; i < 4; ++i) outv[7][i] = 1.f;
          } else {
            const float ratio = a.atm[A_PG * ps + k * cs + col] / a.atm[A_PG * ps + (nz - 1) * cs + col];
#pragma unroll
            for (int i = 0; i < 4; ++i)
              outv[7][i] = (acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3]) * ratio * outv[1][i];
          }
        }
        // running product O_nz ... O_k
        float nacc[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j)
            nacc[i][j] = acc[i][0] * E.O[0][j] + acc[i][1] * E.O[1][j] + acc[i][2] * E.O[2][j] + acc[i][3] * E.O[3][j];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = nacc[i][j];
      }
#pragma unroll
      for (int p = 0; p < NP; ++p)
#pragma unroll
        for (int s = 0; s < 4; ++s) jb[((long)(p * nz + k) * 4 + s) * snumw] = outv[p][s] * gn[s];
    }
  }
}

// ---------------------------------------------------------------- HSRA normalisation
// Equation of state and continuum opacity along the embedded HSRA model
// (normalization.f90:696-754); one thread, heights in order (warm start).
// HSRA's continuum opacities at the regions' central wavelengths
// (normalization.f90:696-754): the EOS of every HSRA height warm-started
// from the height below, as GET_RHO runs it there.  The Saha factors of all
// heights and elements are independent (one thread each), the fixed point is
// sequential in height (one group of RHO_G lanes, the fold in the
// reference's order, so the iteration path is get_rho's bit for bit), and the
// opacities are independent again (one thread per height and region).
__global__ void __launch_bounds__(256) hsra_saha_kernel(double* __restrict__ sb /*[387][92][4]*/) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= FZ_HSRA_LEN * FZ_NELEM) return;
  const int k = t / FZ_NELEM, i = t - k * FZ_NELEM;
  const double T = (double)fz_hsra_t[k];
  const double lam = g_c.csaha1 * pow(T, -1.5);
  double a1, a2, da1, da2;
  saha_element(i, T, lam, -3.0 / (2.0 * T), false, a1, a2, da1, da2);
  double* o = sb + ((long)k * FZ_NELEM + i) * 4;
  o[0] = a1;
  o[1] = a2;
}

__global__ void __launch_bounds__(64) hsra_rho_kernel(const double* __restrict__ sb, double* __restrict__ eos /*[2][387]*/,
                                                      int32_t* __restrict__ err_out /*[2]: code, first bad height*/) {
  __shared__ double sh[RHO_ROW];
  const int q = threadIdx.x;
  if (q >= RHO_G) return;
  double ne = 0.0;
  int err = ERR_OK, kbad = FZ_HSRA_LEN;
  for (int k = 0; k < FZ_HSRA_LEN; ++k) {
    const RhoOut o = rho_group_at(sb + (long)k * FZ_NELEM * 4, 1, q, sh, (double)fz_hsra_t[k],
                                  (double)fz_hsra_pg[k], ne);
    if (o.err) {
      err = o.err;
      kbad = k;
      break;
    }
    ne = o.ne;
    if (q == 0) {
      eos[k] = o.ne;
      eos[FZ_HSRA_LEN + k] = o.nh;
    }
  }
  if (q == 0) {
    err_out[0] = err;
    err_out[1] = kbad;
  }
}

__global__ void __launch_bounds__(256) hsra_opac_kernel(const double* __restrict__ eos, int32_t* __restrict__ err,
                                                        double* __restrict__ kc_out /*[387][nreg]*/) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= FZ_HSRA_LEN * c_cfg.nreg) return;
  const int k = t / c_cfg.nreg, r = t - k * c_cfg.nreg;
  if (k >= err[1]) return;
  Deriv d = {};
  double kc[4];
  const int e = opac_cont((double)fz_hsra_t[k], c_cfg.reg_l0main[r], eos[k], eos[FZ_HSRA_LEN + k], d, false, kc);
  kc_out[k * c_cfg.nreg + r] = kc[0];
  if (e) atomicMax(&err[2], e);
}

// Continuum formal solution through HSRA per wavelength; every step is
// rounded to REAL(4) like ANALYTICAL_SOLVER_HSRA (normalization.f90:466-480).
// The source function uses the last region's central wavelength for every
// pixel, which is what SOURCE_FUNCTION_HSRA's region test evaluates to
// (normalization.f90:679-682).
__global__ void gain_kernel(const double* __restrict__ kc, float* __restrict__ gain) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= c_cfg.numw) return;
  int r = 0;
  while (r + 1 < c_cfg.nreg && l >= c_cfg.reg_pix0[r + 1]) ++r;
  const double wave = c_cfg.reg_wstart[r] + c_cfg.reg_wstep[r] * (double)(l - c_cfg.reg_pix0[r]);
  const double lam = (c_cfg.reg_l0main[c_cfg.nreg - 1] + wave / 1.0e3) * 1.0e-8;
  const double c1 = 2.0 * HPLA * LIGHT * LIGHT / (lam * lam * lam * lam * lam), c2 = HPLA * LIGHT / (lam * KBOL);
  double syn = 0.0;
  for (int k = 0; k < FZ_HSRA_LEN; ++k) {
    const double sf = (double)(float)(c1 * (1.0 / (exp(c2 / (double)fz_hsra_t[k]) - 1.0)));
    if (k == 0) {
      syn = sf;
      continue;
    }
    double eta = kc[k * c_cfg.nreg + r];
    const bool imat = (eta == eta + 1.0);
    if (imat) eta = 1.0e15;
    const double dz = (k == FZ_HSRA_LEN - 1) ? 1.0e5 * fabs((double)fz_hsra_z[k] - (double)fz_hsra_z[k - 1])
                                             : 1.0e5 * fabs((double)fz_hsra_z[k + 1] - (double)fz_hsra_z[k]);
    const double ev = imat ? 1.0 : exp(-eta * dz);
    syn = (double)(float)((1.0 - ev) * sf + ev * syn);
  }
  const float g = (float)(1.0 / syn);
  gain[4 * l + 0] = g;
  gain[4 * l + 1] = g;
  gain[4 * l + 2] = g;
  gain[4 * l + 3] = -g;   // V sign convention (normalization.f90:743,751)
}


// ---------------------------------------------------------------- layout transposes for FORWARD3D
// Fortran cube (NZ,NY,NX) -> device SoA [z][col] with col = y + NY*x
__global__ void cube_to_soa(const float* __restrict__ src, float* __restrict__ dst, int nz, int ncol) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)nz * ncol) return;
  const int col = (int)(i / nz), k = (int)(i - (long)col * nz);
  dst[(long)k * ncol + col] = src[i];
}
__global__ void soa_to_cube(const float* __restrict__ src, float* __restrict__ dst, int nz, int ncol) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)nz * ncol) return;
  const int col = (int)(i / nz), k = (int)(i - (long)col * nz);
  dst[i] = src[(long)k * ncol + col];
}

}  // namespace fz

#include "fz_lm.hpp"
#include "fz_pert.hpp"
#include "fz_invert.hpp"
#include "fz_psf.hpp"
#include "fz_mlp.hpp"

namespace fz {
// Spectral line-spread function (MLSF) on rows of numw floats (Stokes rows,
// per-height or per-node response functions).  Per valid region of n pixels
// the reference multiplies the r2c DFT of the row by the DFT of LSF_KERNEL
// and transforms back with an unnormalised c2r (forward.f90:621-657,
// 814-844), i.e. the circular convolution
//     y[i] = sum_t tap[t] x[(i - t) mod n],  tap = n * LSF_KERNEL.
// One wave per row: the row is staged in LDS (the update is in place), lane i
// accumulates output i over the taps, the tap index is wave-uniform (scalar
// loads) and the LDS reads of consecutive lanes are consecutive.  FP32
// accumulation over n taps of a unit-sum kernel.  blockDim.x / 64 rows per
// workgroup (up to 4, with numw * rows floats of LDS <= 64 KiB).
__global__ void __launch_bounds__(256) lsf_kernel(float* __restrict__ rows, long nrows, const float* __restrict__ taps) {
  extern __shared__ float xs[];
  const int numw = c_cfg.numw;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long row = (long)blockIdx.x * (blockDim.x >> 6) + w;
  float* x = xs + w * numw;
  float* out = rows + row * numw;
  if (row < nrows)
    for (int i = lane; i < numw; i += 64) x[i] = out[i];
  __syncthreads();
  if (row >= nrows) return;
  for (int r = 0; r < c_cfg.nreg; ++r) {
    if (!c_cfg.reg_lsf[r]) continue;
    const int p0 = c_cfg.reg_pix0[r], n = c_cfg.reg_nwave[r];
    const float* xr = x + p0;
    const float* tr = taps + p0;
    for (int i0 = 0; i0 < n; i0 += 64) {
      const int i = i0 + lane;
      float acc = 0.f;
      int j = i < n ? i : 0;           // (i - t) mod n for t = 0; lanes past n idle in range
      for (int t = 0; t < n; ++t) {
        acc += tr[t] * xr[j];
        j = j == 0 ? n - 1 : j - 1;
      }
      if (i < n) out[p0 + i] = acc;
    }
  }
}
}  // namespace fz

// ==================================================================== C ABI
using namespace fz;

// Learned initial model: network shape, FP32 master parameters (flat:
// W_0, b_0, W_1, b_1, ...; W_l row-major [out][in]), their BF16 copy, Adam
// moments, gradient, normalisations and the per-batch work buffers.
struct MlpState {
  bool on = false;
  int nin = 0, nh = 0, nl = 0, nj = 0, nout = 0, nz = 0;
  long np = 0;
  long woff[8] = {0}, boff[8] = {0};
  int din[8] = {0}, dout[8] = {0};
  float *w = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
  bf16* wb = nullptr;
  float* norms = nullptr;   // xmean[nin], xstd[nin], tmean[nout], tstd[nout]
  double* pinv = nullptr;   // [nj][nz] least-squares node fit
  double* coefs = nullptr;  // [nz][nj] hat basis
  long step = 0;
  int cap = 0;              // rows the work buffers hold
  bf16* X = nullptr;
  bf16* act = nullptr;      // [nl][cap][nh]
  float* Y = nullptr;       // [cap][nout]
  float* dYf = nullptr;
  bf16* dYb = nullptr;
  float* dZf[2] = {nullptr, nullptr};
  bf16* dZb[2] = {nullptr, nullptr};
  float* partial = nullptr;
  float* loss = nullptr;
  void release() {
    void* ps[] = {w, g, m1, m2, wb, norms, pinv, coefs, X, act, Y, dYf, dYb, dZf[0], dZf[1], dZb[0], dZb[1],
                  partial, loss};
    for (void* q : ps)
      if (q) (void)hipFree(q);
    *this = MlpState();
  }
  void release_work() {
    void* ps[] = {X, act, Y, dYf, dYb, dZf[0], dZf[1], dZb[0], dZb[1], partial};
    for (void* q : ps)
      if (q) (void)hipFree(q);
    X = nullptr;
    act = nullptr;
    Y = dYf = nullptr;
    dYb = nullptr;
    dZf[0] = dZf[1] = nullptr;
    dZb[0] = dZb[1] = nullptr;
    partial = nullptr;
    cap = 0;
  }
};

struct fz_ctx {
  int device = 0;
  bool configured = false;
  bool have_gain = false;
  Cfg cfg;
  std::string err;
  float* d_gain = nullptr;
  float* d_micro = nullptr;
  size_t micro_cap = 0;
  float* d_scratch = nullptr;
  size_t scratch_cap = 0;
  int tile = 32768;
  int block = 256;
  bool prof = false;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  int rt_launches = 0;
  // microphysics of the next column tile on a second stream while the
  // radiative transfer of the current one runs (default priority); off by
  // default: on the 288x288x96 forward the concurrent kernels slowed the
  // transfer more than they hid (112.5 ms per step against 108)
  bool overlap = false;
  // STOKES SETUP: AUTO (input_file.f90:1294): per-pixel weights from the
  // observations in fz_invert3d_dev
  bool autow = false;
  hipStream_t aux = nullptr;
  std::vector<hipEvent_t> tev;
  // inversion
  bool inv_set = false;
  LmArgs lm;                  // per-cycle constants; per-call pointers filled at launch
  std::vector<int> slot_of;   // [np] physical parameter (= Cartesian slot) of each free parameter
  std::vector<int> kbeg, kend;  // [np] support of each node's hat function
  int* d_eqvnode = nullptr;   // [4][np]: kbeg, kend, Cartesian slot, five-slot index
  int* d_eqvtab_j = nullptr;  // [nparam][nz] node table of the pairwise contraction
  double* d_eqvtab_c = nullptr;  // [nparam][nz][2]
  bool eqv_pairs = false;
  bool lm_big = false;        // more than LM_MAXNP free parameters: work arrays in device memory
  size_t lm_work = 0;         // bytes of one pixel's work arrays (lm_big)
  double* d_lmwork = nullptr;
  size_t lmwork_cap = 0;
  double* d_coefs = nullptr;  // [np][nz]
  double* d_colsum = nullptr; // [np]
  float* d_jtile = nullptr;   // per-height response functions of one tile
  size_t jtile_cap = 0;
  // microphysics tile buffers
  int mtile = 0;               // microphysics columns per tile; 0: from the free memory (auto_mtile)
  float* d_sbuf = nullptr;    // Saha factors [nz][92][4][mtile] (double)
  size_t sbuf_cap = 0;
  float* d_eosb = nullptr;    // n_e, N_H [2][nz][mtile] (double)
  size_t eosb_cap = 0;
  float* d_stat = nullptr;    // per-column status when the caller passes none
  size_t stat_cap = 0;
  float* d_smid = nullptr;    // Saha factors at the RK4 mid-step temperatures (double)
  size_t smid_cap = 0;
  float* d_derb = nullptr;    // EOS derivatives [8][nz][mtile] (double)
  size_t derb_cap = 0;
  bool hydro = false;         // HYDROSTATIC: yes
  double xi = 0.0;            // microturbulence (cm/s); survives fz_configure
  size_t lm_lds = 0;
  bool lm_p0 = false;         // P0 among the inverted parameters
  bool lm_pr = false;         // Pg or rho among the inverted parameters
  // per-pixel spectral windows
  bool win_on = false;
  int win_ncol = 0;
  std::vector<int64_t> h_poff;      // [ncol+1]
  std::vector<int> h_item_first;    // [ncol+1] first RT workgroup of each column
  RtItem* d_items = nullptr;
  int64_t* d_poff = nullptr;
  int32_t* d_win = nullptr;         // [ncol][nreg][3]: first, count, position in the pixel's spectrum
  float* d_lsf = nullptr;     // LSF taps n * LSF_KERNEL [numw]
  double* d_pop = nullptr;    // NLTE departure coefficients [2][nz][nreg]
  bool pop_on = false;
  // GET_PERTURBATION / HESSIAN: normal matrix and solver work arrays, and
  // the device copies of the host-array entry points' inputs
  float* d_pert = nullptr;
  size_t pert_cap = 0;
  float* d_pstage = nullptr;
  size_t pstage_cap = 0;
  MlpState mlp;
  // spatial PSF of the coupled inversion: PSF(NPY, NPX) as given, and the
  // taps of the last map size it was cut for
  std::vector<double> h_psf;
  int psf_npx = 0, psf_npy = 0;
  float* d_taps = nullptr;
  int taps_nx = 0, taps_ny = 0, taps_rx = 0, taps_ry = 0, taps_ntx = 0;
  double taps_sq = 0.0;   // sum of the squared taps
};



// the context whose configuration is in a device's constant memory
static const fz_ctx* g_owner[64];

static int set_err(fz_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(c, x)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) return set_err((c), FZ_EHIP, std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)

// rows of numw floats through the LSF (no-op when it is off)
static int run_lsf(fz_ctx* c, float* d_rows, long nrows, hipStream_t st) {
  if (!c->cfg.lsf_on || nrows <= 0) return FZ_OK;
  const int rpb = std::max(1, std::min(4, 16384 / c->cfg.numw));
  const long nb = (nrows + rpb - 1) / rpb;
  if (nb > 0x7fffffffL) return set_err(c, FZ_EARG, "too many rows for the LSF launch");
  hipLaunchKernelGGL(lsf_kernel, dim3((unsigned)nb), dim3(64 * rpb), sizeof(float) * rpb * c->cfg.numw, st, d_rows,
                     nrows, (const float*)c->d_lsf);
  HIPCHK(c, hipGetLastError());
  return FZ_OK;
}

// microturbulence xi (cm/s) into the configuration: Doppler temperature
// offsets m xi^2 / (2 k) per line
static void set_xi(Cfg& g, const double* matom, double xi) {
  g.xi_on = xi > 0.0 ? 1 : 0;
  for (int i = 0; i < g.nlines; ++i) {
    g.xi_t[i] = xi * xi * matom[g.line[i].zn] * MAMU / (2.0 * KBOL);
    g.xi_tf[i] = (float)g.xi_t[i];
  }
}

static int upload_cfg(fz_ctx* c) {
  if (c->device < 64 && g_owner[c->device] == c) return FZ_OK;
  HIPCHK(c, hipMemcpyToSymbol(HIP_SYMBOL(c_cfg), &c->cfg, sizeof(Cfg)));
  if (c->device < 64) g_owner[c->device] = c;
  return FZ_OK;
}

// device allocation released when it goes out of scope (early error returns included)
struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

// per-device address of the fixed scratch array (fz_get_perturbation_dev_noalloc)
static int scratch_setup(int device);

extern "C" {

int fz_create(fz_ctx** out, int device) {
  if (!out) return FZ_EARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return FZ_EHIP;
  fz_ctx* c = new fz_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete c;
    return FZ_EHIP;
  }
  hipLaunchKernelGGL(fz_init_constants_kernel, dim3(1), dim3(1), 0, 0);
  if (hipDeviceSynchronize() != hipSuccess || scratch_setup(device) != FZ_OK) {
    delete c;
    return FZ_EHIP;
  }
  *out = c;
  return FZ_OK;
}

void fz_destroy(fz_ctx* c) {
  if (!c) return;
  if (c->device < 64 && g_owner[c->device] == c) g_owner[c->device] = nullptr;
  (void)hipSetDevice(c->device);
  for (hipEvent_t e : c->tev) (void)hipEventDestroy(e);
  if (c->aux) (void)hipStreamDestroy(c->aux);
  if (c->d_gain) (void)hipFree(c->d_gain);
  if (c->d_micro) (void)hipFree(c->d_micro);
  if (c->d_scratch) (void)hipFree(c->d_scratch);
  if (c->d_coefs) (void)hipFree(c->d_coefs);
  if (c->d_colsum) (void)hipFree(c->d_colsum);
  if (c->d_eqvnode) (void)hipFree(c->d_eqvnode);
  if (c->d_eqvtab_j) (void)hipFree(c->d_eqvtab_j);
  if (c->d_eqvtab_c) (void)hipFree(c->d_eqvtab_c);
  if (c->d_lmwork) (void)hipFree(c->d_lmwork);
  if (c->d_jtile) (void)hipFree(c->d_jtile);
  if (c->d_sbuf) (void)hipFree(c->d_sbuf);
  if (c->d_eosb) (void)hipFree(c->d_eosb);
  if (c->d_stat) (void)hipFree(c->d_stat);
  if (c->d_smid) (void)hipFree(c->d_smid);
  if (c->d_derb) (void)hipFree(c->d_derb);
  if (c->d_items) (void)hipFree(c->d_items);
  if (c->d_poff) (void)hipFree(c->d_poff);
  if (c->d_win) (void)hipFree(c->d_win);
  if (c->d_lsf) (void)hipFree(c->d_lsf);
  if (c->d_pop) (void)hipFree(c->d_pop);
  if (c->d_pert) (void)hipFree(c->d_pert);
  if (c->d_pstage) (void)hipFree(c->d_pstage);
  if (c->d_taps) (void)hipFree(c->d_taps);
  c->mlp.release();
  for (int i = 0; i < 3; ++i)
    if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
  delete c;
}

const char* fz_last_error(const fz_ctx* c) { return c ? c->err.c_str() : "null context"; }

// Zeeman pattern of one line (magsplit.f90:38-187), components grouped by type
static int zeeman_host(const fz_line& L, std::vector<float>& shift, std::vector<float>& str, int ntype[3]) {
  const int dj = (int)(L.ju - L.jl);
  if (std::abs(dj) > 1) return -1;
  double gl = 0.0, gu = 0.0;
  if (L.jl != 0) gl = 1.5 + (L.sl * (L.sl + 1.0) - L.ll * (L.ll + 1.0)) / (2.0 * L.jl * (L.jl + 1.0));
  if (L.ju != 0) gu = 1.5 + (L.su * (L.su + 1.0) - L.lu * (L.lu + 1.0)) / (2.0 * L.ju * (L.ju + 1.0));
  std::vector<double> sh[3], st[3];
  const int nml = (int)(2 * L.jl + 1), nmu = (int)(2 * L.ju + 1);
  for (int j = 0; j < nml; ++j) {
    const double ml = (double)j - L.jl;
    for (int k = 0; k < nmu; ++k) {
      const double mu = (double)k - L.ju;
      const int dm = (int)(mu - ml);
      if (std::abs(dm) > 1) continue;
      const double jl = L.jl, ju = L.ju;
      double s;
      if (dj == -1) s = dm == -1 ? (jl + mu) * (ju + mu + 2.0) : (dm == 0 ? 2.0 * (jl * jl - mu * mu) : (jl - mu) * (ju - mu + 2.0));
      else if (dj == 0) s = dm == -1 ? (ju - mu) * (ju + mu + 1.0) : (dm == 0 ? 2.0 * mu * mu : (ju + mu) * (ju - mu + 1.0));
      else s = dm == -1 ? (ju - mu) * (jl - mu) : (dm == 0 ? 2.0 * (ju * ju - mu * mu) : (ju + mu) * (jl + mu));
      // shift in mA/G rounded to REAL(4) as in the reference
      sh[dm + 1].push_back((double)(float)(4.66685e-10 * (ml * gl - mu * gu) * L.lambda0 * L.lambda0));
      st[dm + 1].push_back(s);
    }
  }
  for (int t = 0; t < 3; ++t) {
    double sum = 0.0;
    for (double v : st[t]) sum += v;
    ntype[t] = (int)st[t].size();
    for (size_t i = 0; i < st[t].size(); ++i) {
      shift.push_back((float)sh[t][i]);
      str.push_back((float)(st[t][i] / sum));
    }
  }
  return 0;
}

}  // extern "C"

// drops the per-pixel windows (their regions are those of one configuration)
static void clear_windows(fz_ctx* c) {
  if (c->d_items) (void)hipFree(c->d_items);
  if (c->d_poff) (void)hipFree(c->d_poff);
  if (c->d_win) (void)hipFree(c->d_win);
  c->d_items = nullptr;
  c->d_poff = nullptr;
  c->d_win = nullptr;
  c->win_on = false;
  c->win_ncol = 0;
  c->h_poff.clear();
  c->h_item_first.clear();
}

extern "C" {

int fz_configure(fz_ctx* c, const fz_line* lines, int nlines, const fz_region* regs, int nreg, int nz,
                 const double* z_km) {
  if (!c || !lines || !regs || !z_km) return FZ_EARG;
  if (nlines < 1 || nlines > MAXL || nreg < 1 || nreg > MAXR || nz < 2 || nz > MAXNZ)
    return set_err(c, FZ_ECONFIG, "configuration outside limits");
  Cfg& g = c->cfg;
  memset(&g, 0, sizeof(g));
  g.nlines = nlines;
  g.nreg = nreg;
  g.nz = nz;
  std::vector<float> shift, str;
  double matom[FZ_NELEM];
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyFromSymbol(matom, HIP_SYMBOL(fz_matom), sizeof(matom)));
  for (int i = 0; i < nlines; ++i) {
    const fz_line& L = lines[i];
    if (L.zn < 1 || L.zn > FZ_NELEM || L.ion < 1 || L.ion > 3) return set_err(c, FZ_EION, "bad element/ion");
    g.line[i] = {L.zn - 1, L.ion, L.lambda0, L.loggf, L.eplow, L.sigma, L.alpha};
    g.dopfac[i] = (float)(1.0e3 * L.lambda0 / LIGHT * sqrt(2.0 * KBOL / (matom[L.zn - 1] * MAMU)));
    g.idopfac[i] = (float)(1.0 / (1.0e3 * L.lambda0 / LIGHT * sqrt(2.0 * KBOL / (matom[L.zn - 1] * MAMU))));
    g.vfac[i] = (float)(-1.0e3 * L.lambda0 / LIGHT);
    g.comp0[i] = (int)shift.size();
    if (zeeman_host(L, shift, str, g.ntype[i]) != 0) return set_err(c, FZ_ECONFIG, "|Delta J| > 1");
  }
  if ((int)shift.size() > MAXC) return set_err(c, FZ_ECONFIG, "too many Zeeman components");
  for (size_t i = 0; i < shift.size(); ++i) {
    g.comp_shift[i] = shift[i];
    g.comp_str[i] = str[i];
  }
  int numw = 0, ns = 0;
  for (int r = 0; r < nreg; ++r) {
    const fz_region& R = regs[r];
    if (R.nblend < 1 || R.nblend > 4 || R.nwave < 1) return set_err(c, FZ_ECONFIG, "bad region");
    g.reg_nblend[r] = R.nblend;
    g.reg_slot0[r] = ns;
    g.reg_pix0[r] = numw;
    g.reg_nwave[r] = R.nwave;
    g.reg_wstart[r] = R.wave_start;
    g.reg_wstep[r] = R.wave_step;
    g.reg_l0main[r] = lines[R.line[0]].lambda0;
    for (int b = 0; b < R.nblend; ++b) {
      if (R.line[b] < 0 || R.line[b] >= nlines || ns >= MAXS) return set_err(c, FZ_ECONFIG, "bad blend");
      g.slot_line[ns] = R.line[b];
      g.slot_bldiff[ns] = b == 0 ? 0.f : (float)(1.0e3 * (lines[R.line[b]].lambda0 - lines[R.line[0]].lambda0));
      ++ns;
    }
    numw += R.nwave;
  }
  g.nslot = ns;
  g.numw = numw;
  g.nbpc = 0;
  for (int r = 0; r < nreg; ++r)
    for (int l0 = g.reg_pix0[r]; l0 < g.reg_pix0[r] + g.reg_nwave[r]; l0 += RT_WAVE) {
      if (g.nbpc >= MAXBLK) return set_err(c, FZ_ECONFIG, "too many wavelengths");
      g.blk_reg[g.nbpc] = r;
      g.blk_l0[g.nbpc] = l0;
      g.blk_n[g.nbpc] = std::min(RT_WAVE, g.reg_pix0[r] + g.reg_nwave[r] - l0);
      ++g.nbpc;
    }
  g.fgeo0 = F_REG + 4 * nreg + 8 * ns;
  g.nfield = g.fgeo0 + NGEO + 1;      // + the line-optical-depth factor
  set_xi(g, matom, c->xi);
  // layer thickness uses the spacing to the next height, the top one the
  // spacing below (rtesolver.f90:439-443); z rounded to REAL(4) like MODEL(:,13)
  for (int k = 0; k < nz; ++k) {
    const double zk = (double)(float)z_km[k];
    const double zo = (k == nz - 1) ? (double)(float)z_km[k - 1] : (double)(float)z_km[k + 1];
    g.dz_cm[k] = (float)(1.0e5 * fabs(zo - zk));
    g.zz[k] = zk;
  }
  {
    double t0 = g.zz[nz - 1] * 1000.0 * 100.0;
    for (int k = nz - 2; k >= 0; --k) {
      const double dz = (g.zz[k] - g.zz[k + 1]) * 1000.0 * 100.0;
      const double it1 = t0 + dz / 2.0;
      g.he_dz[k] = dz;
      g.he_a[k] = t0 + dz - it1;
      g.he_b[k] = it1 - t0;
      t0 = t0 + dz;
    }
    g.he_cte[0] = 0.f;
    for (int k = 1; k < nz; ++k) g.he_cte[k] = (float)(MAMU / KBOL * GRAV * (g.zz[k] - g.zz[k - 1]) * 1.0e3 * 1.0e2);
  }
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyToSymbol(HIP_SYMBOL(c_cfg), &g, sizeof(Cfg)));
  if (c->device < 64) g_owner[c->device] = c;
  if (c->d_gain) (void)hipFree(c->d_gain);
  HIPCHK(c, hipMalloc(&c->d_gain, sizeof(float) * 4 * numw));
  std::vector<float> ones(4 * numw, 1.0f);
  for (int l = 0; l < numw; ++l) ones[4 * l + 3] = -1.0f;
  HIPCHK(c, hipMemcpy(c->d_gain, ones.data(), sizeof(float) * 4 * numw, hipMemcpyHostToDevice));
  clear_windows(c);
  c->pop_on = false;   // LTE until fz_set_departure (input_file.f90:504-508)
  c->configured = true;
  c->have_gain = false;
  return FZ_OK;
}

int fz_numw(const fz_ctx* c) { return (c && c->configured) ? c->cfg.numw : -1; }

int fz_set_gain(fz_ctx* c, const float* gain) {
  if (!c || !c->configured || !gain) return FZ_EARG;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpy(c->d_gain, gain, sizeof(float) * 4 * c->cfg.numw, hipMemcpyHostToDevice));
  c->have_gain = true;
  return FZ_OK;
}

int fz_set_departure(fz_ctx* c, const double* popl, const double* popu) {
  if (!c || !c->configured) return FZ_EARG;
  if (!popl || !popu) {
    c->pop_on = false;
    return FZ_OK;
  }
  const int nz = c->cfg.nz, nreg = c->cfg.nreg;
  const size_t n = (size_t)nz * nreg;
  std::vector<double> h(2 * n);
  for (size_t i = 0; i < n; ++i) {
    // POPL(r, k) at r + k*nreg, the device layout [2][nz][nreg]
    if (!(popl[i] > 0.0) || !(popu[i] > 0.0) || !std::isfinite(popl[i]) || !std::isfinite(popu[i]))
      return set_err(c, FZ_EARG, "departure coefficients must be finite and positive");
    h[i] = popl[i];
    h[n + i] = popu[i];
  }
  HIPCHK(c, hipSetDevice(c->device));
  if (c->d_pop) (void)hipFree(c->d_pop);
  c->d_pop = nullptr;
  HIPCHK(c, hipMalloc(&c->d_pop, sizeof(double) * 2 * n));
  HIPCHK(c, hipMemcpy(c->d_pop, h.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice));
  c->pop_on = true;
  return FZ_OK;
}

int fz_set_psf(fz_ctx* c, const double* psf, int npy, int npx) {
  if (!c) return FZ_EARG;
  if (!psf) {
    c->h_psf.clear();
    c->psf_npx = c->psf_npy = 0;
    c->taps_nx = c->taps_ny = 0;
    return FZ_OK;
  }
  if (npx < 1 || npy < 1 || npx / 2 > PSF_RMAX || npy / 2 > PSF_RMAX)
    return set_err(c, FZ_EARG, "PSF size outside 1 .. 63 pixels per axis");
  // READ_PSF takes NPX/2 pixels either side of (NPY/2 + 1, NPX/2 + 1): an even
  // side would reach one pixel past the array
  if (npx % 2 == 0 || npy % 2 == 0) return set_err(c, FZ_EARG, "PSF sides must be odd");
  for (long i = 0; i < (long)npx * npy; ++i)
    if (!std::isfinite(psf[i])) return set_err(c, FZ_EARG, "PSF values must be finite");
  c->h_psf.assign(psf, psf + (size_t)npx * npy);
  c->psf_npx = npx;
  c->psf_npy = npy;
  c->taps_nx = c->taps_ny = 0;
  return FZ_OK;
}

}  // extern "C"

// the taps of the PSF for an nx x ny map (READ_PSF's central part of
// half-widths NPX/2, NPY/2 around (NPY/2 + 1, NPX/2 + 1)); the periodic sum
// needs 2 r + 1 <= n along each axis
static int psf_taps(fz_ctx* c, int nx, int ny) {
  if (c->h_psf.empty()) return set_err(c, FZ_ECONFIG, "no PSF set (fz_set_psf)");
  if (c->taps_nx == nx && c->taps_ny == ny && c->d_taps) return FZ_OK;
  const int rx = c->psf_npx / 2, ry = c->psf_npy / 2;
  if (2 * rx + 1 > nx || 2 * ry + 1 > ny) return set_err(c, FZ_ECONFIG, "PSF wider than the map");
  const int x0 = c->psf_npx / 2, y0 = c->psf_npy / 2, nu = 2 * ry + 1;
  const int ntx = (2 * rx + 1 + 7) / 8 * 8;
  // two tap tables [nu][ntx] in the kernel's window order: the convolution's
  // (tap (t, u) = PSF(Y0 + ry - u, X0 + rx - t)) and the adjoint's
  // (PSF(Y0 + u - ry, X0 + t - rx)); zero beyond t = 2 rx
  std::vector<float> t((size_t)2 * nu * ntx, 0.f);
  double sq = 0.0;
  for (int dx = -rx; dx <= rx; ++dx)
    for (int dy = -ry; dy <= ry; ++dy) {
      const float v = (float)c->h_psf[(size_t)(y0 + dy) + (size_t)c->psf_npy * (x0 + dx)];
      t[(size_t)(ry - dy) * ntx + (rx - dx)] = v;
      t[(size_t)nu * ntx + (size_t)(ry + dy) * ntx + (rx + dx)] = v;
      sq += (double)v * (double)v;
    }
  HIPCHK(c, hipSetDevice(c->device));
  if (c->d_taps) (void)hipFree(c->d_taps);
  c->d_taps = nullptr;
  HIPCHK(c, hipMalloc(&c->d_taps, sizeof(float) * t.size()));
  HIPCHK(c, hipMemcpy(c->d_taps, t.data(), sizeof(float) * t.size(), hipMemcpyHostToDevice));
  c->taps_nx = nx;
  c->taps_ny = ny;
  c->taps_rx = rx;
  c->taps_ry = ry;
  c->taps_ntx = ntx;
  c->taps_sq = sq;
  return FZ_OK;
}

static void launch_psf(fz_ctx* c, int nx, int ny, int nmaps, const float* in, float* out, int adjoint, hipStream_t st) {
  PsfArgs a;
  a.nx = nx;
  a.ny = ny;
  a.nmaps = nmaps;
  a.rx = c->taps_rx;
  a.ry = c->taps_ry;
  a.npix = (long)nx * ny;
  const int ntx = c->taps_ntx;
  a.taps = c->d_taps + (adjoint ? (size_t)(2 * a.ry + 1) * ntx : 0);
  a.in = in;
  a.out = out;
  const size_t lds = sizeof(float) * (PSF_TX + ntx - 1) * (PSF_TY + 2 * a.ry);
  const dim3 grid((unsigned)((ny + PSF_TY - 1) / PSF_TY), (unsigned)((nx + PSF_TX - 1) / PSF_TX), (unsigned)nmaps);
  switch (ntx) {
    case 8: hipLaunchKernelGGL(psf_conv_kernel<8>, grid, dim3(256), lds, st, a); break;
    case 16: hipLaunchKernelGGL(psf_conv_kernel<16>, grid, dim3(256), lds, st, a); break;
    case 24: hipLaunchKernelGGL(psf_conv_kernel<24>, grid, dim3(256), lds, st, a); break;
    case 32: hipLaunchKernelGGL(psf_conv_kernel<32>, grid, dim3(256), lds, st, a); break;
    case 40: hipLaunchKernelGGL(psf_conv_kernel<40>, grid, dim3(256), lds, st, a); break;
    case 48: hipLaunchKernelGGL(psf_conv_kernel<48>, grid, dim3(256), lds, st, a); break;
    case 56: hipLaunchKernelGGL(psf_conv_kernel<56>, grid, dim3(256), lds, st, a); break;
    default: hipLaunchKernelGGL(psf_conv_kernel<64>, grid, dim3(256), lds, st, a); break;
  }
}

extern "C" {

int fz_psf_convolve_dev(fz_ctx* c, int nx, int ny, int nmaps, const float* d_in, float* d_out, int adjoint,
                        void* stream) {
  if (!c || nx < 1 || ny < 1 || nmaps < 1 || nmaps > 65535 || !d_in || !d_out || d_in == d_out) return FZ_EARG;
  int r = psf_taps(c, nx, ny);
  if (r) return r;
  HIPCHK(c, hipSetDevice(c->device));
  launch_psf(c, nx, ny, nmaps, d_in, d_out, adjoint != 0, (hipStream_t)stream);
  HIPCHK(c, hipGetLastError());
  return FZ_OK;
}

int fz_set_lsf(fz_ctx* c, const double* lsf_kernel, const int32_t* lsf_valid) {
  if (!c || !c->configured) return FZ_EARG;
  Cfg& g = c->cfg;
  HIPCHK(c, hipSetDevice(c->device));
  g.lsf_on = 0;
  for (int r = 0; r < g.nreg; ++r) g.reg_lsf[r] = 0;
  if (lsf_kernel) {
    std::vector<float> taps(g.numw, 0.f);
    for (int r = 0; r < g.nreg; ++r) {
      g.reg_lsf[r] = lsf_valid ? (lsf_valid[r] != 0) : 1;
      const int p0 = g.reg_pix0[r], n = g.reg_nwave[r];
      for (int i = 0; i < n; ++i) {
        if (!std::isfinite(lsf_kernel[p0 + i])) return set_err(c, FZ_EARG, "non-finite LSF kernel");
        taps[p0 + i] = (float)(lsf_kernel[p0 + i] * (double)n);
      }
    }
    if (!c->d_lsf) HIPCHK(c, hipMalloc(&c->d_lsf, sizeof(float) * RT_WAVE * MAXBLK));
    HIPCHK(c, hipMemcpy(c->d_lsf, taps.data(), sizeof(float) * g.numw, hipMemcpyHostToDevice));
    g.lsf_on = 1;
  }
  HIPCHK(c, hipMemcpyToSymbol(HIP_SYMBOL(c_cfg), &g, sizeof(Cfg)));
  if (c->device < 64) g_owner[c->device] = c;
  return FZ_OK;
}

int fz_set_windows(fz_ctx* c, int ncol, const int32_t* windows) {
  if (!c || !c->configured) return FZ_EARG;
  HIPCHK(c, hipSetDevice(c->device));
  clear_windows(c);
  if (!windows) return FZ_OK;
  if (ncol < 1) return set_err(c, FZ_EARG, "no columns");
  const Cfg& g = c->cfg;
  const int nreg = g.nreg;
  std::vector<int64_t> poff((size_t)ncol + 1, 0);
  std::vector<int> first((size_t)ncol + 1, 0);
  std::vector<RtItem> items;
  std::vector<int32_t> w((size_t)ncol * nreg * 3);
  for (int col = 0; col < ncol; ++col) {
    int pos = 0;
    for (int r = 0; r < nreg; ++r) {
      const int f0 = windows[((size_t)col * nreg + r) * 2], cnt = windows[((size_t)col * nreg + r) * 2 + 1];
      if (cnt < 0 || f0 < 0 || f0 + cnt > g.reg_nwave[r])
        return set_err(c, FZ_EARG, "window outside its region (column " + std::to_string(col) + ")");
      w[((size_t)col * nreg + r) * 3 + 0] = f0;
      w[((size_t)col * nreg + r) * 3 + 1] = cnt;
      w[((size_t)col * nreg + r) * 3 + 2] = pos;
      for (int o = 0; o < cnt; o += RT_WAVE) {
        RtItem it;
        it.col = col;
        it.r = (short)r;
        it.n = (short)std::min(RT_WAVE, cnt - o);
        it.l0 = g.reg_pix0[r] + f0 + o;
        it.lout = pos + o;
        items.push_back(it);
      }
      pos += cnt;
    }
    if (pos == 0) return set_err(c, FZ_EARG, "a pixel without wavelengths (column " + std::to_string(col) + ")");
    poff[col + 1] = poff[col] + pos;
    if (items.size() > (size_t)0x7fffffff) return set_err(c, FZ_EARG, "too many workgroups");
    first[col + 1] = (int)items.size();
  }
  HIPCHK(c, hipMalloc(&c->d_items, sizeof(RtItem) * items.size()));
  HIPCHK(c, hipMalloc(&c->d_poff, sizeof(int64_t) * poff.size()));
  HIPCHK(c, hipMalloc(&c->d_win, sizeof(int32_t) * w.size()));
  HIPCHK(c, hipMemcpy(c->d_items, items.data(), sizeof(RtItem) * items.size(), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->d_poff, poff.data(), sizeof(int64_t) * poff.size(), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->d_win, w.data(), sizeof(int32_t) * w.size(), hipMemcpyHostToDevice));
  c->h_poff = poff;
  c->h_item_first = first;
  c->win_ncol = ncol;
  c->win_on = true;
  return FZ_OK;
}

int fz_window_offsets(const fz_ctx* c, int64_t* offsets) {
  if (!c || !c->win_on || !offsets) return FZ_EARG;
  std::memcpy(offsets, c->h_poff.data(), sizeof(int64_t) * c->h_poff.size());
  return FZ_OK;
}

int fz_gain_hsra(fz_ctx* c, float* gain) {
  if (!c || !c->configured) return FZ_EARG;
  HIPCHK(c, hipSetDevice(c->device));
  int ru = upload_cfg(c);
  if (ru) return ru;
  DevBuf b_kc, b_err, b_sb, b_eos;
  HIPCHK(c, hipMalloc(&b_kc.p, sizeof(double) * FZ_HSRA_LEN * c->cfg.nreg));
  HIPCHK(c, hipMalloc(&b_err.p, sizeof(int32_t) * 3));
  HIPCHK(c, hipMalloc(&b_sb.p, sizeof(double) * FZ_HSRA_LEN * FZ_NELEM * 4));
  HIPCHK(c, hipMalloc(&b_eos.p, sizeof(double) * 2 * FZ_HSRA_LEN));
  double* kc = (double*)b_kc.p;
  int32_t* derr = (int32_t*)b_err.p;
  const int32_t zero3[3] = {0, 0, 0};
  HIPCHK(c, hipMemcpy(derr, zero3, sizeof(zero3), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(hsra_saha_kernel, dim3((FZ_HSRA_LEN * FZ_NELEM + 255) / 256), dim3(256), 0, 0, (double*)b_sb.p);
  hipLaunchKernelGGL(hsra_rho_kernel, dim3(1), dim3(64), 0, 0, (const double*)b_sb.p, (double*)b_eos.p, derr);
  hipLaunchKernelGGL(hsra_opac_kernel, dim3((FZ_HSRA_LEN * c->cfg.nreg + 255) / 256), dim3(256), 0, 0,
                     (const double*)b_eos.p, derr, kc);
  HIPCHK(c, hipGetLastError());
  int32_t eh[3] = {0, 0, 0};
  HIPCHK(c, hipMemcpy(eh, derr, sizeof(eh), hipMemcpyDeviceToHost));
  // the first failure in height order, as the sequential loop meets it: the
  // EOS of height kbad, or an opacity of a height below it
  const int32_t e = eh[2] ? eh[2] : eh[0];
  if (e == ERR_LAMBDA) return set_err(c, e, "HSRA normalisation: wavelength below 1800 A");
  if (e) return set_err(c, e, "HSRA normalisation: equation of state or opacity failed");
  hipLaunchKernelGGL(gain_kernel, dim3((c->cfg.numw + 63) / 64), dim3(64), 0, 0, kc, c->d_gain);
  HIPCHK(c, hipGetLastError());
  if (gain) HIPCHK(c, hipMemcpy(gain, c->d_gain, sizeof(float) * 4 * c->cfg.numw, hipMemcpyDeviceToHost));
  c->have_gain = true;
  return FZ_OK;
}


static int ensure(fz_ctx* c, float** p, size_t* cap, size_t bytes) {
  if (*cap >= bytes) return FZ_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  if (hipMalloc((void**)p, bytes) != hipSuccess) return set_err(c, FZ_ENOMEM, "device allocation failed");
  *cap = bytes;
  return FZ_OK;
}

// floats of radiative-transfer scratch per lane (column-wavelength): the
// incoming Stokes vector of every layer (and W_k with HYDROSTATIC)
static size_t rt_scratch_per_lane(const fz_ctx* c, int rf_mode) {
  if (rf_mode == FZ_RF_NONE) return 0;
  const size_t nz = (size_t)c->cfg.nz;
  return nz * (c->hydro ? 8 : 4);
}

static int rf_nparam(int rf_mode) {
  return rf_mode == FZ_RF_SPHERICAL ? 5
         : rf_mode == FZ_RF_CARTESIAN ? 7
         : rf_mode == FZ_RF_CARTESIAN_P0 ? 8
         : rf_mode == RF_CART_TBV ? 5
         : 0;
}

static int launch_rt(fz_ctx* c, int rf_mode, const RtArgs& a, hipStream_t st, int nitems = 0) {
  if (a.items && nitems == 0) return FZ_OK;
  // packed lanes for one region without windows (the HYDROSTATIC and P0
  // instances keep one region-uniform wave per 64 wavelengths)
  const bool pack = !a.items && c->cfg.nreg == 1 && !c->hydro && rf_mode != FZ_RF_CARTESIAN_P0;
  const unsigned npack = (unsigned)(((long)a.ntile * c->cfg.numw + RT_WAVE - 1) / RT_WAVE);
  const dim3 grid(a.items ? (unsigned)nitems : pack ? npack : (unsigned)a.ntile * (unsigned)c->cfg.nbpc),
      blk(RT_WAVE);
#define FZ_RT_ARGS                                                                                               \
  a.atm, a.micro, a.gain, a.status, a.stokes, a.jac, a.scratch, a.ncol, a.col0, a.ntile, a.jcol0, a.items, a.poff, \
      c->pop_on ? c->d_pop : nullptr, a.fs
  if (pack) {
    if (rf_mode == FZ_RF_NONE) hipLaunchKernelGGL((rt_kernel<FZ_RF_NONE, false, true>), grid, blk, 0, st, FZ_RT_ARGS);
    else if (rf_mode == FZ_RF_SPHERICAL)
      hipLaunchKernelGGL((rt_kernel<FZ_RF_SPHERICAL, false, true>), grid, blk, 0, st, FZ_RT_ARGS);
    else if (rf_mode == RF_CART_TBV)
      hipLaunchKernelGGL((rt_kernel<RF_CART_TBV, false, true>), grid, blk, 0, st, FZ_RT_ARGS);
    else hipLaunchKernelGGL((rt_kernel<FZ_RF_CARTESIAN, false, true>), grid, blk, 0, st, FZ_RT_ARGS);
  }
  else if (rf_mode == FZ_RF_NONE) hipLaunchKernelGGL((rt_kernel<FZ_RF_NONE, false, false>), grid, blk, 0, st, FZ_RT_ARGS);
  else if (!c->hydro && rf_mode == FZ_RF_SPHERICAL)
    hipLaunchKernelGGL((rt_kernel<FZ_RF_SPHERICAL, false, false>), grid, blk, 0, st, FZ_RT_ARGS);
  else if (!c->hydro && rf_mode == FZ_RF_CARTESIAN_P0)
    hipLaunchKernelGGL((rt_kernel<FZ_RF_CARTESIAN_P0, false, false>), grid, blk, 0, st, FZ_RT_ARGS);
  else if (rf_mode == RF_CART_TBV) {
    if (c->hydro) return set_err(c, FZ_ECONFIG, "reduced response set without HYDROSTATIC only");
    hipLaunchKernelGGL((rt_kernel<RF_CART_TBV, false, false>), grid, blk, 0, st, FZ_RT_ARGS);
  }
  else if (c->hydro && rf_mode == FZ_RF_CARTESIAN_P0)
    hipLaunchKernelGGL((rt_kernel<FZ_RF_CARTESIAN_P0, true, false>), grid, blk, 0, st, FZ_RT_ARGS);
  else if (!c->hydro) hipLaunchKernelGGL((rt_kernel<FZ_RF_CARTESIAN, false, false>), grid, blk, 0, st, FZ_RT_ARGS);
  else if (rf_mode == FZ_RF_SPHERICAL)
    hipLaunchKernelGGL((rt_kernel<FZ_RF_SPHERICAL, true, false>), grid, blk, 0, st, FZ_RT_ARGS);
  else hipLaunchKernelGGL((rt_kernel<FZ_RF_CARTESIAN, true, false>), grid, blk, 0, st, FZ_RT_ARGS);
#undef FZ_RT_ARGS
  HIPCHK(c, hipGetLastError());
  return FZ_OK;
}

// Microphysics over all columns, tile by tile; status[ncol] receives each
// column's FZ_E* code (the caller's array, or the context's own).  With
// hydro, Pg is first re-integrated from the top (forward.f90:891-906) and the
// EOS takes N_H and n_e from that integration; with physics == false only the
// hydrostatic integration runs (fz_hydrostatic_dev).
// Microphysics tile when none is set: the electron-density fixed point is
// sequential in height, one thread per column, so a tile should hold as many
// columns as memory allows (up to a quarter of the free memory, at most
// 64 GiB, for the Saha factors and the EOS buffers of every column).
static int auto_mtile(fz_ctx* c, bool hydro) {
  size_t freeb = 0, totb = 0;
  if (hipMemGetInfo(&freeb, &totb) != hipSuccess) return 16384;
  const size_t per_col = sizeof(double) * (size_t)c->cfg.nz * (FZ_NELEM * 4 * (hydro ? 2 : 1) + 2 + 8);
  // what this context already holds for the tile counts as available
  const size_t held = c->sbuf_cap + c->eosb_cap + c->derb_cap + c->smid_cap;
  const size_t budget = std::min((freeb + held) / 4, (size_t)64 << 30);
  const size_t m = budget / per_col;
  return (int)std::max<size_t>(4096, std::min<size_t>(m, 1 << 20));
}

// the second stream and n ordering events of the overlapped forward
static int ensure_aux(fz_ctx* c, int n) {
  if (!c->aux) HIPCHK(c, hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
  while ((int)c->tev.size() < n) {
    hipEvent_t e;
    HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    c->tev.push_back(e);
  }
  return FZ_OK;
}

static int run_micro(fz_ctx* c, int ncol, float* d_atm, int32_t** status, bool der, bool hydro, bool physics,
                     hipStream_t st, int cbeg = 0, int cend = -1) {
  if (cend < 0) cend = ncol;
  const int nz = c->cfg.nz;
  int32_t* stv = *status;
  if (!stv) {
    int r = ensure(c, &c->d_stat, &c->stat_cap, sizeof(int32_t) * (size_t)ncol);
    if (r) return r;
    stv = (int32_t*)c->d_stat;
    *status = stv;
  }
  const int mt = std::min(c->mtile > 0 ? c->mtile : auto_mtile(c, hydro), ncol);
  int r = ensure(c, &c->d_sbuf, &c->sbuf_cap, sizeof(double) * (size_t)nz * FZ_NELEM * 4 * mt);
  if (r) return r;
  r = ensure(c, &c->d_eosb, &c->eosb_cap, sizeof(double) * (size_t)2 * nz * mt);
  if (r) return r;
  if (der && physics) {
    r = ensure(c, &c->d_derb, &c->derb_cap, sizeof(double) * (size_t)8 * nz * mt);
    if (r) return r;
  }
  if (hydro) {
    r = ensure(c, &c->d_smid, &c->smid_cap, sizeof(double) * (size_t)(nz - 1) * FZ_NELEM * 4 * mt);
    if (r) return r;
  }
  for (int c0 = cbeg; c0 < cend; c0 += mt) {
    MicroTile a;
    a.ncol = ncol;
    a.col0 = c0;
    a.ntile = std::min(mt, cend - c0);
    a.der = (der && physics) ? 1 : 0;
    a.atm = d_atm;
    a.micro = c->d_micro;
    a.sbuf = (double*)c->d_sbuf;
    a.eos = physics ? (double*)c->d_eosb : nullptr;
    a.deriv = (der && physics) ? (double*)c->d_derb : nullptr;
    a.status = stv;
    a.pop = c->pop_on ? c->d_pop : nullptr;
    const long nkt = (long)nz * a.ntile;
    hipLaunchKernelGGL(saha_kernel, dim3((unsigned)((nkt + 255) / 256)), dim3(256), 0, st, a);
    if (hydro) {
      const long nmt = (long)(nz - 1) * a.ntile;
      hipLaunchKernelGGL(saha_mid_kernel, dim3((unsigned)((nmt + 255) / 256)), dim3(256), 0, st, a,
                         (double*)c->d_smid);
      hipLaunchKernelGGL(he_rk4_group_kernel, dim3((unsigned)(((long)a.ntile * RHO_G + RHO_BLOCK - 1) / RHO_BLOCK)),
                         dim3(RHO_BLOCK), 0,
                         st, a, (const double*)c->d_smid);
    } else {
      hipLaunchKernelGGL(rho_group_kernel, dim3((unsigned)(((long)a.ntile * RHO_G + RHO_BLOCK - 1) / RHO_BLOCK)),
                         dim3(RHO_BLOCK), 0, st,
                         a);
    }
    if (physics) {
      if (a.der) hipLaunchKernelGGL(deriv_kernel, dim3((unsigned)((nkt + 255) / 256)), dim3(256), 0, st, a);
      hipLaunchKernelGGL(opac_kernel, dim3((unsigned)((nkt + 255) / 256)), dim3(256), 0, st, a);
      hipLaunchKernelGGL(tau_kernel, dim3((a.ntile + 63) / 64), dim3(64), 0, st, a);
    }
    HIPCHK(c, hipGetLastError());
  }
  return FZ_OK;
}

int fz_set_hydrostatic(fz_ctx* c, int on) {
  if (!c) return FZ_EARG;
  c->hydro = on != 0;
  return FZ_OK;
}

int fz_set_microturbulence(fz_ctx* c, double xi_cm_s) {
  if (!c || !std::isfinite(xi_cm_s) || xi_cm_s < 0.0 || xi_cm_s > 1.0e7) return FZ_EARG;
  c->xi = xi_cm_s;
  if (!c->configured) return FZ_OK;
  double matom[FZ_NELEM];
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyFromSymbol(matom, HIP_SYMBOL(fz_matom), sizeof(matom)));
  set_xi(c->cfg, matom, xi_cm_s);
  HIPCHK(c, hipMemcpyToSymbol(HIP_SYMBOL(c_cfg), &c->cfg, sizeof(Cfg)));
  if (c->device < 64) g_owner[c->device] = c;
  return FZ_OK;
}

int fz_set_tile(fz_ctx* c, int tile, int block) {
  if (!c || tile < 1 || block < 64 || block > 1024 || block % 64) return FZ_EARG;
  c->tile = tile;
  c->block = block;
  return FZ_OK;
}

}  // extern "C"

// FORWARD3D on device data; d_full (or null): the Stokes vector after every
// height, [ncol][nz][4][numw] (FULL_STOKES)
static int forward_impl(fz_ctx* c, int ncol, float* d_atm, int rf_mode, float* d_stokes, float* d_jac,
                        float* d_full, int32_t* d_status, void* stream) {
  if (!c || !c->configured || ncol < 1 || !d_atm || !d_stokes) return FZ_EARG;
  if (d_full && c->win_on) return set_err(c, FZ_EARG, "per-pixel windows: no full Stokes output");
  if (rf_mode != FZ_RF_NONE && !d_jac) return set_err(c, FZ_EARG, "jacobian output missing");
  if (rf_mode < 0 || rf_mode > 3) return FZ_EARG;
  if (c->win_on) {
    if (ncol != c->win_ncol) return set_err(c, FZ_EARG, "column count differs from fz_set_windows");
    if (rf_mode != FZ_RF_NONE)
      return set_err(c, FZ_EARG, "per-pixel windows: response functions come through the inversion only");
    if (c->cfg.lsf_on) return set_err(c, FZ_ECONFIG, "per-pixel windows and an LSF together are not supported");
  }
  const Cfg& g = c->cfg;
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(c, hipSetDevice(c->device));
  int r = upload_cfg(c);
  if (r) return r;
  r = ensure(c, &c->d_micro, &c->micro_cap, sizeof(float) * (size_t)g.nfield * g.nz * ncol);
  if (r) return r;
  // columns per radiative-transfer launch: the configured tile, fewer where
  // its scratch would pass 48 GiB
  const size_t per_col = sizeof(float) * rt_scratch_per_lane(c, rf_mode) * g.numw;
  int tile = std::min(c->tile, ncol);
  if (per_col) tile = (int)std::max<size_t>(1, std::min<size_t>((size_t)tile, ((size_t)48 << 30) / per_col));
  if (rf_mode != FZ_RF_NONE) {
    r = ensure(c, &c->d_scratch, &c->scratch_cap, per_col * tile);
    if (r) return r;
  }
  const int ntiles = (ncol + tile - 1) / tile;
  const bool ovl = c->overlap && ntiles > 1 && !c->win_on;
  if (c->prof) HIPCHK(c, hipEventRecord(c->ev[0], st));
  if (ovl) {
    // the microphysics of RT tile i+1 runs on the second stream while tile
    // i's radiative transfer runs on st; tile i's transfer waits for its own
    // microphysics only (the micro scratch is reused tile after tile on the
    // second stream, and the tiles' columns are disjoint)
    r = ensure_aux(c, ntiles + 1);
    if (r) return r;
    if (!d_status) {
      r = ensure(c, &c->d_stat, &c->stat_cap, sizeof(int32_t) * (size_t)ncol);
      if (r) return r;
      d_status = (int32_t*)c->d_stat;
    }
    HIPCHK(c, hipEventRecord(c->tev[0], st));
    HIPCHK(c, hipStreamWaitEvent(c->aux, c->tev[0], 0));
    for (int i = 0; i < ntiles; ++i) {
      r = run_micro(c, ncol, d_atm, &d_status, rf_mode != FZ_RF_NONE, c->hydro, true, c->aux, i * tile,
                    std::min(ncol, (i + 1) * tile));
      if (r) return r;
      HIPCHK(c, hipEventRecord(c->tev[i + 1], c->aux));
    }
  } else {
    r = run_micro(c, ncol, d_atm, &d_status, rf_mode != FZ_RF_NONE, c->hydro, true, st);
    if (r) return r;
  }
  if (ovl) HIPCHK(c, hipStreamWaitEvent(st, c->tev[1], 0));
  // with the overlap, "microphysics" is the first tile's, which nothing hides
  if (c->prof) HIPCHK(c, hipEventRecord(c->ev[1], st));
  c->rt_launches = 0;
  for (int c0 = 0; c0 < ncol; c0 += tile) {
    if (ovl && c0 > 0) HIPCHK(c, hipStreamWaitEvent(st, c->tev[c0 / tile + 1], 0));
    RtArgs a;
    a.ncol = ncol;
    a.col0 = c0;
    a.ntile = std::min(tile, ncol - c0);
    a.atm = d_atm;
    a.micro = c->d_micro;
    a.gain = c->d_gain;
    a.status = d_status;
    a.stokes = d_stokes;
    a.jac = d_jac;
    a.jcol0 = 0;
    a.scratch = c->d_scratch;
    a.items = nullptr;
    a.poff = nullptr;
    a.fs = d_full;
    int nit = 0;
    if (c->win_on) {
      a.items = c->d_items + c->h_item_first[c0];
      a.poff = c->d_poff;
      nit = c->h_item_first[c0 + a.ntile] - c->h_item_first[c0];
    }
    r = launch_rt(c, rf_mode, a, st, nit);
    if (r) return r;
    HIPCHK(c, hipGetLastError());
    c->rt_launches++;
  }
  r = run_lsf(c, d_stokes, (long)ncol * 4, st);
  if (r) return r;
  if (rf_mode != FZ_RF_NONE) {
    r = run_lsf(c, d_jac, (long)ncol * rf_nparam(rf_mode) * g.nz * 4, st);
    if (r) return r;
  }
  if (c->prof) HIPCHK(c, hipEventRecord(c->ev[2], st));
  return FZ_OK;
}

extern "C" {

int fz_forward_dev(fz_ctx* c, int ncol, float* d_atm, int rf_mode, float* d_stokes, float* d_jac, int32_t* d_status,
                   void* stream) {
  return forward_impl(c, ncol, d_atm, rf_mode, d_stokes, d_jac, nullptr, d_status, stream);
}

int fz_forward_full_dev(fz_ctx* c, int ncol, float* d_atm, int rf_mode, float* d_stokes, float* d_jac, float* d_full,
                        int32_t* d_status, void* stream) {
  if (!d_full) return FZ_EARG;
  return forward_impl(c, ncol, d_atm, rf_mode, d_stokes, d_jac, d_full, d_status, stream);
}

int fz_line_tau_dev(fz_ctx* c, int ncol, float* d_atm, float* d_tau, int32_t* d_status, void* stream) {
  if (!c || !c->configured || ncol < 1 || !d_atm || !d_tau) return FZ_EARG;
  if (c->win_on) return set_err(c, FZ_EARG, "per-pixel windows: no line optical depth");
  const Cfg& g = c->cfg;
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(c, hipSetDevice(c->device));
  int r = upload_cfg(c);
  if (r) return r;
  r = ensure(c, &c->d_micro, &c->micro_cap, sizeof(float) * (size_t)g.nfield * g.nz * ncol);
  if (r) return r;
  r = run_micro(c, ncol, d_atm, &d_status, false, c->hydro, true, st);
  if (r) return r;
  RtArgs a = {};
  a.ncol = ncol;
  a.col0 = 0;
  a.ntile = ncol;
  a.atm = d_atm;
  a.micro = c->d_micro;
  a.gain = c->d_gain;
  a.status = d_status;
  const long n = (long)ncol * g.numw;
  hipLaunchKernelGGL(taulin_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, d_tau);
  HIPCHK(c, hipGetLastError());
  return FZ_OK;
}

int fz_set_overlap(fz_ctx* c, int on) {
  if (!c) return FZ_EARG;
  c->overlap = on != 0;
  return FZ_OK;
}

int fz_set_auto_weights(fz_ctx* c, int on) {
  if (!c) return FZ_EARG;
  c->autow = on != 0;
  return FZ_OK;
}

// SET_WEIGHTS3D's per-pixel work (get_dmodel.f90:2051-2170): SET_WEIGHTS1D
// (:2191-2224) on the pixel's observed row, WSTK = those weights in REAL(4),
// then SET_ISIGMAP (:356-373).  One lane per pixel; the row is nstk blocks of
// numw values, the blocks of the inverted Stokes parameters in I, Q, U, V
// order.  STDDEV (:2228-2251): mean as the sum of x_i / n, then the n - 1
// variance, both in sequence order.
__global__ void auto_weights_kernel(int ncol, int numw, int inv_mask, double isigma, const double* __restrict__ obs,
                                    double* __restrict__ isig) {
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  int nstk = 0;
  for (int s = 0; s < 4; ++s) nstk += (inv_mask >> s) & 1;
  const int nf = nstk * numw;
  const double* o = obs + col * nf;
  double w[4], minw = 1.0e60;
  const double dn = (double)numw;
  for (int s = 0, b = 0; s < 4; ++s) {
    w[s] = 0.0;
    if (!((inv_mask >> s) & 1)) continue;
    const double* x = o + (long)b * numw;
    double mn = 0.0;
    for (int i = 0; i < numw; ++i) mn += x[i] / dn;
    double v = 0.0;
    for (int i = 0; i < numw; ++i) v += (x[i] - mn) * (x[i] - mn);
    const double sd = sqrt(v / (dn - 1.0));
    w[s] = sqrt(1.0 / (sd * sd));
    if (w[s] < minw) minw = w[s];
    ++b;
  }
  minw *= 10.0;
  double ss = 0.0;
  for (int s = 0; s < 4; ++s) {
    if (w[s] > minw) w[s] = minw;
    ss += w[s] * w[s];
  }
  const double nrm = sqrt(ss) / sqrt((double)nstk);
  double* out = isig + col * nf;
  for (int s = 0, b = 0; s < 4; ++s) {
    if (!((inv_mask >> s) & 1)) continue;
    const float ws = (float)(w[s] / nrm);
    const double v = (double)(ws * ws) / (isigma * isigma);
    for (int i = 0; i < numw; ++i) out[(long)b * numw + i] = v;
    ++b;
  }
}

int fz_auto_weights_dev(fz_ctx* c, int ncol, const double* d_obs, const int32_t* invstk, float noise,
                        double* d_isigmap, void* stream) {
  if (!c || !c->configured || ncol < 1 || !d_obs || !invstk || !d_isigmap) return FZ_EARG;
  if (c->win_on) return set_err(c, FZ_ECONFIG, "automatic Stokes weights take the dense spectral grid");
  if (c->cfg.numw < 2) return set_err(c, FZ_ECONFIG, "automatic Stokes weights need two wavelengths or more");
  int mask = 0;
  for (int s = 0; s < 4; ++s) mask |= invstk[s] ? 1 << s : 0;
  if (!mask) return set_err(c, FZ_ECONFIG, "no Stokes parameter inverted");
  HIPCHK(c, hipSetDevice(c->device));
  const double isigma = (double)(noise > 0 ? noise : 4.0e-3f);
  hipLaunchKernelGGL(auto_weights_kernel, dim3((ncol + 255) / 256), dim3(256), 0, (hipStream_t)stream, ncol,
                     c->cfg.numw, mask, isigma, d_obs, d_isigmap);
  HIPCHK(c, hipGetLastError());
  return FZ_OK;
}

int fz_set_profiling(fz_ctx* c, int on) {
  if (!c) return FZ_EARG;
  HIPCHK(c, hipSetDevice(c->device));
  if (on && !c->ev[0])
    for (int i = 0; i < 3; ++i) HIPCHK(c, hipEventCreate(&c->ev[i]));
  c->prof = on != 0;
  return FZ_OK;
}

int fz_last_timing(fz_ctx* c, float* micro_ms, float* rt_ms, int* launches) {
  if (!c || !c->prof) return FZ_EARG;
  HIPCHK(c, hipEventSynchronize(c->ev[2]));
  float a = 0.f, b = 0.f;
  HIPCHK(c, hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
  HIPCHK(c, hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
  if (micro_ms) *micro_ms = a;
  if (rt_ms) *rt_ms = b;
  if (launches) *launches = c->rt_launches;
  return FZ_OK;
}

int fz_autotune(fz_ctx* c, int ncol, float* d_atm, int rf_mode, float* d_stokes, float* d_jac, int32_t* d_status,
                void* stream, int* tile_out, int* mtile_out) {
  if (!c || !c->configured || ncol < 1 || !d_atm || !d_stokes) return FZ_EARG;
  HIPCHK(c, hipSetDevice(c->device));
  const bool prof = c->prof;
  const int mtile0 = c->mtile, tile0 = c->tile;
  // on a failed trial the context keeps the caller's tiles and profiling switch
  auto fail = [&](int code) {
    c->mtile = mtile0;
    c->tile = tile0;
    c->prof = prof;
    return code;
  };
  int r = fz_set_profiling(c, 1);
  if (r) return r;
  const int nz = c->cfg.nz, numw = c->cfg.numw;
  size_t freeb = 0, totb = 0;
  if (hipMemGetInfo(&freeb, &totb) != hipSuccess) return fail(set_err(c, FZ_EHIP, "hipMemGetInfo failed"));
  const size_t budget = freeb / 3;
  // microphysics tile: Saha factors, EOS and derivative buffers per column
  const size_t per_mcol = sizeof(double) * (size_t)nz * (FZ_NELEM * 4 * (c->hydro ? 2 : 1) + 2 + 8);
  // radiative-transfer tile: the Stokes scratch per column
  const size_t per_rcol = sizeof(float) * rt_scratch_per_lane(c, rf_mode) * numw;
  const int mcand[] = {4096, 8192, 16384, 32768, 65536, 131072, 262144};
  const int rcand[] = {2048, 4096, 8192, 16384, 32768};
  float best = 1e30f;
  int best_m = c->mtile, best_t = c->tile;
  for (int m : mcand) {
    if (m > 2 * ncol && m != mcand[0]) break;
    if ((size_t)std::min(m, ncol) * per_mcol > budget) break;
    c->mtile = m;
    r = fz_forward_dev(c, ncol, d_atm, rf_mode, d_stokes, d_jac, d_status, stream);
    if (r) return fail(r);
    float ms = 0.f;
    r = fz_last_timing(c, &ms, nullptr, nullptr);
    if (r) return fail(r);
    if (ms < best) {
      best = ms;
      best_m = m;
    }
  }
  c->mtile = best_m;
  best = 1e30f;
  if (rf_mode != FZ_RF_NONE) {
    for (int t : rcand) {
      if (t > 2 * ncol && t != rcand[0]) break;
      if ((size_t)std::min(t, ncol) * per_rcol > budget) break;
      c->tile = t;
      r = fz_forward_dev(c, ncol, d_atm, rf_mode, d_stokes, d_jac, d_status, stream);
      if (r) return fail(r);
      float ms = 0.f;
      r = fz_last_timing(c, nullptr, &ms, nullptr);
      if (r) return fail(r);
      if (ms < best) {
        best = ms;
        best_t = t;
      }
    }
  }
  c->tile = best_t;
  // release what the larger trials left behind; the next call allocates its size
  if (c->d_sbuf) (void)hipFree(c->d_sbuf);
  if (c->d_eosb) (void)hipFree(c->d_eosb);
  if (c->d_derb) (void)hipFree(c->d_derb);
  if (c->d_smid) (void)hipFree(c->d_smid);
  if (c->d_scratch) (void)hipFree(c->d_scratch);
  c->d_sbuf = c->d_eosb = c->d_derb = c->d_smid = c->d_scratch = nullptr;
  c->sbuf_cap = c->eosb_cap = c->derb_cap = c->smid_cap = c->scratch_cap = 0;
  c->prof = prof;
  if (tile_out) *tile_out = c->tile;
  if (mtile_out) *mtile_out = c->mtile;
  return FZ_OK;
}

int fz_hydrostatic_dev(fz_ctx* c, int ncol, float* d_atm, int32_t* d_status, void* stream) {
  if (!c || !c->configured || ncol < 1 || !d_atm) return FZ_EARG;
  HIPCHK(c, hipSetDevice(c->device));
  int r = upload_cfg(c);
  if (r) return r;
  r = run_micro(c, ncol, d_atm, &d_status, false, true, false, (hipStream_t)stream);
  return r;
}

int fz_forward3d(fz_ctx* c, int nx, int ny, float* tem3d, float* pg3d, float* rho3d, float* pel3d, float* mw3d,
                 const float* bx3d, const float* by3d, const float* bz3d, const float* vx3d, const float* vy3d,
                 const float* vz3d, float* tau3d5, const int32_t* am_i_done, int rf_mode, double* syn3d, float* dsyn) {
  if (!c || !c->configured || nx < 1 || ny < 1 || !syn3d) return FZ_EARG;
  const Cfg& g = c->cfg;
  const int nz = g.nz, numw = g.numw;
  const int np = rf_nparam(rf_mode);
  if (np && !dsyn) return FZ_EARG;
  // columns still to be computed (AM_I_DONE == 0, forward.f90:120-140)
  std::vector<int> cols;
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      if (!am_i_done || am_i_done[y + (size_t)ny * x] == 0) cols.push_back(y + ny * x);
  const int n = (int)cols.size();
  if (n == 0) return FZ_OK;
  const float* cubes[NATM] = {tem3d, pg3d, rho3d, pel3d, mw3d, bx3d, by3d, bz3d, vx3d, vy3d, vz3d, tau3d5};
  const size_t colbytes = sizeof(float) * nz;
  std::vector<float> h(colbytes / sizeof(float) * n);
  DevBuf b_atm, b_tmp, b_st, b_jac, b_status;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMalloc(&b_atm.p, colbytes * NATM * n));
  HIPCHK(c, hipMalloc(&b_tmp.p, colbytes * n));
  HIPCHK(c, hipMalloc(&b_st.p, sizeof(float) * 4 * numw * n));
  HIPCHK(c, hipMalloc(&b_status.p, sizeof(int32_t) * n));
  if (np) HIPCHK(c, hipMalloc(&b_jac.p, sizeof(float) * (size_t)np * nz * 4 * numw * n));
  float *d_atm = (float*)b_atm.p, *d_tmp = (float*)b_tmp.p, *d_st = (float*)b_st.p, *d_jac = (float*)b_jac.p;
  int32_t* d_status = (int32_t*)b_status.p;
  const unsigned tb = 256, tg = (unsigned)(((long)nz * n + tb - 1) / tb);
  for (int p = 0; p < NATM; ++p) {
    for (int i = 0; i < n; ++i) memcpy(&h[(size_t)i * nz], cubes[p] + (size_t)cols[i] * nz, colbytes);
    HIPCHK(c, hipMemcpy(d_tmp, h.data(), colbytes * n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(cube_to_soa, dim3(tg), dim3(tb), 0, 0, d_tmp, d_atm + (size_t)p * nz * n, nz, n);
  }
  int rc = fz_forward_dev(c, n, d_atm, rf_mode, d_st, d_jac, d_status, nullptr);
  if (rc == FZ_OK) {
    std::vector<int32_t> stat(n);
    HIPCHK(c, hipMemcpy(stat.data(), d_status, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n && rc == FZ_OK; ++i)
      if (stat[i]) rc = set_err(c, stat[i], "column " + std::to_string(cols[i]) + " failed");
  }
  if (rc == FZ_OK) {
    // model updated in place: Pg (re-integrated with HYDROSTATIC), rho, Pe,
    // mu, log tau5 (SPLIT_MODEL3D, forward.f90:1225-1270)
    float* outs[NATM] = {nullptr, pg3d, rho3d, pel3d, mw3d, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, tau3d5};
    for (int p = 0; p < NATM; ++p) {
      if (!outs[p]) continue;
      hipLaunchKernelGGL(soa_to_cube, dim3(tg), dim3(tb), 0, 0, d_atm + (size_t)p * nz * n, d_tmp, nz, n);
      HIPCHK(c, hipMemcpy(h.data(), d_tmp, colbytes * n, hipMemcpyDeviceToHost));
      for (int i = 0; i < n; ++i) memcpy(outs[p] + (size_t)cols[i] * nz, &h[(size_t)i * nz], colbytes);
    }
    std::vector<float> st((size_t)4 * numw * n);
    HIPCHK(c, hipMemcpy(st.data(), d_st, sizeof(float) * st.size(), hipMemcpyDeviceToHost));
    const size_t nf = (size_t)4 * numw;
    for (int i = 0; i < n; ++i)
      for (size_t f = 0; f < nf; ++f) syn3d[(size_t)cols[i] * nf + f] = (double)st[(size_t)i * nf + f];
    if (np) {
      const size_t per = (size_t)np * nz * 4 * numw;
      std::vector<float> jh(per);
      for (int i = 0; i < n; ++i) {
        HIPCHK(c, hipMemcpy(jh.data(), d_jac + (size_t)i * per, sizeof(float) * per, hipMemcpyDeviceToHost));
        // device [np][nz][4][numw] -> DSYN(NFREQ=4*numw, IFREEP=np*nz): same order
        memcpy(dsyn + (size_t)cols[i] * per, jh.data(), sizeof(float) * per);
      }
    }
  }
  return rc;
}

int fz_malloc(fz_ctx* c, void** p, uint64_t bytes) {
  if (!c || !p) return FZ_EARG;
  HIPCHK(c, hipSetDevice(c->device));
  if (hipMalloc(p, bytes) != hipSuccess) return set_err(c, FZ_ENOMEM, "hipMalloc failed");
  return FZ_OK;
}
int fz_free(fz_ctx* c, void* p) {
  if (!c) return FZ_EARG;
  HIPCHK(c, hipFree(p));
  return FZ_OK;
}
int fz_memcpy_h2d(fz_ctx* c, void* d, const void* s, uint64_t b) {
  HIPCHK(c, hipMemcpy(d, s, b, hipMemcpyHostToDevice));
  return FZ_OK;
}
int fz_memcpy_d2h(fz_ctx* c, void* d, const void* s, uint64_t b) {
  HIPCHK(c, hipMemcpy(d, s, b, hipMemcpyDeviceToHost));
  return FZ_OK;
}
int fz_synchronize(fz_ctx* c) {
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipDeviceSynchronize());
  return FZ_OK;
}

// ------------------------------------------------------------------ inversion

// GET_POSITIONS + GET_INT_COEFS (get_dmodel.f90:463-477, 563-644)
int fz_int_coefs(int nz, int nj, double* coefs) {
  if (!coefs || nz < 1 || nj < 1 || nj > nz) return FZ_EARG;
  std::fill(coefs, coefs + (size_t)nz * nj, 0.0);
  if (nj == 1) {
    for (int k = 0; k < nz; ++k) coefs[k] = 1.0;
    return FZ_OK;
  }
  if (nj == nz) {
    for (int k = 0; k < nz; ++k) coefs[(size_t)k * nz + k] = 1.0;
    return FZ_OK;
  }
  std::vector<double> pos(nj);
  const double step = (double)(nz - 1) / (double)(nj - 1);
  for (int j = 0; j < nj; ++j) pos[j] = ((double)(j + 1) - 1.0) * step + 1.0;
  for (int j = 0; j < nj; ++j) {
    const double lo = pos[std::max(j - 1, 0)], hi = pos[std::min(j + 1, nj - 1)], mid = pos[j];
    for (int i = 1; i <= nz; ++i) {
      const double x = (double)i;
      if (x < lo || x > hi) continue;
      double v;
      if (x == mid) v = 1.0;
      else if (x < mid) v = (1.0 - 0.0) / (mid - lo) * (x - lo);
      else v = (1.0 - 0.0) / (hi - mid) * (hi - x);
      coefs[(size_t)j * nz + (i - 1)] = v;
    }
  }
  return FZ_OK;
}

int fz_set_inversion(fz_ctx* c, const fz_inversion* inv, int* ifreep, int* nfreq) {
  if (!c || !inv) return FZ_EARG;
  if (!c->configured) return set_err(c, FZ_ECONFIG, "fz_configure first");
  const int nz = c->cfg.nz, numw = c->cfg.numw;
  LmArgs& L = c->lm;
  L = LmArgs();
  L.nz = nz;
  L.numw = numw;
  int nstk = 0;
  for (int s = 0; s < 4; ++s)
    if (inv->invstk[s]) L.rowstk[nstk++] = s;
  if (nstk == 0) return set_err(c, FZ_ECONFIG, "no Stokes parameter inverted");
  L.nfreq = nstk * numw;
  static const int mcol[8] = {A_T, A_PG, A_RHO, A_BX, A_BY, A_BZ, A_VZ, A_PG};
  std::vector<double> coefs;
  c->slot_of.clear();
  int np = 0, npar = 0;
  for (int p = 0; p < 8; ++p) {
    if (!inv->atmpar[p]) continue;
    const int nj = inv->nodes[p];
    if (nj < 1 || nj > nz) return set_err(c, FZ_ECONFIG, "nodes must be in 1..nz");
    if (np + nj > LM_MAXNP_BIG) return set_err(c, FZ_ECONFIG, "more than 2048 free parameters");
    L.par[npar] = p;
    L.slot[npar] = p;  // Cartesian slots: T|Pg, Pg|T, rho|T, Bx, By, Bz, v_z, P0
    L.mcol[npar] = mcol[p];
    L.j0[npar] = np;
    L.nj[npar] = nj;
    std::vector<double> cf((size_t)nz * nj);
    fz_int_coefs(nz, nj, cf.data());
    coefs.insert(coefs.end(), cf.begin(), cf.end());
    for (int j = 0; j < nj; ++j) c->slot_of.push_back(p);
    np += nj;
    ++npar;
  }
  if (np == 0) return set_err(c, FZ_ECONFIG, "no parameter inverted");
  L.np = np;
  L.nparam = npar;
  std::vector<double> colsum(np);
  c->kbeg.assign(np, 0);
  c->kend.assign(np, 0);
  for (int j = 0; j < np; ++j) {
    double w = 0.0;
    int kb = nz, ke = -1;
    for (int k = 0; k < nz; ++k) {
      const double v = coefs[(size_t)j * nz + k];
      w += v;
      if (v != 0.0) {
        kb = std::min(kb, k);
        ke = std::max(ke, k);
      }
    }
    colsum[j] = w;
    c->kbeg[j] = kb;
    c->kend[j] = ke;
  }
  L.regul = inv->regul ? 1 : 0;
  L.solver = inv->solver;
  if (L.solver != 1 && L.solver != 2) return set_err(c, FZ_ECONFIG, "solver must be 1 or 2");
  L.node_expansion = inv->node_expansion;
  if (L.node_expansion != 1 && L.node_expansion != 2) return set_err(c, FZ_ECONFIG, "node_expansion must be 1 or 2");
  for (int p = 0; p < 8; ++p) {
    L.pen_typ[p] = inv->pen_typ[p];
    L.pen_alp[p] = inv->pen_alp[p];
  }
  L.svdtol = inv->svdtol == 0.0 ? 1.0e-4 : inv->svdtol;
  L.inu = inv->inu == 0.0 ? (double)L.nfreq : inv->inu;
  L.mpert = inv->mpert == 0.0 ? 0.15 : inv->mpert;
  L.maxsteps = (double)inv->maxsteps;
  // up to LM_MAXNP free parameters the step's work arrays live in LDS;
  // beyond, in a per-pixel slice of device memory
  c->lm_big = np > LM_MAXNP;
  c->lm_lds = c->lm_big ? 0 : lm_lds_bytes(np, nz, L.regul != 0, LM_THREADS);
  c->lm_work = c->lm_big ? (lm_lds_bytes(np, nz, L.regul != 0, LM_BIG_THREADS) + 255) / 256 * 256 : 0;
  if (c->lm_lds > 160 * 1024 - 512) return set_err(c, FZ_ECONFIG, "inversion set-up exceeds the LDS of one workgroup");
  HIPCHK(c, hipSetDevice(c->device));
  if (c->d_coefs) (void)hipFree(c->d_coefs);
  if (c->d_colsum) (void)hipFree(c->d_colsum);
  if (c->d_eqvnode) (void)hipFree(c->d_eqvnode);
  c->d_coefs = nullptr;
  c->d_colsum = nullptr;
  c->d_eqvnode = nullptr;
  // node table of the contraction: supports, Cartesian slot, five-slot index
  static const int tbv_slot[8] = {0, -1, -1, 1, 2, 3, 4, -1};
  std::vector<int> node((size_t)4 * np);
  for (int j = 0; j < np; ++j) {
    node[j] = c->kbeg[j];
    node[np + j] = c->kend[j];
    node[2 * np + j] = c->slot_of[j];
    node[3 * np + j] = tbv_slot[c->slot_of[j]];
  }
  HIPCHK(c, hipMalloc((void**)&c->d_coefs, sizeof(double) * coefs.size()));
  HIPCHK(c, hipMalloc((void**)&c->d_colsum, sizeof(double) * np));
  HIPCHK(c, hipMalloc((void**)&c->d_eqvnode, sizeof(int) * node.size()));
  HIPCHK(c, hipMemcpy(c->d_coefs, coefs.data(), sizeof(double) * coefs.size(), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->d_colsum, colsum.data(), sizeof(double) * np, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->d_eqvnode, node.data(), sizeof(int) * node.size(), hipMemcpyHostToDevice));
  // the pairwise contraction applies when every parameter's node supports
  // overlap at most pairwise and move up with the node index (the hat basis
  // of GET_INT_COEFS); its per-height table
  c->eqv_pairs = true;
  for (int ip = 0; ip < npar; ++ip)
    for (int j = L.j0[ip]; j < L.j0[ip] + L.nj[ip]; ++j) {
      const int je = L.j0[ip] + L.nj[ip] - 1;
      if (c->kbeg[j] > c->kend[j]) c->eqv_pairs = false;
      if (j + 1 <= je && (c->kbeg[j + 1] < c->kbeg[j] || c->kend[j + 1] < c->kend[j])) c->eqv_pairs = false;
      if (j + 2 <= je && c->kbeg[j + 2] <= c->kend[j]) c->eqv_pairs = false;
    }
  if (c->d_eqvtab_j) (void)hipFree(c->d_eqvtab_j);
  if (c->d_eqvtab_c) (void)hipFree(c->d_eqvtab_c);
  c->d_eqvtab_j = nullptr;
  c->d_eqvtab_c = nullptr;
  if (c->eqv_pairs) {
    std::vector<int> tj((size_t)npar * nz, 0);
    std::vector<double> tc((size_t)npar * nz * 2, 0.0);
    for (int ip = 0; ip < npar; ++ip) {
      const int jb = L.j0[ip], je = jb + L.nj[ip] - 1;
      int jl = jb;
      for (int k = c->kbeg[jb]; k <= c->kend[je]; ++k) {
        while (k > c->kend[jl]) ++jl;
        int w = jl;
        if (k >= c->kbeg[jl]) {
          w |= 1 << 30;
          tc[((size_t)ip * nz + k) * 2] = coefs[(size_t)jl * nz + k];
        }
        if (jl < je && k >= c->kbeg[jl + 1]) {
          w |= 1 << 29;
          tc[((size_t)ip * nz + k) * 2 + 1] = coefs[(size_t)(jl + 1) * nz + k];
        }
        tj[(size_t)ip * nz + k] = w;
      }
    }
    HIPCHK(c, hipMalloc((void**)&c->d_eqvtab_j, sizeof(int) * tj.size()));
    HIPCHK(c, hipMalloc((void**)&c->d_eqvtab_c, sizeof(double) * tc.size()));
    HIPCHK(c, hipMemcpy(c->d_eqvtab_j, tj.data(), sizeof(int) * tj.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_eqvtab_c, tc.data(), sizeof(double) * tc.size(), hipMemcpyHostToDevice));
  }
  if (!c->lm_big)
    HIPCHK(c, hipFuncSetAttribute((const void*)lm_kernel<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)c->lm_lds));
  L.coefs = c->d_coefs;
  L.colsum = c->d_colsum;
  c->inv_set = true;
  c->lm_p0 = inv->atmpar[7] != 0;
  c->lm_pr = inv->atmpar[1] != 0 || inv->atmpar[2] != 0;
  if (ifreep) *ifreep = np;
  if (nfreq) *nfreq = L.nfreq;
  return FZ_OK;
}

int fz_forward_nodes_dev(fz_ctx* c, int ncol, float* d_atm, float* d_stokes, float* d_eqv, int32_t* d_status,
                         void* stream) {
  if (!c || !c->configured || ncol < 1 || !d_atm || !d_stokes || !d_eqv) return FZ_EARG;
  if (!c->inv_set) return set_err(c, FZ_ECONFIG, "fz_set_inversion first");
  if (c->win_on) {
    if (ncol != c->win_ncol) return set_err(c, FZ_EARG, "column count differs from fz_set_windows");
    if (c->cfg.lsf_on) return set_err(c, FZ_ECONFIG, "per-pixel windows and an LSF together are not supported");
  }
  const Cfg& g = c->cfg;
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(c, hipSetDevice(c->device));
  int r = upload_cfg(c);
  if (r) return r;
  r = ensure(c, &c->d_micro, &c->micro_cap, sizeof(float) * (size_t)g.nfield * g.nz * ncol);
  if (r) return r;
  // the P0 slot is synthesised only when P0 is inverted; the Pg and rho
  // slots only when they are inverted or feed another response (P0, HYDROSTATIC)
  const bool p0 = c->lm_p0;
  const bool tbv = !p0 && !c->lm_pr && !c->hydro;
  const int rf = p0 ? FZ_RF_CARTESIAN_P0 : (tbv ? RF_CART_TBV : FZ_RF_CARTESIAN);
  const int nslot = rf_nparam(rf);
  // columns per tile: the configured tile, fewer where one tile's response
  // functions and scratch would pass 24 GiB; sized by the widest tile (the
  // wavelengths of its columns: ragged with per-pixel windows)
  const size_t per_w = sizeof(float) * ((size_t)nslot * g.nz * 4 + rt_scratch_per_lane(c, rf));
  const size_t budget = (size_t)24 << 30;
  int tile = std::min(c->tile, ncol);
  size_t wmax = 0;
  for (;;) {
    wmax = 0;
    for (int c0 = 0; c0 < ncol; c0 += tile) {
      const int n = std::min(tile, ncol - c0);
      wmax = std::max(wmax, c->win_on ? (size_t)(c->h_poff[c0 + n] - c->h_poff[c0]) : (size_t)n * g.numw);
    }
    if (wmax * per_w <= budget || tile == 1) break;
    tile = std::max(1, (int)((double)tile * (double)budget / (double)(wmax * per_w)));
  }
  r = ensure(c, &c->d_scratch, &c->scratch_cap, sizeof(float) * rt_scratch_per_lane(c, rf) * wmax);
  if (r) return r;
  r = ensure(c, &c->d_jtile, &c->jtile_cap, sizeof(float) * (size_t)nslot * g.nz * 4 * wmax);
  if (r) return r;
  if (c->prof) HIPCHK(c, hipEventRecord(c->ev[0], st));
  r = run_micro(c, ncol, d_atm, &d_status, true, c->hydro, true, st);
  if (r) return r;
  if (c->prof) HIPCHK(c, hipEventRecord(c->ev[1], st));
  c->rt_launches = 0;
  EqvArgs e;
  e.nz = g.nz;
  e.numw = g.numw;
  e.nfreq = c->lm.nfreq;
  e.np = c->lm.np;
  e.nstk = c->lm.nfreq / g.numw;
  e.nslot = nslot;
  for (int s = 0; s < 4; ++s) e.rowstk[s] = c->lm.rowstk[s];
  const int np = e.np;
  e.kbeg = c->d_eqvnode;
  e.kend = c->d_eqvnode + np;
  e.slot_of = c->d_eqvnode + (tbv ? 3 : 2) * np;
  e.nparam = c->lm.nparam;
  for (int ip = 0; ip < e.nparam; ++ip) {
    e.j0[ip] = c->lm.j0[ip];
    e.nj[ip] = c->lm.nj[ip];
  }
  e.pairs = c->eqv_pairs ? 1 : 0;
  e.tab_j = c->d_eqvtab_j;
  e.tab_c = c->d_eqvtab_c;
  e.coefs = c->d_coefs;
  e.eqv = d_eqv;
  e.poff = c->win_on ? c->d_poff : nullptr;
  for (int c0 = 0; c0 < ncol; c0 += tile) {
    RtArgs a;
    a.ncol = ncol;
    a.col0 = c0;
    a.ntile = std::min(tile, ncol - c0);
    a.atm = d_atm;
    a.micro = c->d_micro;
    a.gain = c->d_gain;
    a.status = d_status;
    a.stokes = d_stokes;
    a.jac = c->d_jtile;
    a.jcol0 = c0;
    a.scratch = c->d_scratch;
    a.items = nullptr;
    a.poff = nullptr;
    a.fs = nullptr;
    int nit = 0;
    if (c->win_on) {
      a.items = c->d_items + c->h_item_first[c0];
      a.poff = c->d_poff;
      nit = c->h_item_first[c0 + a.ntile] - c->h_item_first[c0];
    }
    r = launch_rt(c, rf, a, st, nit);
    if (r) return r;
    e.ncol_tile = a.ntile;
    e.col0 = c0;
    e.jac = c->d_jtile;
    const long ne = c->win_on ? (long)e.nstk * (c->h_poff[c0 + a.ntile] - c->h_poff[c0]) : (long)a.ntile * e.nfreq;
    hipLaunchKernelGGL(eqv_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, e);
    HIPCHK(c, hipGetLastError());
    c->rt_launches++;
  }
  // the LSF acts on the node-contracted response functions, after GAIN1D
  // (forward.f90:608-657), and on the Stokes rows (:810-844)
  r = run_lsf(c, d_stokes, (long)ncol * 4, st);
  if (r) return r;
  r = run_lsf(c, d_eqv, (long)ncol * e.np * e.nstk, st);
  if (r) return r;
  if (c->prof) HIPCHK(c, hipEventRecord(c->ev[2], st));
  return FZ_OK;
}

}  // extern "C"

// per-pixel windows: the LM step reads ragged rows and normalises each
// pixel's chi^2 by its own number of wavelengths times inverted Stokes
static void set_lm_windows(const fz_ctx* c, LmArgs& a) {
  a.poff = c->win_on ? c->d_poff : nullptr;
  a.nstk = a.nfreq / a.numw;
  a.inu_pix = c->win_on ? 1 : 0;
}

// lm_kernel over nrows rows: one wave per row with the work arrays in LDS,
// or, beyond LM_MAXNP free parameters, four waves per row with the work
// arrays in device memory, in batches of rows whose slices fit 2 GiB
static int launch_lm(fz_ctx* c, LmArgs a, int nrows, hipStream_t st) {
  if (!c->lm_big) {
    a.work = nullptr;
    a.work_stride = 0;
    a.row0 = 0;
    hipLaunchKernelGGL(lm_kernel<false>, dim3(nrows), dim3(LM_THREADS), c->lm_lds, st, a);
    HIPCHK(c, hipGetLastError());
    return FZ_OK;
  }
  const int batch = (int)std::max<size_t>(1, std::min<size_t>((size_t)nrows, ((size_t)2 << 30) / c->lm_work));
  int r = ensure(c, (float**)&c->d_lmwork, &c->lmwork_cap, c->lm_work * batch);
  if (r) return r;
  a.work = c->d_lmwork;
  a.work_stride = (long)(c->lm_work / sizeof(double));
  for (int r0 = 0; r0 < nrows; r0 += batch) {
    a.row0 = r0;
    hipLaunchKernelGGL(lm_kernel<true>, dim3(std::min(batch, nrows - r0)), dim3(LM_BIG_THREADS), 0, st, a);
    HIPCHK(c, hipGetLastError());
  }
  return FZ_OK;
}

extern "C" {

int fz_lm_step_dev(fz_ctx* c, int ncol, const double* d_obs, const float* d_stokes, double* d_best_syn, double* d_sla,
                   const float* d_eqv, float* d_eqv_best, const double* d_isigmap, float* d_atm, float* d_atm_best,
                   int32_t* d_done, double* d_delta, void* stream) {
  if (!c || ncol < 1 || !d_obs || !d_stokes || !d_best_syn || !d_sla || !d_eqv || !d_eqv_best || !d_isigmap ||
      !d_atm || !d_atm_best)
    return FZ_EARG;
  if (!c->inv_set) return set_err(c, FZ_ECONFIG, "fz_set_inversion first");
  if (c->win_on && ncol != c->win_ncol) return set_err(c, FZ_EARG, "column count differs from fz_set_windows");
  HIPCHK(c, hipSetDevice(c->device));
  LmArgs a = c->lm;
  a.ncol = ncol;
  a.obs = d_obs;
  a.stokes = d_stokes;
  a.best_syn = d_best_syn;
  a.sla = d_sla;
  a.eqv = d_eqv;
  a.eqv_best = d_eqv_best;
  a.isigmap = d_isigmap;
  a.atm = d_atm;
  a.atm_best = d_atm_best;
  a.done = d_done;
  a.delta = d_delta;
  a.pix_of = nullptr;
  set_lm_windows(c, a);
  return launch_lm(c, a, ncol, (hipStream_t)stream);
}

}  // extern "C"

// ------------------------------------------------------------------ GET_PERTURBATION / HESSIAN

// doubles of device memory that run_pert needs beyond its inputs and outputs:
// the normal matrix with the gradient column, and the solver's work arrays
// when they do not fit the LDS
static size_t pert_scratch_doubles(int np) {
  size_t n = (size_t)np * (np + 1);
  if (np > LM_MAXNP) n += pert_work_doubles(np, LM_BIG_THREADS);
  return (n + 31) / 32 * 32;
}

static int pert_check(fz_ctx* c, int np, int nf, int dtype) {
  if (np < 1 || nf < 1) return set_err(c, FZ_EARG, "GET_PERTURBATION: NP and NF must be positive");
  if (np > LM_MAXNP_BIG) return set_err(c, FZ_EARG, "GET_PERTURBATION: more than 2048 parameters");
  // lm.f90:194-201: any other DTYPE stops the reference
  if (dtype != 1 && dtype != 2) return set_err(c, FZ_EARG, "LM does not support DTYPE " + std::to_string(dtype));
  return FZ_OK;
}

// the normal matrix of X (np columns at x[p*sp + f*sf]) and, with y, the
// gradient X^T y into gram[np][np+1] (upper triangle)
static int launch_gram(fz_ctx* c, int np, int nf, const double* x, long sp, long sf, const double* y, double* gram,
                       hipStream_t st) {
  GramArgs g;
  g.np = np;
  g.nf = nf;
  g.x = x;
  g.sp = sp;
  g.sf = sf;
  g.y = y;
  g.gram = gram;
  const unsigned nti = (unsigned)((np + GR_TILE - 1) / GR_TILE), ntj = (unsigned)((np + 1 + GR_TILE - 1) / GR_TILE);
  hipLaunchKernelGGL(gram_kernel, dim3(ntj, nti), dim3(GR_THREADS), 0, st, g);
  HIPCHK(c, hipGetLastError());
  return FZ_OK;
}

// GET_PERTURBATION on device arrays with the caller's scratch (at least
// pert_scratch_doubles(np) doubles)
static int run_pert(fz_ctx* c, int np, int nf, const double* d_ijac, const double* d_diff, double lambda,
                    double* d_delta, const double* d_pen, const double* d_hss, int dtype, double svdtol,
                    double* scratch, hipStream_t st) {
  int r = launch_gram(c, np, nf, d_ijac, nf, 1, d_diff, scratch, st);
  if (r) return r;
  PertArgs a;
  a.np = np;
  a.dtype = dtype;
  a.lambda = lambda;
  a.svdtol = svdtol;
  a.gram = scratch;
  a.reg_pen = d_pen;
  a.reg_hss = d_hss;
  a.delta = d_delta;
  if (np <= LM_MAXNP) {
    const size_t lds = pert_work_doubles(np, LM_THREADS) * sizeof(double);
    HIPCHK(c, hipFuncSetAttribute((const void*)pert_solve_kernel<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds));
    a.work = nullptr;
    hipLaunchKernelGGL(pert_solve_kernel<false>, dim3(1), dim3(LM_THREADS), lds, st, a);
  } else {
    a.work = scratch + (size_t)np * (np + 1);
    hipLaunchKernelGGL(pert_solve_kernel<true>, dim3(1), dim3(LM_BIG_THREADS), 0, st, a);
  }
  HIPCHK(c, hipGetLastError());
  return FZ_OK;
}

// fz_get_perturbation_dev_noalloc: the normal matrix and the solver's work
// arrays in a fixed device array instead of memory the context allocates.
// Every call on a GPU shares that GPU's copy (calls that overlap on one GPU
// race); each call writes every entry it reads, so nothing is cleared.
constexpr int NOALLOC_MAXNP = 512;
constexpr size_t NOALLOC_DOUBLES = (size_t)NOALLOC_MAXNP * (NOALLOC_MAXNP + 1) +
                                   2 * (size_t)NOALLOC_MAXNP * (NOALLOC_MAXNP + 1) + 8 * (size_t)NOALLOC_MAXNP +
                                   LM_BIG_THREADS + 32;   // 6.3 MB
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];
// the address of each GPU's copy of scratch, from hipGetSymbolAddress with
// that GPU current (the array's name is a host-side stand-in)
static double* g_scratch_dev[64];

static int scratch_setup(int device) {
  static std::mutex mu;
  std::lock_guard<std::mutex> lock(mu);
  if (device < 0 || device >= 64) return FZ_EARG;
  if (g_scratch_dev[device]) return FZ_OK;
  void* ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return FZ_EHIP;
  g_scratch_dev[device] = (double*)ptr;
  return FZ_OK;
}

extern "C" {

int fz_get_perturbation_dev_noalloc(fz_ctx* c, int np, int nf, const double* d_ijacobian, const double* d_diff,
                                    double lambdap, double* d_delta, const double* d_reg_pen,
                                    const double* d_reg_hss, int dtype, double svdtol, void* stream) {
  if (!c || !d_ijacobian || !d_diff || !d_delta) return FZ_EARG;
  int r = pert_check(c, np, nf, dtype);
  if (r) return r;
  if (np > NOALLOC_MAXNP)
    return set_err(c, FZ_EARG, "fz_get_perturbation_dev_noalloc: at most 512 parameters (fixed scratch)");
  static_assert(NOALLOC_DOUBLES >= (size_t)NOALLOC_MAXNP * (NOALLOC_MAXNP + 1) + 2 * (size_t)NOALLOC_MAXNP *
                                        (NOALLOC_MAXNP + 1) + 8 * (size_t)NOALLOC_MAXNP + LM_BIG_THREADS,
                "scratch holds pert_scratch_doubles(NOALLOC_MAXNP)");
  if (pert_scratch_doubles(np) > NOALLOC_DOUBLES) return set_err(c, FZ_EARG, "scratch too small");
  double* ptr = g_scratch_dev[c->device];
  if (!ptr) return set_err(c, FZ_EHIP, "scratch address of this device not set");
  HIPCHK(c, hipSetDevice(c->device));
  return run_pert(c, np, nf, d_ijacobian, d_diff, lambdap, d_delta, d_reg_pen, d_reg_hss, dtype, svdtol, ptr,
                  (hipStream_t)stream);
}

int fz_get_perturbation_dev(fz_ctx* c, int np, int nf, const double* d_ijacobian, const double* d_diff,
                            double lambdap, double* d_delta, const double* d_reg_pen, const double* d_reg_hss,
                            int dtype, double svdtol, void* stream) {
  if (!c || !d_ijacobian || !d_diff || !d_delta) return FZ_EARG;
  int r = pert_check(c, np, nf, dtype);
  if (r) return r;
  HIPCHK(c, hipSetDevice(c->device));
  r = ensure(c, &c->d_pert, &c->pert_cap, sizeof(double) * pert_scratch_doubles(np));
  if (r) return r;
  return run_pert(c, np, nf, d_ijacobian, d_diff, lambdap, d_delta, d_reg_pen, d_reg_hss, dtype, svdtol,
                  (double*)c->d_pert, (hipStream_t)stream);
}

int fz_get_perturbation(fz_ctx* c, int np, int nf, const double* ijacobian, const double* diff, double lambdap,
                        double* delta, const double* reg_pen, const double* reg_hss, int dtype, double svdtol) {
  if (!c || !ijacobian || !diff || !delta) return FZ_EARG;
  int r = pert_check(c, np, nf, dtype);
  if (r) return r;
  HIPCHK(c, hipSetDevice(c->device));
  const size_t nj = (size_t)np * nf, nh = reg_hss ? (size_t)np * np : 0, npn = reg_pen ? (size_t)np : 0;
  r = ensure(c, &c->d_pstage, &c->pstage_cap, sizeof(double) * (nj + nf + npn + nh + np));
  if (r) return r;
  r = ensure(c, &c->d_pert, &c->pert_cap, sizeof(double) * pert_scratch_doubles(np));
  if (r) return r;
  double* dj = (double*)c->d_pstage;
  double* dd = dj + nj;
  double* dp = dd + nf;
  double* dh = dp + npn;
  double* dl = dh + nh;
  HIPCHK(c, hipMemcpy(dj, ijacobian, sizeof(double) * nj, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(dd, diff, sizeof(double) * nf, hipMemcpyHostToDevice));
  if (npn) HIPCHK(c, hipMemcpy(dp, reg_pen, sizeof(double) * npn, hipMemcpyHostToDevice));
  if (nh) HIPCHK(c, hipMemcpy(dh, reg_hss, sizeof(double) * nh, hipMemcpyHostToDevice));
  r = run_pert(c, np, nf, dj, dd, lambdap, dl, npn ? dp : nullptr, nh ? dh : nullptr, dtype, svdtol,
               (double*)c->d_pert, nullptr);
  if (r) return r;
  HIPCHK(c, hipMemcpy(delta, dl, sizeof(double) * np, hipMemcpyDeviceToHost));
  return FZ_OK;
}

int fz_hessian(fz_ctx* c, int np, int nf, const double* jacobian, double* hss) {
  if (!c || !jacobian || !hss) return FZ_EARG;
  if (np < 1 || nf < 1 || np > LM_MAXNP_BIG) return set_err(c, FZ_EARG, "HESSIAN: NP in 1..2048 and NF >= 1");
  HIPCHK(c, hipSetDevice(c->device));
  const size_t nj = (size_t)np * nf, ng = (size_t)np * (np + 1);
  int r = ensure(c, &c->d_pstage, &c->pstage_cap, sizeof(double) * (nj + ng));
  if (r) return r;
  double* dj = (double*)c->d_pstage;
  double* dg = dj + nj;
  HIPCHK(c, hipMemcpy(dj, jacobian, sizeof(double) * nj, hipMemcpyHostToDevice));
  // JACOBIAN(NP,NF): parameter p of spectral row f at jacobian[f*np + p]
  r = launch_gram(c, np, nf, dj, 1, np, nullptr, dg, nullptr);
  if (r) return r;
  std::vector<double> g(ng);
  HIPCHK(c, hipMemcpy(g.data(), dg, sizeof(double) * ng, hipMemcpyDeviceToHost));
  for (int j = 0; j < np; ++j)
    for (int i = 0; i < np; ++i) hss[(size_t)j * np + i] = g[(size_t)std::min(i, j) * (np + 1) + std::max(i, j)];
  return FZ_OK;
}

}  // extern "C"

extern "C" {

int fz_get_dmodel3ds(fz_ctx* c, int nx, int ny, const double* obs3d, const double* syn3d, double* best_syn,
                     double* inv_mas, int ld_inv_mas, const float* dsyn, float* best_dsyn, const double* isigmap3d,
                     float* const* model, float* const* best_model, int32_t* am_i_done) {
  if (!c || nx < 1 || ny < 1 || !obs3d || !syn3d || !best_syn || !inv_mas || ld_inv_mas < 9 || !dsyn ||
      !best_dsyn || !isigmap3d || !model || !best_model || !am_i_done)
    return FZ_EARG;
  for (int p = 0; p < NATM; ++p)
    if (!model[p] || !best_model[p]) return FZ_EARG;
  if (!c->inv_set) return set_err(c, FZ_ECONFIG, "fz_set_inversion first");
  if (c->win_on) return set_err(c, FZ_EARG, "per-pixel windows: use fz_lm_step_dev");
  const int nz = c->cfg.nz, nf = c->lm.nfreq, np = c->lm.np;
  // pixels with AM_I_DONE == 0, in the reference's (y fastest) order
  std::vector<int> cols;
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      if (am_i_done[y + (size_t)ny * x] == 0) cols.push_back(y + ny * x);
  const int n = (int)cols.size();
  if (n == 0) return FZ_OK;
  // one weight vector when ISIGMAP3D is the same in every pixel
  bool shared = true;
  for (int i = 1; i < n && shared; ++i)
    shared = memcmp(isigmap3d + (size_t)cols[i] * nf, isigmap3d + (size_t)cols[0] * nf, sizeof(double) * nf) == 0;
  const size_t nfn = (size_t)nf * n, ne = (size_t)np * nf;
  DevBuf b_obs, b_syn, b_bsyn, b_sla, b_eqv, b_eqvb, b_isig, b_atm, b_best, b_tmp, b_done;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMalloc(&b_obs.p, sizeof(double) * nfn));
  HIPCHK(c, hipMalloc(&b_syn.p, sizeof(double) * nfn));
  HIPCHK(c, hipMalloc(&b_bsyn.p, sizeof(double) * nfn));
  HIPCHK(c, hipMalloc(&b_sla.p, sizeof(double) * 9 * n));
  HIPCHK(c, hipMalloc(&b_eqv.p, sizeof(float) * ne * n));
  HIPCHK(c, hipMalloc(&b_eqvb.p, sizeof(float) * ne * n));
  HIPCHK(c, hipMalloc(&b_isig.p, sizeof(double) * (shared ? (size_t)nf : nfn)));
  HIPCHK(c, hipMalloc(&b_atm.p, sizeof(float) * NATM * nz * n));
  HIPCHK(c, hipMalloc(&b_best.p, sizeof(float) * NATM * nz * n));
  HIPCHK(c, hipMalloc(&b_tmp.p, sizeof(float) * nz * n));
  HIPCHK(c, hipMalloc(&b_done.p, sizeof(int32_t) * n));
  // gather the active pixels' rows (GET_DMODEL3DS sends the same arrays per pixel)
  auto gather_d = [&](const double* src, size_t len, void* dst) -> int {
    std::vector<double> h(len * n);
    for (int i = 0; i < n; ++i) memcpy(&h[(size_t)i * len], src + (size_t)cols[i] * len, sizeof(double) * len);
    HIPCHK(c, hipMemcpy(dst, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    return FZ_OK;
  };
  auto gather_f = [&](const float* src, size_t len, void* dst) -> int {
    std::vector<float> h(len * n);
    for (int i = 0; i < n; ++i) memcpy(&h[(size_t)i * len], src + (size_t)cols[i] * len, sizeof(float) * len);
    HIPCHK(c, hipMemcpy(dst, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
    return FZ_OK;
  };
  int r;
  if ((r = gather_d(obs3d, nf, b_obs.p)) || (r = gather_d(syn3d, nf, b_syn.p)) ||
      (r = gather_d(best_syn, nf, b_bsyn.p)) || (r = gather_f(dsyn, ne, b_eqv.p)) ||
      (r = gather_f(best_dsyn, ne, b_eqvb.p)))
    return r;
  if (shared) HIPCHK(c, hipMemcpy(b_isig.p, isigmap3d + (size_t)cols[0] * nf, sizeof(double) * nf, hipMemcpyHostToDevice));
  else if ((r = gather_d(isigmap3d, nf, b_isig.p))) return r;
  {
    std::vector<double> h((size_t)9 * n);
    for (int i = 0; i < n; ++i) memcpy(&h[(size_t)i * 9], inv_mas + (size_t)cols[i] * ld_inv_mas, sizeof(double) * 9);
    HIPCHK(c, hipMemcpy(b_sla.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
  }
  HIPCHK(c, hipMemset(b_done.p, 0, sizeof(int32_t) * n));
  const unsigned tb = 256, tg = (unsigned)(((long)nz * n + tb - 1) / tb);
  const size_t colbytes = sizeof(float) * nz;
  std::vector<float> h((size_t)nz * n);
  float* d_tmp = (float*)b_tmp.p;
  for (int which = 0; which < 2; ++which) {
    float* const* cubes = which ? best_model : model;
    float* d = (float*)(which ? b_best.p : b_atm.p);
    for (int p = 0; p < NATM; ++p) {
      for (int i = 0; i < n; ++i) memcpy(&h[(size_t)i * nz], cubes[p] + (size_t)cols[i] * nz, colbytes);
      HIPCHK(c, hipMemcpy(d_tmp, h.data(), colbytes * n, hipMemcpyHostToDevice));
      hipLaunchKernelGGL(cube_to_soa, dim3(tg), dim3(tb), 0, 0, d_tmp, d + (size_t)p * nz * n, nz, n);
    }
  }
  LmArgs a = c->lm;
  a.ncol = n;
  a.obs = (const double*)b_obs.p;
  a.syn_d = (const double*)b_syn.p;
  a.stokes = (const float*)b_syn.p;   // not read: the FP64 synthesis replaces it
  a.best_syn = (double*)b_bsyn.p;
  a.sla = (double*)b_sla.p;
  a.eqv = (const float*)b_eqv.p;
  a.eqv_best = (float*)b_eqvb.p;
  a.isigmap = (const double*)b_isig.p;
  a.isig_stride = shared ? 0 : nf;
  a.atm = (float*)b_atm.p;
  a.atm_best = (float*)b_best.p;
  a.done = (int32_t*)b_done.p;
  a.delta = nullptr;
  a.pix_of = nullptr;
  set_lm_windows(c, a);
  r = launch_lm(c, a, n, nullptr);
  if (r) return r;
  HIPCHK(c, hipDeviceSynchronize());
  // scatter back: models, INV_MAS(1:9), BEST_DSYN, BEST_SYN, AM_I_DONE
  for (int which = 0; which < 2; ++which) {
    float* const* cubes = which ? best_model : model;
    const float* d = (const float*)(which ? b_best.p : b_atm.p);
    for (int p = 0; p < NATM; ++p) {
      hipLaunchKernelGGL(soa_to_cube, dim3(tg), dim3(tb), 0, 0, d + (size_t)p * nz * n, d_tmp, nz, n);
      HIPCHK(c, hipMemcpy(h.data(), d_tmp, colbytes * n, hipMemcpyDeviceToHost));
      for (int i = 0; i < n; ++i) memcpy(cubes[p] + (size_t)cols[i] * nz, &h[(size_t)i * nz], colbytes);
    }
  }
  {
    std::vector<double> hs((size_t)9 * n), hb(nfn);
    HIPCHK(c, hipMemcpy(hs.data(), b_sla.p, sizeof(double) * hs.size(), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(hb.data(), b_bsyn.p, sizeof(double) * nfn, hipMemcpyDeviceToHost));
    std::vector<int32_t> hd(n);
    HIPCHK(c, hipMemcpy(hd.data(), b_done.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) {
      memcpy(inv_mas + (size_t)cols[i] * ld_inv_mas, &hs[(size_t)i * 9], sizeof(double) * 9);
      memcpy(best_syn + (size_t)cols[i] * nf, &hb[(size_t)i * nf], sizeof(double) * nf);
      if (hd[i]) am_i_done[cols[i]] = 1;
    }
    std::vector<float> he(ne * n);
    HIPCHK(c, hipMemcpy(he.data(), b_eqvb.p, sizeof(float) * he.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) memcpy(best_dsyn + (size_t)cols[i] * ne, &he[(size_t)i * ne], sizeof(float) * ne);
  }
  return FZ_OK;
}

// fz_lm_step_dev over the na active pixels pix_of[0..na): stokes and eqv are
// the compacted rows of those pixels, everything else is in pixel space.
static int lm_step_active(fz_ctx* c, int ncol, int na, const int32_t* pix_of, const double* d_obs,
                          const float* d_stokes_c, double* d_best_syn, double* d_sla, const float* d_eqv_c,
                          float* d_eqv_best, const double* d_isigmap, long isig_stride, float* d_atm,
                          float* d_atm_best, int32_t* d_done, hipStream_t st) {
  LmArgs a = c->lm;
  a.ncol = ncol;
  a.obs = d_obs;
  a.stokes = d_stokes_c;
  a.best_syn = d_best_syn;
  a.sla = d_sla;
  a.eqv = d_eqv_c;
  a.eqv_best = d_eqv_best;
  a.isigmap = d_isigmap;
  a.isig_stride = isig_stride;
  a.atm = d_atm;
  a.atm_best = d_atm_best;
  a.done = d_done;
  a.delta = nullptr;
  a.pix_of = pix_of;
  set_lm_windows(c, a);
  if (c->win_on && (pix_of || ncol != c->win_ncol)) return set_err(c, FZ_EARG, "per-pixel windows: every pixel keeps its row");
  return launch_lm(c, a, na, st);
}

// GET_NJEVALS (get_dmodel.f90:375-407) with TOFFSET = 1 (pre_post_duties.f90:645)
static int njevals(int apar, int curic, int maxiter, int cycpow, int nz) {
  const int toffset = 1;
  if (apar == 1) {
    if (curic == 0 || curic == maxiter) return nz;
    return std::min(nz, (int)std::pow((double)cycpow, (double)(curic + toffset - 1)));
  }
  if (curic == 0) return nz;
  if (curic == 1) return 1;
  if (curic == maxiter) return nz;
  return std::min(nz, (int)std::pow((double)cycpow, (double)(curic - 1)));
}

// SET_TAU's assists on device data (fz_assist_dev, the first cycle of fz_invert3d_dev)
static int run_assist(fz_ctx* c, int ncol, const double* d_obs, int flags, float* d_atm, hipStream_t st) {
  const Cfg& g = c->cfg;
  AssistArgs a;
  a.ncol = ncol;
  a.nz = g.nz;
  a.nfreq = c->lm.nfreq;
  a.numw = g.numw;
  a.itz = g.reg_nwave[0];
  a.nstk = c->lm.nfreq / g.numw;
  a.flags = flags;
  a.kz = (int)std::ceil((double)(float)g.nz / 3.0) - 1;   // CEILING(REAL(NZ)/3.)
  a.lk_hc = (g.reg_l0main[0] * 1.0e-8 * KBOL) / (HPLA * LIGHT);
  a.obs = d_obs;
  a.atm = d_atm;
  DevBuf b_vec, b_pix, b_mx;
  HIPCHK(c, hipMalloc(&b_vec.p, sizeof(double) * (3 * (size_t)a.itz + 1)));
  HIPCHK(c, hipMalloc(&b_pix.p, sizeof(double) * 5 * (size_t)ncol));
  HIPCHK(c, hipMalloc(&b_mx.p, sizeof(double) * 4));
  a.vec = (double*)b_vec.p;
  a.pix = (double*)b_pix.p;
  a.mx = (double*)b_mx.p;
  const unsigned gp = (unsigned)((ncol + 255) / 256);
  if (flags & (ASSIST_B | ASSIST_V)) {
    hipLaunchKernelGGL(assist_sum_kernel, dim3((unsigned)a.itz), dim3(256), 0, st, a);
    hipLaunchKernelGGL(assist_vec_kernel, dim3(1), dim3(64), 0, st, a);
  } else {
    HIPCHK(c, hipMemsetAsync(a.vec, 0, sizeof(double) * (3 * (size_t)a.itz + 1), st));
  }
  hipLaunchKernelGGL(assist_pix_kernel, dim3(gp), dim3(256), 0, st, a);
  hipLaunchKernelGGL(assist_max_kernel, dim3(1), dim3(256), 0, st, a);
  hipLaunchKernelGGL(assist_apply_kernel, dim3(gp), dim3(256), 0, st, a);
  HIPCHK(c, hipGetLastError());
  // the scratch is released at return: wait for the kernels that read it
  HIPCHK(c, hipStreamSynchronize(st));
  return FZ_OK;
}

extern "C" int fz_assist_dev(fz_ctx* c, int ncol, const double* d_obs, int flags, float* d_atm, void* stream) {
  if (!c || ncol < 1 || !d_obs || !d_atm || (flags & ~15)) return FZ_EARG;
  if (!c->inv_set) return set_err(c, FZ_ECONFIG, "fz_set_inversion first");
  if (c->win_on) return set_err(c, FZ_EARG, "per-pixel windows: the assists read a common first region");
  if (flags == 0) return FZ_OK;
  HIPCHK(c, hipSetDevice(c->device));
  return run_assist(c, ncol, d_obs, flags, d_atm, (hipStream_t)stream);
}

int fz_invert3d_dev(fz_ctx* c, int nx, int ny, float* d_atm, const double* d_obs, const fz_invert_control* ctl,
                    float* d_stokes, double* d_chi2, int* iterations, void* stream) {
  if (!c || !c->configured || nx < 1 || ny < 1 || !d_atm || !d_obs || !ctl || !d_stokes) return FZ_EARG;
  const int nz = c->cfg.nz, numw = c->cfg.numw, ncol = nx * ny;
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(c, hipSetDevice(c->device));
  // MAXITER (pre_post_duties.f90:635-646) from the parameters' maximum nodes
  int nfreev = 0, nfreep = 0, infreepmax = 0;
  for (int p = 0; p < 8; ++p) {
    if (!ctl->atmpar[p]) continue;
    const int vmax = ctl->nslb_max[p] > 0 ? std::min(ctl->nslb_max[p], nz) : nz;
    ++nfreev;
    nfreep += vmax;
    infreepmax = std::max(infreepmax, vmax);
  }
  if (nfreev == 0) return set_err(c, FZ_ECONFIG, "no parameter inverted");
  const int cycpow = ctl->cycpow > 0 ? ctl->cycpow : 2;
  const int maxiter = (nfreep == nfreev) ? 1 : (int)std::floor(std::log((double)infreepmax) / std::log((double)cycpow));
  if (maxiter < 1) return set_err(c, FZ_ECONFIG, "no inversion cycle");
  const int maxsteps = ctl->maxsteps > 0 ? ctl->maxsteps : 5;
  const int itcap = ctl->max_iterations > 0 ? ctl->max_iterations : 1000;
  // Stokes weights WSTK / MINW in REAL(4) (input_file.f90:1298-1303; weights of
  // Stokes parameters not inverted are zero) and ISIGMAP = WSTK^2 / ISIGMA^2
  // with ISIGMA a REAL(8) taken from a REAL(4) value (input_file.f90:181,433)
  int nstk = 0;
  float wsum = 0.f, wst[4];
  for (int s = 0; s < 4; ++s) {
    wst[s] = ctl->invstk[s] ? ctl->wstk[s] : 0.f;
    if (ctl->invstk[s]) ++nstk;
    wsum += wst[s] * wst[s];
  }
  if (nstk == 0) return set_err(c, FZ_ECONFIG, "no Stokes parameter inverted");
  const float minw = sqrtf(wsum) / sqrtf((float)nstk);
  const double isigma = (double)(float)(ctl->noise > 0 ? ctl->noise : 4.0e-3);
  const int nfreq = nstk * numw;
  // per-pixel windows: ragged rows of nstk*numw_pix (obs, best synthesis,
  // response functions), Stokes of 4*numw_pix; every pixel keeps its row
  const bool win = c->win_on;
  if (win && ncol != c->win_ncol) return set_err(c, FZ_EARG, "column count differs from fz_set_windows");
  const size_t nwtot = win ? (size_t)c->h_poff[ncol] : (size_t)ncol * numw;
  std::vector<double> isig(nfreq);
  {
    int b = 0;
    for (int s = 0; s < 4; ++s) {
      if (!ctl->invstk[s]) continue;
      const float w = wst[s] / minw;
      for (int l = 0; l < numw; ++l) isig[b * numw + l] = (double)(w * w) / (isigma * isigma);
      ++b;
    }
  }
  int npmax = 0;
  for (int p = 0; p < 8; ++p)
    if (ctl->atmpar[p]) npmax += ctl->nslb_max[p] > 0 ? std::min(ctl->nslb_max[p], nz) : nz;
  if (npmax > LM_MAXNP_BIG) return set_err(c, FZ_ECONFIG, "more than 2048 free parameters in the last cycle");
  DevBuf b_isig, b_best, b_eqv, b_eqvb, b_bsyn, b_sla, b_done, b_stat, b_t0, b_t1, b_atmc, b_idx, b_cnt;
  const size_t natm = (size_t)12 * nz * ncol;
  // STOKES SETUP: AUTO replaces the map-wide weights by every pixel's own
  // (SET_WEIGHTS, get_dmodel.f90:2024-2047), set once from the observations
  const long isig_stride = c->autow ? nfreq : 0;
  if (c->autow && win) return set_err(c, FZ_ECONFIG, "automatic Stokes weights take the dense spectral grid");
  HIPCHK(c, hipMalloc(&b_isig.p, sizeof(double) * nfreq * (c->autow ? (size_t)ncol : 1)));
  HIPCHK(c, hipMalloc(&b_best.p, sizeof(float) * natm));
  HIPCHK(c, hipMalloc(&b_eqv.p, sizeof(float) * nwtot * npmax * nstk));
  HIPCHK(c, hipMalloc(&b_eqvb.p, sizeof(float) * nwtot * npmax * nstk));
  HIPCHK(c, hipMalloc(&b_bsyn.p, sizeof(double) * nwtot * nstk));
  HIPCHK(c, hipMalloc(&b_sla.p, sizeof(double) * (size_t)ncol * 9));
  HIPCHK(c, hipMalloc(&b_done.p, sizeof(int32_t) * ncol));
  HIPCHK(c, hipMalloc(&b_stat.p, sizeof(double) * 2));
  HIPCHK(c, hipMalloc(&b_t0.p, sizeof(double) * (size_t)nz * ncol));
  HIPCHK(c, hipMalloc(&b_t1.p, sizeof(double) * (size_t)nz * ncol));
  // active pixels: FORWARD3D models only the columns with AM_I_DONE == 0
  // (forward.f90:43-306); they are gathered into a compact cube
  HIPCHK(c, hipMalloc(&b_atmc.p, sizeof(float) * natm));
  HIPCHK(c, hipMalloc(&b_idx.p, sizeof(int32_t) * ncol));
  HIPCHK(c, hipMalloc(&b_cnt.p, sizeof(int32_t)));
  if (c->autow) {
    int r = fz_auto_weights_dev(c, ncol, d_obs, ctl->invstk, ctl->noise, (double*)b_isig.p, stream);
    if (r) return r;
  } else {
    HIPCHK(c, hipMemcpyAsync(b_isig.p, isig.data(), sizeof(double) * nfreq, hipMemcpyHostToDevice, st));
  }
  float* best = (float*)b_best.p;
  double* sla = (double*)b_sla.p;
  int32_t* done = (int32_t*)b_done.p;
  const unsigned gatm = (unsigned)((natm + 255) / 256), grow = (unsigned)(((size_t)nz * ncol + 255) / 256);
  static const int rowof[8] = {A_T, A_PG, A_RHO, A_BX, A_BY, A_BZ, A_VZ, A_PG};
  int total_it = 0;
  const int cyc0 = std::min(std::max(1, ctl->first_cycle), maxiter);
  for (int cyc = cyc0; cyc <= maxiter; ++cyc) {
    // INITIALIZE_CYCLE (invert3d.f90:228-325)
    fz_inversion inv;
    std::memset(&inv, 0, sizeof(inv));
    for (int p = 0; p < 8; ++p) {
      if (!ctl->atmpar[p]) continue;
      const int vmax = ctl->nslb_max[p] > 0 ? std::min(ctl->nslb_max[p], nz) : nz;
      inv.atmpar[p] = 1;
      inv.nodes[p] = std::min(njevals(p + 1, cyc, maxiter, cycpow, nz), vmax);
      inv.pen_typ[p] = ctl->pen_typ[p];
      inv.pen_alp[p] = ctl->pen_alp[p];
    }
    for (int s = 0; s < 4; ++s) inv.invstk[s] = ctl->invstk[s];
    inv.regul = ctl->regul;
    inv.solver = ctl->solver > 0 ? ctl->solver : 1;
    inv.node_expansion = ctl->node_expansion > 0 ? ctl->node_expansion : 1;
    inv.maxsteps = maxsteps;
    inv.svdtol = ctl->svdtol;
    int r = fz_set_inversion(c, &inv, nullptr, nullptr);
    if (r) return r;
    hipLaunchKernelGGL(inv_init_kernel, dim3((ncol + 255) / 256), dim3(256), 0, st, ncol, sla, done);
    // SET_TAU (:492-727): a synthesis updates log tau5 when Pg and P0 are not
    // inverted; in the first cycle the assist initialisers change the start
    // model before it
    if (!ctl->atmpar[1] && !ctl->atmpar[7]) {
      if (cyc == cyc0 && ctl->assist && !c->win_on) {
        r = run_assist(c, ncol, d_obs, ctl->assist & 15, d_atm, st);
        if (r) return r;
      }
      r = fz_forward_dev(c, ncol, d_atm, FZ_RF_NONE, d_stokes, nullptr, nullptr, stream);
      if (r) return r;
    }
    // UPDATE_CURRENT_TO_BEST, then SMOOTHING_B on the best model into both
    hipLaunchKernelGGL(copy_f32_kernel, dim3(gatm), dim3(256), 0, st, d_atm, best, (long)natm);
    if (ctl->smoothing) {
      for (int p = 0; p < 8; ++p) {
        if (!ctl->atmpar[p]) continue;
        float* rb = best + (size_t)rowof[p] * nz * ncol;
        float* rc = d_atm + (size_t)rowof[p] * nz * ncol;
        double *t0 = (double*)b_t0.p, *t1 = (double*)b_t1.p;
        hipLaunchKernelGGL(to_double_kernel, dim3(grow), dim3(256), 0, st, rb, t0, (long)nz * ncol);
        const int ext[3] = {nx, ny, nz};
        for (int ax = 0; ax < 3; ++ax) {
          if (ext[ax] <= 4) continue;
          hipLaunchKernelGGL(smooth_axis_kernel, dim3(grow), dim3(256), 0, st, t0, t1, nz, ny, nx, ax);
          std::swap(t0, t1);
        }
        hipLaunchKernelGGL(from_double_kernel, dim3(grow), dim3(256), 0, st, t0, rc, rb, (long)nz * ncol);
      }
    }
    HIPCHK(c, hipGetLastError());
    const double need = (double)std::max(1, maxsteps - (maxiter - cyc));
    int na = ncol;   // active pixels (every pixel at a cycle start)
    for (int it = 0; it < itcap; ++it) {
      // MODELLING -> SOLVE_DMODEL -> MAKE_DECISION
      if (na == ncol) {
        r = fz_forward_nodes_dev(c, ncol, d_atm, d_stokes, (float*)b_eqv.p, nullptr, stream);
        if (r) return r;
        r = lm_step_active(c, ncol, ncol, nullptr, d_obs, d_stokes, (double*)b_bsyn.p, sla, (const float*)b_eqv.p,
                           (float*)b_eqvb.p, (const double*)b_isig.p, isig_stride, d_atm, best, done, st);
        if (r) return r;
      } else {
        const int32_t* idx = (const int32_t*)b_idx.p;
        float* atmc = (float*)b_atmc.p;
        const unsigned gc = (unsigned)(((long)12 * nz * na + 255) / 256);
        hipLaunchKernelGGL(gather_cols_kernel, dim3(gc), dim3(256), 0, st, (const float*)d_atm, atmc, idx, na, ncol,
                           12 * nz);
        r = fz_forward_nodes_dev(c, na, atmc, d_stokes, (float*)b_eqv.p, nullptr, stream);
        if (r) return r;
        hipLaunchKernelGGL(scatter_cols_kernel, dim3(gc), dim3(256), 0, st, (const float*)atmc, d_atm, idx, na,
                           ncol, 12 * nz);
        r = lm_step_active(c, ncol, na, idx, d_obs, d_stokes, (double*)b_bsyn.p, sla, (const float*)b_eqv.p,
                           (float*)b_eqvb.p, (const double*)b_isig.p, isig_stride, d_atm, best, done, st);
        if (r) return r;
      }
      hipLaunchKernelGGL(stats_kernel, dim3(1), dim3(256), 0, st, ncol, sla, done, (double*)b_stat.p);
      double hs[2];
      HIPCHK(c, hipMemcpyAsync(hs, b_stat.p, sizeof(hs), hipMemcpyDeviceToHost, st));
      HIPCHK(c, hipStreamSynchronize(st));
      ++total_it;
      if (ncol - (int)hs[1] <= 0) break;
      if (hs[0] / (double)ncol >= need) break;
      const int nact = ncol - (int)hs[1];
      if (nact < ncol && !win) {
        hipLaunchKernelGGL(compact_kernel, dim3(1), dim3(1024), 0, st, ncol, (const int32_t*)done, (int32_t*)b_idx.p,
                           (int32_t*)b_cnt.p);
        HIPCHK(c, hipGetLastError());
      }
      na = win ? ncol : nact;
    }
    // END_CYCLE (:328-403): current := best, synthesis without response functions
    hipLaunchKernelGGL(copy_f32_kernel, dim3(gatm), dim3(256), 0, st, best, d_atm, (long)natm);
    r = fz_forward_dev(c, ncol, d_atm, FZ_RF_NONE, d_stokes, nullptr, nullptr, stream);
    if (r) return r;
  }
  if (d_chi2) {
    HIPCHK(c, hipMemcpy2DAsync(d_chi2, sizeof(double), sla + 2, sizeof(double) * 9, sizeof(double), ncol,
                               hipMemcpyDeviceToDevice, st));
  }
  HIPCHK(c, hipStreamSynchronize(st));
  if (iterations) *iterations = total_it;
  return FZ_OK;
}

// INVERT3D with the spatial PSF coupling the pixels (the coupled inversion
// whose buffers coupled_inversion.f90:10-207 allocates; the reference does
// not wire its mathematics, get_dmodel.f90:1707-1711).  The model of every
// pixel is fitted to the observations through the PSF: chi^2 of the map =
// sum over pixels of the weighted squared residual obs - PSF * syn.  A
// Levenberg-Marquardt iteration over the whole map:
//   * forward with node-contracted response functions of the current model,
//     PSF convolution of the synthesis, residual and chi^2 of the map;
//   * acceptance by the map's chi^2 (pixels are coupled, so one pixel's
//     chi^2 does not decide its own step), lambda divided by sqrt(10) on an
//     accepted step and multiplied by 10 on a rejected one (the rules of
//     UPDATE_LAMBDA, get_dmodel.f90:917-956, applied to the map), the best
//     model, response functions and residual kept;
//   * the exact gradient of the map's chi^2 for each pixel, J_q^T W (PSF^T r)_q
//     with the adjoint convolution of the best residual, and the pixel's own
//     block of the coupled Gauss-Newton matrix, (sum PSF^2) J_q^T W J_q; the
//     damped solve and the node-basis update of lm_kernel.
// The cycle schedule, SET_TAU, the 1-2-1 smoothing and the assists are those
// of fz_invert3d_dev.  The cycle ends after MAXSTEPS rejected steps in a row
// with no net gain (the map's step count as INV_SLA(1) counts a pixel's) or
// the iteration cap.
static int coupled_forward(fz_ctx* c, int nx, int ny, float* d_atm, float* d_stokes, float* d_eqv, float* d_t1,
                           float* d_t2, hipStream_t st, bool nodes) {
  const int numw = c->cfg.numw, ncol = nx * ny;
  int r = nodes ? fz_forward_nodes_dev(c, ncol, d_atm, d_stokes, d_eqv, nullptr, st)
                : fz_forward_dev(c, ncol, d_atm, FZ_RF_NONE, d_stokes, nullptr, nullptr, st);
  if (r) return r;
  const long nch = 4L * numw;
  // [ncol][4 numw] -> [4 numw][ncol], convolve
  hipLaunchKernelGGL((transpose_kernel<float, float>), dim3((unsigned)((nch + 31) / 32), (unsigned)((ncol + 31) / 32)),
                     dim3(256), 0, st, (const float*)d_stokes, d_t1, (long)ncol, nch);
  launch_psf(c, nx, ny, (int)nch, d_t1, d_t2, 0, st);
  HIPCHK(c, hipGetLastError());
  return FZ_OK;
}

extern "C" {

int fz_coupled_invert3d_dev(fz_ctx* c, int nx, int ny, float* d_atm, const double* d_obs,
                            const fz_invert_control* ctl, float* d_stokes, double* d_chi2, int* iterations,
                            void* stream) {
  if (!c || !c->configured || nx < 1 || ny < 1 || !d_atm || !d_obs || !ctl || !d_stokes) return FZ_EARG;
  if (c->win_on) return set_err(c, FZ_ECONFIG, "the coupled inversion takes the dense spectral grid");
  if (c->autow) return set_err(c, FZ_ECONFIG, "the coupled inversion takes the map-wide Stokes weights");
  // the map transposes put 32 pixels per workgroup row on the grid's y axis
  if ((long)nx * ny > 32L * 65535) return set_err(c, FZ_EARG, "more than 2097120 pixels in one coupled map");
  const int nz = c->cfg.nz, numw = c->cfg.numw, ncol = nx * ny;
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(c, hipSetDevice(c->device));
  int r = psf_taps(c, nx, ny);
  if (r) return r;
  int nfreev = 0, nfreep = 0, infreepmax = 0;
  for (int p = 0; p < 8; ++p) {
    if (!ctl->atmpar[p]) continue;
    const int vmax = ctl->nslb_max[p] > 0 ? std::min(ctl->nslb_max[p], nz) : nz;
    ++nfreev;
    nfreep += vmax;
    infreepmax = std::max(infreepmax, vmax);
  }
  if (nfreev == 0) return set_err(c, FZ_ECONFIG, "no parameter inverted");
  const int cycpow = ctl->cycpow > 0 ? ctl->cycpow : 2;
  const int maxiter = (nfreep == nfreev) ? 1 : (int)std::floor(std::log((double)infreepmax) / std::log((double)cycpow));
  if (maxiter < 1) return set_err(c, FZ_ECONFIG, "no inversion cycle");
  const int maxsteps = ctl->maxsteps > 0 ? ctl->maxsteps : 5;
  const int itcap = ctl->max_iterations > 0 ? ctl->max_iterations : 1000;
  int nstk = 0;
  float wsum = 0.f, wst[4];
  for (int s = 0; s < 4; ++s) {
    wst[s] = ctl->invstk[s] ? ctl->wstk[s] : 0.f;
    if (ctl->invstk[s]) ++nstk;
    wsum += wst[s] * wst[s];
  }
  if (nstk == 0) return set_err(c, FZ_ECONFIG, "no Stokes parameter inverted");
  const float minw = sqrtf(wsum) / sqrtf((float)nstk);
  const double isigma = (double)(float)(ctl->noise > 0 ? ctl->noise : 4.0e-3);
  const int nfreq = nstk * numw;
  std::vector<double> isig(nfreq);
  {
    int b = 0;
    for (int s = 0; s < 4; ++s) {
      if (!ctl->invstk[s]) continue;
      const float w = wst[s] / minw;
      for (int l = 0; l < numw; ++l) isig[b * numw + l] = (double)(w * w) / (isigma * isigma);
      ++b;
    }
  }
  int npmax = 0;
  for (int p = 0; p < 8; ++p)
    if (ctl->atmpar[p]) npmax += ctl->nslb_max[p] > 0 ? std::min(ctl->nslb_max[p], nz) : nz;
  if (npmax > LM_MAXNP_BIG) return set_err(c, FZ_ECONFIG, "more than 2048 free parameters in the last cycle");
  const size_t natm = (size_t)12 * nz * ncol, nrow = (size_t)ncol * nfreq, nmap = (size_t)4 * numw * ncol;
  DevBuf b_isig, b_best, b_eqv, b_eqvb, b_syn, b_res, b_resb, b_gres, b_chi, b_sum, b_t1, b_t2, b_s0, b_s1;
  HIPCHK(c, hipMalloc(&b_isig.p, sizeof(double) * nfreq));
  HIPCHK(c, hipMalloc(&b_best.p, sizeof(float) * natm));
  HIPCHK(c, hipMalloc(&b_eqv.p, sizeof(float) * nrow * npmax));
  HIPCHK(c, hipMalloc(&b_eqvb.p, sizeof(float) * nrow * npmax));
  HIPCHK(c, hipMalloc(&b_syn.p, sizeof(double) * nrow));
  HIPCHK(c, hipMalloc(&b_res.p, sizeof(double) * nrow));
  HIPCHK(c, hipMalloc(&b_resb.p, sizeof(double) * nrow));
  HIPCHK(c, hipMalloc(&b_gres.p, sizeof(double) * nrow));
  HIPCHK(c, hipMalloc(&b_chi.p, sizeof(double) * ncol));
  HIPCHK(c, hipMalloc(&b_sum.p, sizeof(double)));
  HIPCHK(c, hipMalloc(&b_t1.p, sizeof(float) * std::max(nmap, nrow)));
  HIPCHK(c, hipMalloc(&b_t2.p, sizeof(float) * std::max(nmap, nrow)));
  HIPCHK(c, hipMalloc(&b_s0.p, sizeof(double) * (size_t)nz * ncol));
  HIPCHK(c, hipMalloc(&b_s1.p, sizeof(double) * (size_t)nz * ncol));
  HIPCHK(c, hipMemcpyAsync(b_isig.p, isig.data(), sizeof(double) * nfreq, hipMemcpyHostToDevice, st));
  float* best = (float*)b_best.p;
  float *eqv = (float*)b_eqv.p, *eqvb = (float*)b_eqvb.p;
  float *t1 = (float*)b_t1.p, *t2 = (float*)b_t2.p;
  double *res = (double*)b_res.p, *resb = (double*)b_resb.p, *gres = (double*)b_gres.p;
  const unsigned gatm = (unsigned)((natm + 255) / 256), grow = (unsigned)(((size_t)nz * ncol + 255) / 256);
  static const int rowof[8] = {A_T, A_PG, A_RHO, A_BX, A_BY, A_BZ, A_VZ, A_PG};
  CoupledArgs ca;
  ca.npix = ncol;
  ca.nfreq = nfreq;
  ca.numw = numw;
  ca.nstk = nstk;
  ca.conv = t2;
  ca.obs = d_obs;
  ca.isig = (const double*)b_isig.p;
  ca.syn_c = (double*)b_syn.p;
  ca.chi = (double*)b_chi.p;
  // the map's chi^2 of the current model (forward, convolution, residual)
  auto map_chi2 = [&](bool nodes, double* rdst, double& chi) -> int {
    int rr = coupled_forward(c, nx, ny, d_atm, d_stokes, eqv, t1, t2, st, nodes);
    if (rr) return rr;
    ca.res = rdst;
    for (int b = 0; b < 4; ++b) ca.rowstk[b] = c->lm.rowstk[b];
    ca.inu = (double)nfreq;
    hipLaunchKernelGGL(coupled_resid_kernel, dim3((unsigned)((ncol + 255) / 256)), dim3(256), 0, st, ca);
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, st, (const double*)ca.chi, ncol, (double*)b_sum.p);
    HIPCHK(c, hipMemcpyAsync(&chi, b_sum.p, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    chi /= (double)ncol;
    return FZ_OK;
  };
  int total_it = 0;
  const int cyc0 = std::min(std::max(1, ctl->first_cycle), maxiter);
  for (int cyc = cyc0; cyc <= maxiter; ++cyc) {
    fz_inversion inv;
    std::memset(&inv, 0, sizeof(inv));
    for (int p = 0; p < 8; ++p) {
      if (!ctl->atmpar[p]) continue;
      const int vmax = ctl->nslb_max[p] > 0 ? std::min(ctl->nslb_max[p], nz) : nz;
      inv.atmpar[p] = 1;
      inv.nodes[p] = std::min(njevals(p + 1, cyc, maxiter, cycpow, nz), vmax);
      inv.pen_typ[p] = ctl->pen_typ[p];
      inv.pen_alp[p] = ctl->pen_alp[p];
    }
    for (int s = 0; s < 4; ++s) inv.invstk[s] = ctl->invstk[s];
    inv.regul = ctl->regul;
    inv.solver = ctl->solver > 0 ? ctl->solver : 1;
    inv.node_expansion = ctl->node_expansion > 0 ? ctl->node_expansion : 1;
    inv.maxsteps = maxsteps;
    inv.svdtol = ctl->svdtol;
    r = fz_set_inversion(c, &inv, nullptr, nullptr);
    if (r) return r;
    // SET_TAU with its assists, then the smoothing of the start model
    if (!ctl->atmpar[1] && !ctl->atmpar[7]) {
      if (cyc == cyc0 && ctl->assist) {
        r = run_assist(c, ncol, d_obs, ctl->assist & 15, d_atm, st);
        if (r) return r;
      }
      r = fz_forward_dev(c, ncol, d_atm, FZ_RF_NONE, d_stokes, nullptr, nullptr, stream);
      if (r) return r;
    }
    if (ctl->smoothing) {
      for (int p = 0; p < 8; ++p) {
        if (!ctl->atmpar[p]) continue;
        float* rc = d_atm + (size_t)rowof[p] * nz * ncol;
        double *s0 = (double*)b_s0.p, *s1 = (double*)b_s1.p;
        hipLaunchKernelGGL(to_double_kernel, dim3(grow), dim3(256), 0, st, rc, s0, (long)nz * ncol);
        const int ext[3] = {nx, ny, nz};
        for (int ax = 0; ax < 3; ++ax) {
          if (ext[ax] <= 4) continue;
          hipLaunchKernelGGL(smooth_axis_kernel, dim3(grow), dim3(256), 0, st, s0, s1, nz, ny, nx, ax);
          std::swap(s0, s1);
        }
        hipLaunchKernelGGL(from_double_kernel, dim3(grow), dim3(256), 0, st, s0, rc, rc, (long)nz * ncol);
      }
    }
    HIPCHK(c, hipGetLastError());
    const size_t neqv = nrow * (size_t)c->lm.np;
    double chibest = 1.0e29, lam = 10.0, steps = 0.0, precyc = 0.0;
    for (int it = 0; it < itcap; ++it) {
      double chi = 0.0;
      r = map_chi2(true, res, chi);
      if (r) return r;
      ++total_it;
      // UPDATE_LAMBDA on the map's chi^2
      const bool accept = chi < chibest;
      if (accept) {
        if ((chibest - chi) > 0.01 * chibest) steps = std::max(0.0, steps - 1.0);
        else steps = steps + 1.0;
        chibest = chi;
        lam = lam / std::sqrt(10.0);
        hipLaunchKernelGGL(copy_f32_kernel, dim3(gatm), dim3(256), 0, st, (const float*)d_atm, best, (long)natm);
        HIPCHK(c, hipMemcpyAsync(eqvb, eqv, sizeof(float) * neqv, hipMemcpyDeviceToDevice, st));
        HIPCHK(c, hipMemcpyAsync(resb, res, sizeof(double) * nrow, hipMemcpyDeviceToDevice, st));
      } else {
        if (std::fabs(precyc - 1.0) < 1.0e-2) {
          precyc = 0.0;
          lam = lam * 3.0;
        } else {
          lam = lam * 10.0;
        }
        steps = steps + 1.0;
        hipLaunchKernelGGL(copy_f32_kernel, dim3(gatm), dim3(256), 0, st, (const float*)best, d_atm, (long)natm);
      }
      lam = std::max(1.0e-2, lam);
      // MAKE_DECISION (invert3d.f90:820-859) on the map's step count
      if (steps >= (double)std::max(1, maxsteps - (maxiter - cyc))) break;
      // gradient residual: the adjoint PSF convolution of the best residual
      hipLaunchKernelGGL((transpose_kernel<double, float>), dim3((unsigned)((nfreq + 31) / 32), (unsigned)((ncol + 31) / 32)),
                         dim3(256), 0, st, (const double*)resb, t1, (long)ncol, (long)nfreq);
      launch_psf(c, nx, ny, nfreq, t1, t2, 1, st);
      hipLaunchKernelGGL((transpose_kernel<float, double>), dim3((unsigned)((ncol + 31) / 32), (unsigned)((nfreq + 31) / 32)),
                         dim3(256), 0, st, (const float*)t2, gres, (long)nfreq, (long)ncol);
      HIPCHK(c, hipGetLastError());
      LmArgs a = c->lm;
      a.ncol = ncol;
      a.obs = d_obs;
      a.stokes = d_stokes;
      a.syn_d = (const double*)b_syn.p;
      a.best_syn = nullptr;
      a.sla = nullptr;
      a.eqv = eqvb;
      a.eqv_best = nullptr;
      a.isigmap = (const double*)b_isig.p;
      a.isig_stride = 0;
      a.atm = d_atm;
      a.atm_best = nullptr;
      a.done = nullptr;
      a.delta = nullptr;
      a.pix_of = nullptr;
      a.poff = nullptr;
      a.gres = gres;
      a.hscale = c->taps_sq;
      a.ext_lambda = lam;
      r = launch_lm(c, a, ncol, st);
      if (r) return r;
    }
    // END_CYCLE: current := best
    hipLaunchKernelGGL(copy_f32_kernel, dim3(gatm), dim3(256), 0, st, (const float*)best, d_atm, (long)natm);
    HIPCHK(c, hipGetLastError());
  }
  // the convolved synthesis of the final model and each pixel's chi^2
  double chi = 0.0;
  r = map_chi2(false, res, chi);
  if (r) return r;
  hipLaunchKernelGGL((transpose_kernel<float, float>), dim3((unsigned)((ncol + 31) / 32), (unsigned)((4 * numw + 31) / 32)),
                     dim3(256), 0, st, (const float*)t2, d_stokes, (long)(4 * numw), (long)ncol);
  if (d_chi2) HIPCHK(c, hipMemcpyAsync(d_chi2, b_chi.p, sizeof(double) * ncol, hipMemcpyDeviceToDevice, st));
  HIPCHK(c, hipStreamSynchronize(st));
  if (iterations) *iterations = total_it;
  return FZ_OK;
}

}  // extern "C"

}  // extern "C"

// INVERT3D on the host cubes: upload, fz_invert3d_dev or (coupled) fz_coupled_invert3d_dev, download
static int invert3d_host(fz_ctx* c, int nx, int ny, float* tem3d, float* pg3d, float* rho3d, float* pel3d,
                         float* mw3d, float* bx3d, float* by3d, float* bz3d, const float* vx3d, const float* vy3d,
                         float* vz3d, float* tau3d5, const double* obs3d, const fz_invert_control* ctl,
                         double* syn3d, double* chi2, int* iterations, bool coupled) {
  if (!c || !c->configured || nx < 1 || ny < 1 || !obs3d || !ctl || !syn3d) return FZ_EARG;
  if (c->win_on) return set_err(c, FZ_ECONFIG, "per-pixel windows: the cubes of fz_invert3d are dense");
  const int nz = c->cfg.nz, numw = c->cfg.numw, ncol = nx * ny;
  int nstk = 0;
  for (int s = 0; s < 4; ++s) nstk += ctl->invstk[s] != 0;
  const size_t nfreq = (size_t)nstk * numw;
  float* cubes[NATM] = {tem3d, pg3d, rho3d, pel3d, mw3d, bx3d, by3d, bz3d, (float*)vx3d, (float*)vy3d, vz3d, tau3d5};
  HIPCHK(c, hipSetDevice(c->device));
  DevBuf b_atm, b_tmp, b_obs, b_st, b_chi;
  const size_t cube = sizeof(float) * (size_t)nz * ncol;
  HIPCHK(c, hipMalloc(&b_atm.p, cube * NATM));
  HIPCHK(c, hipMalloc(&b_tmp.p, cube));
  HIPCHK(c, hipMalloc(&b_obs.p, sizeof(double) * nfreq * ncol));
  HIPCHK(c, hipMalloc(&b_st.p, sizeof(float) * 4 * numw * (size_t)ncol));
  HIPCHK(c, hipMalloc(&b_chi.p, sizeof(double) * ncol));
  float* d_atm = (float*)b_atm.p;
  const unsigned tb = 256, tg = (unsigned)(((long)nz * ncol + tb - 1) / tb);
  // the cubes (NZ,NY,NX) are column-major with column index y + ny*x, the
  // same column order as the device layout [parameter][height][column]
  for (int p = 0; p < NATM; ++p) {
    HIPCHK(c, hipMemcpy(b_tmp.p, cubes[p], cube, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(cube_to_soa, dim3(tg), dim3(tb), 0, 0, (const float*)b_tmp.p, d_atm + (size_t)p * nz * ncol,
                       nz, ncol);
  }
  // OBS3D(NFREQ,NY,NX) is already [column][NFREQ]
  HIPCHK(c, hipMemcpy(b_obs.p, obs3d, sizeof(double) * nfreq * ncol, hipMemcpyHostToDevice));
  const int rc = coupled ? fz_coupled_invert3d_dev(c, nx, ny, d_atm, (const double*)b_obs.p, ctl, (float*)b_st.p,
                                                    (double*)b_chi.p, iterations, nullptr)
                         : fz_invert3d_dev(c, nx, ny, d_atm, (const double*)b_obs.p, ctl, (float*)b_st.p,
                                           (double*)b_chi.p, iterations, nullptr);
  if (rc) return rc;
  // the model (every parameter the inversion or the EOS changed) back to the cubes
  const int outp[] = {A_T, A_PG, A_RHO, A_PEL, A_MU, A_BX, A_BY, A_BZ, A_VZ, A_TAU};
  for (int p : outp) {
    hipLaunchKernelGGL(soa_to_cube, dim3(tg), dim3(tb), 0, 0, d_atm + (size_t)p * nz * ncol, (float*)b_tmp.p, nz,
                       ncol);
    HIPCHK(c, hipMemcpy(cubes[p], b_tmp.p, cube, hipMemcpyDeviceToHost));
  }
  std::vector<float> st((size_t)4 * numw * ncol);
  HIPCHK(c, hipMemcpy(st.data(), b_st.p, sizeof(float) * st.size(), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < (size_t)4 * numw * ncol; ++i) syn3d[i] = (double)st[i];
  if (chi2) HIPCHK(c, hipMemcpy(chi2, b_chi.p, sizeof(double) * ncol, hipMemcpyDeviceToHost));
  return FZ_OK;
}

extern "C" {

int fz_invert3d(fz_ctx* c, int nx, int ny, float* tem3d, float* pg3d, float* rho3d, float* pel3d, float* mw3d,
                float* bx3d, float* by3d, float* bz3d, const float* vx3d, const float* vy3d, float* vz3d,
                float* tau3d5, const double* obs3d, const fz_invert_control* ctl, double* syn3d, double* chi2,
                int* iterations) {
  return invert3d_host(c, nx, ny, tem3d, pg3d, rho3d, pel3d, mw3d, bx3d, by3d, bz3d, vx3d, vy3d, vz3d, tau3d5, obs3d,
                       ctl, syn3d, chi2, iterations, false);
}

int fz_coupled_invert3d(fz_ctx* c, int nx, int ny, float* tem3d, float* pg3d, float* rho3d, float* pel3d,
                        float* mw3d, float* bx3d, float* by3d, float* bz3d, const float* vx3d, const float* vy3d,
                        float* vz3d, float* tau3d5, const double* obs3d, const fz_invert_control* ctl, double* syn3d,
                        double* chi2, int* iterations) {
  return invert3d_host(c, nx, ny, tem3d, pg3d, rho3d, pel3d, mw3d, bx3d, by3d, bz3d, vx3d, vy3d, vz3d, tau3d5, obs3d,
                       ctl, syn3d, chi2, iterations, true);
}

}  // extern "C"

// ---------------------------------------------------------------- learned initial model
namespace {

uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

int mlp_work(fz_ctx* c, int n) {
  MlpState& M = c->mlp;
  if (M.cap >= n) return FZ_OK;
  M.release_work();
  const size_t r = (size_t)n;
  const int nmax = std::max(M.nh, std::max(M.nin, M.nout));
  bool ok = hipMalloc(&M.X, sizeof(bf16) * r * M.nin) == hipSuccess &&
            hipMalloc(&M.act, sizeof(bf16) * r * M.nh * M.nl) == hipSuccess &&
            hipMalloc(&M.Y, sizeof(float) * r * M.nout) == hipSuccess &&
            hipMalloc(&M.dYf, sizeof(float) * r * M.nout) == hipSuccess &&
            hipMalloc(&M.dYb, sizeof(bf16) * r * M.nout) == hipSuccess &&
            hipMalloc(&M.partial, sizeof(float) * ((r * M.nout + 255) / 256 + 1)) == hipSuccess;
  for (int i = 0; i < 2 && ok; ++i)
    ok = hipMalloc(&M.dZf[i], sizeof(float) * r * nmax) == hipSuccess &&
         hipMalloc(&M.dZb[i], sizeof(bf16) * r * nmax) == hipSuccess;
  if (!ok) {
    M.release_work();
    return set_err(c, FZ_ENOMEM, "device allocation failed (network work buffers)");
  }
  M.cap = n;
  return FZ_OK;
}

void mlp_gemm(const GemmArgs& g, hipStream_t st) {
  const dim3 grid((unsigned)((g.N + MLP_TN - 1) / MLP_TN), (unsigned)((g.M + MLP_TM - 1) / MLP_TM));
  hipLaunchKernelGGL(mlp_gemm_kernel, grid, dim3(256), 0, st, g);
}

GemmArgs gemm0() {
  GemmArgs g;
  memset(&g, 0, sizeof(g));
  return g;
}

// forward pass of n rows already in M.X; the output (standardised) in M.Y
void mlp_forward_rows(MlpState& M, int n, hipStream_t st) {
  for (int l = 0; l <= M.nl; ++l) {
    GemmArgs g = gemm0();
    g.M = n;
    g.N = M.dout[l];
    g.K = M.din[l];
    g.A = l == 0 ? M.X : M.act + (size_t)(l - 1) * M.cap * M.nh;
    g.sam = M.din[l];
    g.sak = 1;
    g.B = M.wb + M.woff[l];
    g.sbn = M.din[l];
    g.sbk = 1;
    g.bias = M.w + M.boff[l];
    if (l < M.nl) {
      g.relu = 1;
      g.Cb = M.act + (size_t)l * M.cap * M.nh;
      g.sbm2 = M.dout[l];
      g.sbn2 = 1;
    } else {
      g.Cf = M.Y;
      g.scm = M.nout;
      g.scn = 1;
    }
    mlp_gemm(g, st);
  }
}

// loss and gradient of the parameters into M.g for n rows (targets t, rows
// gathered by idx when given); the mean squared error into M.loss[0]
void mlp_backward_rows(MlpState& M, int n, const float* d_t, const int32_t* d_idx, hipStream_t st) {
  const float* tmean = M.norms + 2 * M.nin;
  const float* tstd = tmean + M.nout;
  const long ne = (long)n * M.nout;
  const unsigned nb = (unsigned)((ne + 255) / 256);
  hipLaunchKernelGGL(mlp_loss_kernel, dim3(nb), dim3(256), 0, st, n, M.nout, (const float*)M.Y, d_t, d_idx, tmean,
                     tstd, M.dYf, M.dYb, M.partial);
  hipLaunchKernelGGL(mlp_reduce_kernel, dim3(1), dim3(256), 0, st, (int)nb, (const float*)M.partial, (float)ne, M.loss);
  const float* cf = M.dYf;
  const bf16* cb = M.dYb;
  int buf = 0;
  for (int l = M.nl; l >= 0; --l) {
    const bf16* ain = l == 0 ? M.X : M.act + (size_t)(l - 1) * M.cap * M.nh;
    const int di = M.din[l], dO = M.dout[l];
    GemmArgs g = gemm0();   // dW_l[o][k] = sum_m dZ[m][o] A_in[m][k]
    g.M = dO;
    g.N = di;
    g.K = n;
    g.A = cb;
    g.sam = 1;
    g.sak = dO;
    g.B = ain;
    g.sbn = 1;
    g.sbk = di;
    g.Cf = M.g + M.woff[l];
    g.scm = di;
    g.scn = 1;
    mlp_gemm(g, st);
    hipLaunchKernelGGL(mlp_colsum_kernel, dim3((unsigned)((dO + 255) / 256)), dim3(256), 0, st, n, dO, cf,
                       M.g + M.boff[l]);
    if (l == 0) break;
    GemmArgs h = gemm0();   // dA_in[m][k] = sum_o dZ[m][o] W_l[o][k], times ReLU'(A_in)
    h.M = n;
    h.N = di;
    h.K = dO;
    h.A = cb;
    h.sam = dO;
    h.sak = 1;
    h.B = M.wb + M.woff[l];
    h.sbn = 1;
    h.sbk = di;
    h.mask = ain;
    h.smm = di;
    h.smn = 1;
    h.Cf = M.dZf[buf];
    h.scm = di;
    h.scn = 1;
    h.Cb = M.dZb[buf];
    h.sbm2 = di;
    h.sbn2 = 1;
    mlp_gemm(h, st);
    cf = M.dZf[buf];
    cb = M.dZb[buf];
    buf ^= 1;
  }
}

int mlp_input(fz_ctx* c, int n, const float* d_x, const double* d_xd, const int32_t* d_idx, hipStream_t st) {
  MlpState& M = c->mlp;
  const long ne = (long)n * M.nin;
  hipLaunchKernelGGL(mlp_input_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, n, M.nin, d_x, d_xd,
                     d_idx, (const float*)M.norms, (const float*)(M.norms + M.nin), M.X);
  HIPCHK(c, hipGetLastError());
  return FZ_OK;
}

}  // namespace

extern "C" {

int fz_mlp_create(fz_ctx* c, int nin, int nhidden, int nlayers, int nodes, uint64_t seed) {
  if (!c || !c->configured || nin < 1 || nhidden < 1 || nlayers < 1 || nlayers > 6 || nodes < 1 ||
      nodes > MLP_MAXNODE || nodes > c->cfg.nz)
    return set_err(c, FZ_EARG, "network shape outside limits");
  HIPCHK(c, hipSetDevice(c->device));
  c->mlp.release();
  MlpState& M = c->mlp;
  M.nin = nin;
  M.nh = nhidden;
  M.nl = nlayers;
  M.nj = nodes;
  M.nout = MLP_NQ * nodes;
  M.nz = c->cfg.nz;
  long off = 0;
  for (int l = 0; l <= nlayers; ++l) {
    M.din[l] = l == 0 ? nin : nhidden;
    M.dout[l] = l == nlayers ? M.nout : nhidden;
    M.woff[l] = off;
    off += (long)M.din[l] * M.dout[l];
    M.boff[l] = off;
    off += M.dout[l];
  }
  M.np = off;
  // He-uniform weights for ReLU layers (bound sqrt(6 / fan_in)), the output
  // layer scaled by 1/4; zero biases
  std::vector<float> w(off, 0.f);
  uint64_t s = seed;
  for (int l = 0; l <= nlayers; ++l) {
    const double b = std::sqrt(6.0 / M.din[l]) * (l == nlayers ? 0.25 : 1.0);
    for (long i = 0; i < (long)M.din[l] * M.dout[l]; ++i)
      w[M.woff[l] + i] = (float)(b * (2.0 * ((splitmix64(s) >> 11) * (1.0 / 9007199254740992.0)) - 1.0));
  }
  bool ok = hipMalloc(&M.w, sizeof(float) * off) == hipSuccess && hipMalloc(&M.g, sizeof(float) * off) == hipSuccess &&
            hipMalloc(&M.m1, sizeof(float) * off) == hipSuccess &&
            hipMalloc(&M.m2, sizeof(float) * off) == hipSuccess && hipMalloc(&M.wb, sizeof(bf16) * off) == hipSuccess &&
            hipMalloc(&M.norms, sizeof(float) * 2 * (nin + M.nout)) == hipSuccess &&
            hipMalloc(&M.pinv, sizeof(double) * nodes * M.nz) == hipSuccess &&
            hipMalloc(&M.coefs, sizeof(double) * nodes * M.nz) == hipSuccess &&
            hipMalloc(&M.loss, sizeof(float) * 4) == hipSuccess;
  if (!ok) {
    M.release();
    return set_err(c, FZ_ENOMEM, "device allocation failed (network)");
  }
  HIPCHK(c, hipMemcpy(M.w, w.data(), sizeof(float) * off, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemset(M.m1, 0, sizeof(float) * off));
  HIPCHK(c, hipMemset(M.m2, 0, sizeof(float) * off));
  HIPCHK(c, hipMemset(M.g, 0, sizeof(float) * off));
  hipLaunchKernelGGL(mlp_to_bf16_kernel, dim3((unsigned)((off + 255) / 256)), dim3(256), 0, 0, off, (const float*)M.w,
                     M.wb);
  HIPCHK(c, hipGetLastError());
  std::vector<float> nrm(2 * (nin + M.nout));
  for (int i = 0; i < nin; ++i) {
    nrm[i] = 0.f;
    nrm[nin + i] = 1.f;
  }
  for (int i = 0; i < M.nout; ++i) {
    nrm[2 * nin + i] = 0.f;
    nrm[2 * nin + M.nout + i] = 1.f;
  }
  HIPCHK(c, hipMemcpy(M.norms, nrm.data(), sizeof(float) * nrm.size(), hipMemcpyHostToDevice));
  // hat basis of GET_INT_COEFS (fz_int_coefs: node-major [nj][nz]) as
  // cf[k * nj + j], and its least-squares pseudo-inverse pv[j * nz + k]
  std::vector<double> cf((size_t)M.nz * nodes), pv((size_t)nodes * M.nz);
  {
    std::vector<double> cj((size_t)M.nz * nodes);
    const int r0 = fz_int_coefs(M.nz, nodes, cj.data());
    if (r0) return set_err(c, r0, "node basis");
    for (int j = 0; j < nodes; ++j)
      for (int k = 0; k < M.nz; ++k) cf[(size_t)k * nodes + j] = cj[(size_t)j * M.nz + k];
  }
  {
    // normal equations (C^T C) P = C^T, solved by Gauss-Jordan on nodes x nodes
    const int nj = nodes, nz = M.nz;
    std::vector<double> a((size_t)nj * nj, 0.0), inv((size_t)nj * nj, 0.0);
    for (int i = 0; i < nj; ++i)
      for (int j = 0; j < nj; ++j)
        for (int k = 0; k < nz; ++k) a[i * nj + j] += cf[(size_t)k * nj + i] * cf[(size_t)k * nj + j];
    for (int i = 0; i < nj; ++i) inv[i * nj + i] = 1.0;
    for (int p = 0; p < nj; ++p) {
      int piv = p;
      for (int i = p + 1; i < nj; ++i)
        if (std::fabs(a[i * nj + p]) > std::fabs(a[piv * nj + p])) piv = i;
      if (a[piv * nj + p] == 0.0) return set_err(c, FZ_ECONFIG, "singular node basis");
      for (int j = 0; j < nj; ++j) {
        std::swap(a[p * nj + j], a[piv * nj + j]);
        std::swap(inv[p * nj + j], inv[piv * nj + j]);
      }
      const double d = a[p * nj + p];
      for (int j = 0; j < nj; ++j) {
        a[p * nj + j] /= d;
        inv[p * nj + j] /= d;
      }
      for (int i = 0; i < nj; ++i) {
        if (i == p) continue;
        const double f = a[i * nj + p];
        for (int j = 0; j < nj; ++j) {
          a[i * nj + j] -= f * a[p * nj + j];
          inv[i * nj + j] -= f * inv[p * nj + j];
        }
      }
    }
    for (int j = 0; j < nj; ++j)
      for (int k = 0; k < nz; ++k) {
        double v = 0.0;
        for (int i = 0; i < nj; ++i) v += inv[j * nj + i] * cf[(size_t)k * nj + i];
        pv[(size_t)j * nz + k] = v;
      }
  }
  HIPCHK(c, hipMemcpy(M.coefs, cf.data(), sizeof(double) * cf.size(), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(M.pinv, pv.data(), sizeof(double) * pv.size(), hipMemcpyHostToDevice));
  M.on = true;
  return FZ_OK;
}

int64_t fz_mlp_size(const fz_ctx* c) {
  if (!c || !c->mlp.on) return -1;
  return c->mlp.np + 2 * (c->mlp.nin + c->mlp.nout);
}

int fz_mlp_get_params(fz_ctx* c, float* params) {
  if (!c || !c->mlp.on || !params) return FZ_EARG;
  MlpState& M = c->mlp;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpy(params, M.w, sizeof(float) * M.np, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(params + M.np, M.norms, sizeof(float) * 2 * (M.nin + M.nout), hipMemcpyDeviceToHost));
  return FZ_OK;
}

int fz_mlp_set_params(fz_ctx* c, const float* params) {
  if (!c || !c->mlp.on || !params) return FZ_EARG;
  MlpState& M = c->mlp;
  HIPCHK(c, hipSetDevice(c->device));
  for (long i = 0; i < M.np + 2 * (M.nin + M.nout); ++i)
    if (!std::isfinite(params[i])) return set_err(c, FZ_EARG, "non-finite network parameter");
  HIPCHK(c, hipMemcpy(M.w, params, sizeof(float) * M.np, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(M.norms, params + M.np, sizeof(float) * 2 * (M.nin + M.nout), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(mlp_to_bf16_kernel, dim3((unsigned)((M.np + 255) / 256)), dim3(256), 0, 0, M.np,
                     (const float*)M.w, M.wb);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemset(M.m1, 0, sizeof(float) * M.np));
  HIPCHK(c, hipMemset(M.m2, 0, sizeof(float) * M.np));
  M.step = 0;
  return FZ_OK;
}

int fz_mlp_fit_norms_dev(fz_ctx* c, int n, const float* d_x, const float* d_t, void* stream) {
  if (!c || !c->mlp.on || n < 2 || !d_x || !d_t) return FZ_EARG;
  MlpState& M = c->mlp;
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(c, hipSetDevice(c->device));
  hipLaunchKernelGGL(mlp_stats_kernel, dim3((unsigned)((M.nin + 63) / 64)), dim3(64), 0, st, n, M.nin, d_x, M.norms,
                     M.norms + M.nin);
  hipLaunchKernelGGL(mlp_stats_kernel, dim3((unsigned)((M.nout + 63) / 64)), dim3(64), 0, st, n, M.nout, d_t,
                     M.norms + 2 * M.nin, M.norms + 2 * M.nin + M.nout);
  HIPCHK(c, hipGetLastError());
  return FZ_OK;
}

int fz_mlp_targets_dev(fz_ctx* c, int ncol, const float* d_atm, float* d_t, void* stream) {
  if (!c || !c->mlp.on || ncol < 1 || !d_atm || !d_t) return FZ_EARG;
  if (c->cfg.nz != c->mlp.nz) return set_err(c, FZ_ECONFIG, "height grid changed since fz_mlp_create");
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(c, hipSetDevice(c->device));
  hipLaunchKernelGGL(mlp_targets_kernel, dim3((unsigned)((ncol + 127) / 128)), dim3(128), 0, st, ncol, c->mlp.nz,
                     c->mlp.nj, d_atm, (const double*)c->mlp.pinv, d_t);
  HIPCHK(c, hipGetLastError());
  return FZ_OK;
}

int fz_mlp_forward_dev(fz_ctx* c, int n, const float* d_x, const double* d_xd, float* d_y, void* stream) {
  if (!c || !c->mlp.on || n < 1 || (!d_x && !d_xd) || !d_y) return FZ_EARG;
  MlpState& M = c->mlp;
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(c, hipSetDevice(c->device));
  const int chunk = 65536;
  int r = mlp_work(c, std::min(n, chunk));
  if (r) return r;
  for (int r0 = 0; r0 < n; r0 += M.cap) {
    const int nr = std::min(M.cap, n - r0);
    r = mlp_input(c, nr, d_x ? d_x + (size_t)r0 * M.nin : nullptr, d_xd ? d_xd + (size_t)r0 * M.nin : nullptr,
                  nullptr, st);
    if (r) return r;
    mlp_forward_rows(M, nr, st);
    HIPCHK(c, hipMemcpyAsync(d_y + (size_t)r0 * M.nout, M.Y, sizeof(float) * nr * M.nout, hipMemcpyDeviceToDevice,
                             st));
  }
  HIPCHK(c, hipGetLastError());
  return FZ_OK;
}

int fz_mlp_gradient_dev(fz_ctx* c, int n, const float* d_x, const float* d_t, const int32_t* d_idx, float* d_loss,
                        float* d_grad, void* stream) {
  if (!c || !c->mlp.on || n < 1 || !d_x || !d_t) return FZ_EARG;
  MlpState& M = c->mlp;
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(c, hipSetDevice(c->device));
  int r = mlp_work(c, n);
  if (r) return r;
  r = mlp_input(c, n, d_x, nullptr, d_idx, st);
  if (r) return r;
  mlp_forward_rows(M, n, st);
  mlp_backward_rows(M, n, d_t, d_idx, st);
  if (d_loss) HIPCHK(c, hipMemcpyAsync(d_loss, M.loss, sizeof(float), hipMemcpyDeviceToDevice, st));
  if (d_grad) HIPCHK(c, hipMemcpyAsync(d_grad, M.g, sizeof(float) * M.np, hipMemcpyDeviceToDevice, st));
  HIPCHK(c, hipGetLastError());
  return FZ_OK;
}

int fz_mlp_train_step_dev(fz_ctx* c, int n, const float* d_x, const float* d_t, const int32_t* d_idx, float lr,
                          float* d_loss, void* stream) {
  int r = fz_mlp_gradient_dev(c, n, d_x, d_t, d_idx, d_loss, nullptr, stream);
  if (r) return r;
  MlpState& M = c->mlp;
  const float b1 = 0.9f, b2 = 0.999f;
  ++M.step;
  const float c1 = (float)(1.0 - std::pow((double)b1, (double)M.step));
  const float c2 = (float)(1.0 - std::pow((double)b2, (double)M.step));
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((unsigned)((M.np + 255) / 256)), dim3(256), 0, (hipStream_t)stream, M.np,
                     M.w, (const float*)M.g, M.m1, M.m2, M.wb, lr, b1, b2, 1e-8f, c1, c2);
  HIPCHK(c, hipGetLastError());
  return FZ_OK;
}

int fz_learned_init_dev(fz_ctx* c, int ncol, const double* d_obs, float* d_atm, void* stream) {
  if (!c || !c->mlp.on || ncol < 1 || !d_obs || !d_atm) return FZ_EARG;
  if (c->cfg.nz != c->mlp.nz) return set_err(c, FZ_ECONFIG, "height grid changed since fz_mlp_create");
  MlpState& M = c->mlp;
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(c, hipSetDevice(c->device));
  DevBuf y;
  HIPCHK(c, hipMalloc(&y.p, sizeof(float) * (size_t)ncol * M.nout));
  int r = fz_mlp_forward_dev(c, ncol, nullptr, d_obs, (float*)y.p, stream);
  if (r) return r;
  hipLaunchKernelGGL(mlp_decode_kernel, dim3((unsigned)((ncol + 127) / 128)), dim3(128), 0, st, ncol, M.nz, M.nj,
                     (const float*)y.p, (const float*)(M.norms + 2 * M.nin), (const float*)(M.norms + 2 * M.nin + M.nout),
                     (const double*)M.coefs, d_atm);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(st));
  return FZ_OK;
}

}  // extern "C"
