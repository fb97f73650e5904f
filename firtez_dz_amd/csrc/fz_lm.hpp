// Per-pixel Levenberg-Marquardt step of the FIRTEZ-dz inversion on the GPU.
//
//   eqv_kernel : node contraction of the per-height response functions
//                (CALCULATE_EQUIVALENT_DSYN, forward.f90:552-681): one thread
//                per (pixel, spectral row), FP64 accumulation over the support
//                of each node's hat function, REAL(4) store.
//   lm_kernel  : one 256-thread workgroup per pixel, FP64 throughout except
//                the REAL(4) model (GET_PERT_1D, get_dmodel.f90:1201-1244):
//                chi^2 and accept/reject (:960-1032), lambda update
//                (:917-956), Jacobian normalisation (:725-770), Gauss-Newton
//                normal matrix and gradient streamed through LDS in chunks of
//                64 spectral rows, damped pseudo-inverse by a parallel cyclic
//                Jacobi eigen-decomposition in LDS (the symmetric form of the
//                truncated SVD of lm.f90:61-133), and the limited
//                perturbation applied through the node basis (:1715-2016).
#pragma once

namespace fz {

constexpr int LM_MAXNP = 64;     // free parameters per pixel with the work arrays in LDS
// one wave per pixel: the Jacobi rounds and reductions are wave-local, so
// the barriers are cheap and a CU holds many pixels at once
constexpr int LM_THREADS = 64;
constexpr int LM_CH = 32;        // spectral rows per chunk of the normal matrix
// more free parameters (up to NZ nodes per parameter, the reference's
// default): the same step with its work arrays in a per-pixel slice of
// device memory, four waves per pixel
constexpr int LM_MAXNP_BIG = 2048;
constexpr int LM_BIG_THREADS = 256;

// get_dmodel.f90:31 (PARAM_NORMS) and :1903-1919 (hard bounds)
__constant__ double c_lm_norm[8] = {6000.0, 1.0e5, 1.0e-5, 1.0e3, 1.0e3, 1.0e3, 1.0e6, 6.0e2};
__constant__ double c_lm_low[8] = {2500.0, 1.0e-2, 1.0e-10, -5.0e4, -5.0e4, -5.0e4, -1.0e7, 1.0e-8};
__constant__ double c_lm_upp[8] = {1.8e4, 1.0e12, 1.0e10, 5.0e4, 5.0e4, 5.0e4, 1.0e7, 1.0e12};

struct LmArgs {
  int ncol, nz, numw, nfreq, np;
  int nparam;                  // inverted physical parameters
  int par[8];                  // physical index 0..7 (T, Pg, rho, Bx, By, Bz, vz, P0)
  int slot[8];                 // response-function slot in the Cartesian per-height output
  int mcol[8];                 // model column
  int j0[8], nj[8];            // first free parameter, nodes
  int rowstk[4];               // Stokes index of each inverted Stokes block
  int regul, solver, node_expansion, pen_typ[8];
  double pen_alp[8];
  double svdtol, inu, mpert, maxsteps;
  const double* coefs;         // [np][nz]
  const double* colsum;        // [np] column sums of the node basis
  // per pixel
  const double* obs;           // [ncol][nfreq]
  const float* stokes;         // [ncol][4][numw] current synthesis (times GAIN1D)
  const double* syn_d;
  double* best_syn;            // [ncol][nfreq]
  double* sla;                 // [ncol][9]
  const float* eqv;            // [ncol][np][nfreq] current equivalent response functions
  float* eqv_best;             // [ncol][np][nfreq]
  const double* isigmap;       // [nfreq], or [ncol][nfreq] with isig_stride = nfreq
  long isig_stride;            // 0: one weight vector for every pixel (ISIGMAP3D equal over the map)
  // synthesis of the inverted Stokes parameters in FP64, [ncol][nfreq] like
  // obs (SYN3D of the host-array entry point), or null: the float Stokes rows
  float* atm;                  // [12][nz][ncol] current model (in: model of eqv; out: new trial)
  float* atm_best;             // [12][nz][ncol]
  int32_t* done;               // [ncol] AM_I_DONE
  double* delta;               // [ncol][np] applied node perturbation, or null
  // active-pixel launch: workgroup b handles pixel pix_of[b], and stokes/eqv
  // hold the compacted rows (row b); null: workgroup b is pixel b
  const int32_t* pix_of;
  // per-pixel windows (poff null: none): pixel pix has numw_pix =
  // poff[pix+1]-poff[pix] wavelengths; its obs/best_syn rows start at
  // nstk*poff[pix], its Stokes at 4*poff[pix], its eqv/eqv_best at
  // np*nstk*poff[pix]; isigmap is read per Stokes block (isigmap[b*numw]);
  // inu_pix: chi^2 normalised by the pixel's own NFREQ
  const int64_t* poff;
  int nstk, inu_pix;
  // work arrays in device memory (more than LM_MAXNP free parameters):
  // workgroup b handles row row0 + b and uses work[b * work_stride ...]
  double* work;
  long work_stride;
  int row0;
  // externally controlled step (the PSF-coupled inversion, whose acceptance
  // and damping are decided over the whole map): gres non-null gives the
  // residual of the gradient J^T W gres per pixel ([ncol][nfreq], the
  // adjoint-convolved residual), hscale multiplies J^T W J (the sum of the
  // squared PSF taps: the pixel's own block of the coupled normal matrix) and
  // ext_lambda is the damping; chi^2, acceptance, INV_SLA, the best model
  // and AM_I_DONE are left to the caller, eqv is the Jacobian of the model
  // in atm
  const double* gres = nullptr;
  double hscale = 1.0, ext_lambda = 0.0;
};

// ---------------------------------------------------------------- node contraction
struct EqvArgs {
  int ncol_tile, col0, nz, numw, nfreq, np, nstk;
  int nslot;                   // response-function slots per column in the tile (7, or 8 with P0)
  int rowstk[4];
  const int* slot_of;          // [np] response-function slot of each free parameter
  const int* kbeg;             // [np] support of each node's hat function
  const int* kend;             // [np]
  // per parameter and height (pairs): the lowest unfinished node in bits
  // 0-15, bit 30 set when it takes height k, bit 29 when the next node does;
  // tab_c the two coefficients, so each height's loads are independent of
  // the running node index
  const int* tab_j;            // [nparam][nz]
  const double* tab_c;         // [nparam][nz][2]
  // the nodes of inverted parameter ip are j0[ip] .. j0[ip]+nj[ip]-1; pairs:
  // every height lies in the support of at most two consecutive nodes of a
  // parameter and the supports move up with the node index (hat functions),
  // so one pass over the heights feeds two running sums
  int nparam, pairs;
  int j0[8], nj[8];
  const double* coefs;         // [np][nz]
  const float* jac;            // tile: [ntile][nslot][nz][4][numw], ragged with windows
  float* eqv;                  // [ncol][np][nfreq], ragged with windows
  // per-pixel windows (poff null: none): the tile's response functions are
  // [nslot][nz][4][numw_col] per column from nslot*nz*4*(poff[col]-poff[col0]);
  // the output block of pixel col is [np][nstk][numw_col] from np*nstk*poff[col]
  const int64_t* poff;
};

// the column of the tile [c0, c0+n) whose ragged range [poff[c], poff[c+1]) holds g
__device__ __forceinline__ int ragged_column(const int64_t* poff, int c0, int n, int64_t g) {
  int lo = c0, hi = c0 + n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (poff[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// EQVDSYN (forward.f90:552-681): one thread per (column, inverted Stokes row,
// wavelength), all free parameters; FP64 sums over the node's heights
__global__ void __launch_bounds__(256) eqv_kernel(EqvArgs a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const float* jc;
  float* out;
  long cs, fstride;
  if (a.poff) {
    const int64_t base = a.poff[a.col0];
    if (t >= (long)a.nstk * (a.poff[a.col0 + a.ncol_tile] - base)) return;
    const int col = ragged_column(a.poff, a.col0, a.ncol_tile, base + t / a.nstk);
    const int64_t p0 = a.poff[col];
    const int nw = (int)(a.poff[col + 1] - p0);
    const int f = (int)(t - (long)a.nstk * (p0 - base));
    const int b = f / nw, l = f - b * nw;
    cs = (long)4 * nw;
    jc = a.jac + (long)a.nslot * a.nz * 4 * (p0 - base) + (long)a.rowstk[b] * nw + l;
    fstride = (long)a.nstk * nw;
    out = a.eqv + (long)a.np * a.nstk * p0 + f;
  } else {
    if (t >= (long)a.ncol_tile * a.nfreq) return;
    const int ct = (int)(t / a.nfreq);
    const int f = (int)(t - (long)ct * a.nfreq);
    const int b = f / a.numw, l = f - b * a.numw;
    cs = (long)4 * a.numw;
    jc = a.jac + (long)ct * a.nslot * a.nz * cs + a.rowstk[b] * a.numw + l;
    fstride = a.nfreq;
    out = a.eqv + ((long)(a.col0 + ct) * a.np) * a.nfreq + f;
  }
  if (a.pairs) {
    // each node's sum is formed in the same order (heights ascending from
    // its support's start) as the per-node loop below: identical results,
    // half the reads where the hat supports overlap
    for (int ip = 0; ip < a.nparam; ++ip) {
      const int jb = a.j0[ip], je = jb + a.nj[ip] - 1;
      const float* jp = jc + (long)a.slot_of[jb] * a.nz * cs;
      const int* tj = a.tab_j + (long)ip * a.nz;
      const double* tc = a.tab_c + (long)ip * a.nz * 2;
      int jl = jb;                                   // lowest unfinished node
      double alo = 0.0, ahi = 0.0;                   // sums of nodes jl, jl+1
      const int k0 = a.kbeg[jb], k1 = a.kend[je];
      // the loads of eight heights are issued together (one strided read per
      // height and lane: eight in flight instead of one), the sums then take
      // them in height order as before
      for (int kb = k0; kb <= k1; kb += 8) {
        float vb[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) vb[i] = (kb + i <= k1) ? jp[(long)(kb + i) * cs] : 0.f;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int k = kb + i;
          if (k > k1) break;
          const int w = tj[k];
          while (jl < (w & 0xffff)) {
            out[(long)jl * fstride] = (float)alo;
            alo = ahi;
            ahi = 0.0;
            ++jl;
          }
          const double v = (double)vb[i];
          if (w & (1 << 30)) alo += v * tc[2 * k];
          if (w & (1 << 29)) ahi += v * tc[2 * k + 1];
        }
      }
      out[(long)jl * fstride] = (float)alo;
      if (jl < je) out[(long)(jl + 1) * fstride] = (float)ahi;
    }
    return;
  }
  for (int j = 0; j < a.np; ++j) {
    const float* jp = jc + (long)a.slot_of[j] * a.nz * cs;
    const double* cf = a.coefs + (long)j * a.nz;
    double acc = 0.0;
    for (int k = a.kbeg[j]; k <= a.kend[j]; ++k) acc += (double)jp[k * cs] * cf[k];
    out[(long)j * fstride] = (float)acc;
  }
}

// ---------------------------------------------------------------- LM step
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (t < s) red[t] += red[t + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// GET_IT_PENALTIES (get_dmodel.f90:1041-1197) by one thread: pen_res[np],
// pen_hss[np*np] (row-major, block diagonal) from the model columns in LDS;
// sm and pen are LDS work vectors of np entries.
__device__ void lm_penalties(const LmArgs& a, const float* mod, double* pen_res, double* pen_hss, double* sm,
                             double* pen) {
  const int np = a.np, nz = a.nz;
  for (int i = 0; i < np; ++i) pen_res[i] = 0.0;
  for (int i = 0; i < np * np; ++i) pen_hss[i] = 0.0;
  for (int ip = 0; ip < a.nparam; ++ip) {
    const int p = a.par[ip], ji = a.j0[ip], jd = a.nj[ip];
    const float* col = mod + a.mcol[ip] * nz;
    const double alp = sqrt(a.pen_alp[p] / (double)np);
    for (int j = 0; j < jd; ++j) {
      const double* cf = a.coefs + (long)(ji + j) * nz;
      double s = 0.0;
      for (int k = 0; k < nz; ++k) s = s + (double)col[k] * cf[k];
      sm[j] = s / a.colsum[ji + j] / c_lm_norm[p];
    }
    // L: type 0 -> diag(1/jd), type 1 -> first differences (first row empty)
    for (int k = 0; k < jd; ++k) {
      double v = 0.0;
      if (a.pen_typ[p] == 0) v = (1.0 * (1.0 / (double)jd)) * alp * sm[k];
      else if (a.pen_typ[p] == 1 && k > 0) v = (1.0 * 1.0) * alp * sm[k] + (-1.0 * 1.0) * alp * sm[k - 1];
      pen[k] = v;
    }
    for (int j = 0; j < jd; ++j) {
      double s = 0.0;
      if (a.pen_typ[p] == 0) s = (1.0 / (double)jd) * alp * pen[j];
      else if (a.pen_typ[p] == 1) {
        if (j > 0) s += alp * pen[j];
        if (j + 1 < jd) s += -alp * pen[j + 1];
      }
      pen_res[ji + j] = s;
    }
    // L^T L into the zeroed block: type 0 is diagonal; type 1, with
    // L_kk = alp (k>0) and L_k,k-1 = -alp, is tridiagonal:
    // (L^T L)_ii = sum over k = i (i>0), then k = i+1 (i+1<jd) of alp^2,
    // (L^T L)_i,i+-1 = -alp^2 from the single k = max(i, j)
    for (int i = 0; i < jd; ++i) {
      double* row = pen_hss + (long)(ji + i) * np + ji;
      if (a.pen_typ[p] == 0) {
        row[i] = ((1.0 / (double)jd) * alp) * ((1.0 / (double)jd) * alp);
      } else if (a.pen_typ[p] == 1) {
        double d = 0.0;
        if (i > 0) d += alp * alp;
        if (i + 1 < jd) d += (-alp) * (-alp);
        row[i] = d;
        if (i > 0) row[i - 1] = alp * -alp;
        if (i + 1 < jd) row[i + 1] = -alp * alp;
      }
    }
  }
}

// Damped solve of GET_PERTURBATION (lm.f90:137-204) on a symmetric matrix
// A[np][ld] (ld = np + 1; rows up to np + (np & 1) zero-padded) and the
// right-hand side g, by all NT threads of the workgroup; V must hold the
// identity.  solver 1: parallel cyclic Jacobi (tournament order) and the
// truncated pseudo-inverse of GET_INVERSE (lm.f90:61-133: values below
// svdtol * max dropped); solver 2: the CG iteration of
// NEW_CALCULATE_DELTA_UBICGSTAB (lm.f90:375-453).  Work vectors: dv[np],
// red[NT], rot[np + 1], nod[4 np].  Result in del[np]; A and V are destroyed.
template <int NT>
__device__ void damped_solve(int solver, double svdtol, int np, double* A, double* V, const double* g,
                             double* dv, double* del, double* red, double* rot, double* nod) {
  const int t = threadIdx.x;
  const int ld = np + 1;
  const int npe = np + (np & 1);
  if (solver == 1) {
    const int half = npe / 2;
    for (int sweep = 0; sweep < 60; ++sweep) {
      double o = 0.0, tt = 0.0;
      for (int e = t; e < np * np; e += NT) {
        const int i = e / np, j = e - i * np;
        const double x = A[i * ld + j];
        tt += x * x;
        if (i != j) o += x * x;
      }
      const double off = block_sum<NT>(o, red), tot = block_sum<NT>(tt, red);
      if (off <= 1e-30 * tot || off == 0.0) break;
      for (int rnd = 0; rnd < npe - 1; ++rnd) {
        // pair i: players L[i], L[npe-1-i] with L = [0, (rnd+k)%(npe-1)+1 ...]
        for (int i = t; i < half; i += NT) {
          const int p0 = (i == 0) ? 0 : ((rnd + i - 1) % (npe - 1)) + 1;
          const int q0 = ((rnd + (npe - 1 - i) - 1) % (npe - 1)) + 1;
          const int p = min(p0, q0), q = max(p0, q0);
          double cs = 1.0, sn = 0.0;
          if (q < np) {
            const double apq = A[p * ld + q];
            if (apq != 0.0) {
              const double th = (A[q * ld + q] - A[p * ld + p]) / (2.0 * apq);
              const double tq = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(1.0 + th * th));
              cs = 1.0 / sqrt(1.0 + tq * tq);
              sn = tq * cs;
            }
          }
          rot[2 * i] = cs;
          rot[2 * i + 1] = sn;
        }
        __syncthreads();
        // columns of A and V: A <- A J, V <- V J
        for (int e = t; e < half * npe; e += NT) {
          const int i = e / npe, k = e - i * npe;
          const int p0 = (i == 0) ? 0 : ((rnd + i - 1) % (npe - 1)) + 1;
          const int q0 = ((rnd + (npe - 1 - i) - 1) % (npe - 1)) + 1;
          const int p = min(p0, q0), q = max(p0, q0);
          const double cs = rot[2 * i], sn = rot[2 * i + 1];
          if (sn == 0.0 && cs == 1.0) continue;
          const double akp = A[k * ld + p], akq = A[k * ld + q];
          A[k * ld + p] = cs * akp - sn * akq;
          A[k * ld + q] = sn * akp + cs * akq;
          const double vkp = V[k * ld + p], vkq = V[k * ld + q];
          V[k * ld + p] = cs * vkp - sn * vkq;
          V[k * ld + q] = sn * vkp + cs * vkq;
        }
        __syncthreads();
        // rows of A: A <- J^T A
        for (int e = t; e < half * npe; e += NT) {
          const int i = e / npe, k = e - i * npe;
          const int p0 = (i == 0) ? 0 : ((rnd + i - 1) % (npe - 1)) + 1;
          const int q0 = ((rnd + (npe - 1 - i) - 1) % (npe - 1)) + 1;
          const int p = min(p0, q0), q = max(p0, q0);
          const double cs = rot[2 * i], sn = rot[2 * i + 1];
          if (sn == 0.0 && cs == 1.0) continue;
          const double apk = A[p * ld + k], aqk = A[q * ld + k];
          A[p * ld + k] = cs * apk - sn * aqk;
          A[q * ld + k] = sn * apk + cs * aqk;
        }
        __syncthreads();
      }
    }
    // truncated pseudo-inverse: keep |d| >= svdtol * max|d|
    double mx = 0.0;
    for (int i = 0; i < np; ++i) mx = fmax(mx, fabs(A[i * ld + i]));
    const double thr = mx * (svdtol > 0 ? svdtol : 10.0e-4);
    for (int e = t; e < np; e += NT) {
      const double d = A[e * ld + e];
      double pr = 0.0;
      if (!(fabs(d) < thr)) {
        for (int i = 0; i < np; ++i) pr += V[i * ld + e] * g[i];
        pr /= d;
      }
      dv[e] = pr;
    }
    __syncthreads();
    for (int i = t; i < np; i += NT) {
      double s = 0.0;
      for (int e = 0; e < np; ++e) s += V[i * ld + e] * dv[e];
      del[i] = s;
    }
  } else {
    // NEW_CALCULATE_DELTA_UBICGSTAB (lm.f90:375-453): the matrix-vector and
    // dot products spread over the workgroup; every thread holds the same
    // scalars, so the iteration and its stopping test are uniform
    double* xv = nod;
    double* pv = nod + np;
    double* rv = nod + 2 * np;
    double* qv = nod + 3 * np;
    auto dot = [&](const double* x, const double* y) {
      double s = 0.0;
      for (int i = t; i < np; i += NT) s += x[i] * y[i];
      return block_sum<NT>(s, red);
    };
    // A is exactly symmetric here (its lower triangle mirrored, a symmetric
    // penalty Hessian, the diagonal scaled): row i is read as column i, so
    // the lanes of a wave read consecutive entries
    auto matvec = [&](const double* x, double* y) {
      for (int i = t; i < np; i += NT) {
        double s = 0.0;
        for (int j = 0; j < np; ++j) s += A[j * ld + i] * x[j];
        y[i] = s;
      }
      __syncthreads();
    };
    for (int i = t; i < np; i += NT) {
      xv[i] = 0.0;
      pv[i] = g[i];
      rv[i] = g[i];   // bet - alp x with x = 0
    }
    __syncthreads();
    double crho = dot(rv, rv);
    matvec(pv, qv);
    double calpha = crho / dot(pv, qv);
    for (int i = t; i < np; i += NT) {
      xv[i] = xv[i] + calpha * pv[i];
      rv[i] = rv[i] - calpha * qv[i];
    }
    __syncthreads();
    double pcond = 1.0e29, dcond = -1.0e29;
    bool conv = false;
    // each iteration forms the next direction before its stopping test, so
    // one pass over A gives both A x (the test) and A p (the next step);
    // each of the two sums keeps the order of a separate product
    {
      const double crho0 = crho;
      crho = dot(rv, rv);
      const double cbeta = crho / crho0;
      for (int i = t; i < np; i += NT) pv[i] = rv[i] + cbeta * pv[i];
      __syncthreads();
      matvec(pv, qv);
    }
    for (int it = 1; it <= 100000; ++it) {
      calpha = crho / dot(pv, qv);
      for (int i = t; i < np; i += NT) {
        xv[i] = xv[i] + calpha * pv[i];
        rv[i] = rv[i] - calpha * qv[i];
      }
      __syncthreads();
      const double crho0 = crho;
      crho = dot(rv, rv);
      const double cbeta = crho / crho0;
      for (int i = t; i < np; i += NT) pv[i] = rv[i] + cbeta * pv[i];
      __syncthreads();
      double cs = 0.0;
      for (int i = t; i < np; i += NT) {
        double s = 0.0, sq = 0.0;
        for (int j = 0; j < np; ++j) {
          const double aij = A[j * ld + i];
          s += aij * xv[j];
          sq += aij * pv[j];
        }
        qv[i] = sq;
        const double e = (s - g[i]) / (fabs(g[i]) + 1.0e-5);
        cs += e * e;
      }
      const double cond = block_sum<NT>(cs, red) / (double)np;
      if (cond < 1.0e-3 || dcond > 1.0e30) {
        for (int i = t; i < np; i += NT) del[i] = xv[i];
        conv = true;
        break;
      }
      dcond = cond - pcond;
      pcond = cond;
    }
    if (!conv)
      for (int i = t; i < np; i += NT) del[i] = xv[i] / 10.0;
  }
}

template <bool BIG>
__global__ void __launch_bounds__(BIG ? LM_BIG_THREADS : LM_THREADS) lm_kernel(LmArgs a) {
  constexpr int NT = BIG ? LM_BIG_THREADS : LM_THREADS;
  extern __shared__ double smem_dyn[];
  const int t = threadIdx.x;
  const int row = a.row0 + (int)blockIdx.x;
  double* smem = BIG ? a.work + (long)blockIdx.x * a.work_stride : smem_dyn;
  const int pix = a.pix_of ? a.pix_of[row] : row;
  if (a.done && a.done[pix] != 0) return;
  const int np = a.np, nz = a.nz;
  int nf = a.nfreq, numw = a.numw;
  long o_obs = (long)pix * a.nfreq, o_stk = (long)row * 4 * a.numw;
  long o_ecur = (long)row * np * a.nfreq, o_ebest = (long)pix * np * a.nfreq;
  double inu = a.inu;
  if (a.poff) {
    const int64_t p0 = a.poff[pix];
    numw = (int)(a.poff[pix + 1] - p0);
    nf = a.nstk * numw;
    o_obs = (long)a.nstk * p0;
    o_stk = 4 * p0;
    o_ecur = o_ebest = (long)np * a.nstk * p0;
    if (a.inu_pix) inu = (double)nf;
  }
  // inverse noise weight of row f (one value per Stokes block with windows)
  const double* isg = a.isigmap + (long)pix * a.isig_stride;
  auto isig = [&](int f) -> double {
    return a.poff ? isg[(long)(f / numw) * a.numw] : isg[f];
  };
  const int ld = np + 1;                       // LDS row stride of A and V
  const int npe = np + (np & 1);               // even size for the tournament ordering
  // LDS carve-up
  double* A = smem;                            // [npe][ld]  (ld >= npe when np odd: ld = np+1 = npe)
  double* V = A + npe * ld;
  double* Jn = V + npe * ld;                   // [LM_CH][np]
  double* rch = Jn + LM_CH * np;               // [LM_CH]
  double* g = rch + LM_CH;                     // [np]
  double* dv = g + np;                         // [np] eigenvalues / projections
  double* del = dv + np;                       // [np]
  double* red = del + np;                      // [NT]
  double* rot = red + NT;              // [npe/2][2]
  double* nzp = rot + npe;                     // [nz]
  double* nod = nzp + nz;                      // [4][np] node work vectors
  double* pen_res = nod + 4 * np;              // [np]
  double* pen_hss = pen_res + np;              // [np*np] (only when regularising)
  float* mod = (float*)(pen_hss + (a.regul ? np * np : 0));   // [12][nz] current model
  float* bum = mod + 12 * nz;                  // [nz]
  float* ftmp = bum + nz;                      // [nz]
  int* izero = (int*)(ftmp + nz);              // [np] zeroed response columns
  int* qtab = izero + np;                      // [ntri + np] (i << 16 | j) of each normal-matrix / gradient sum
  __shared__ double s_sla[9];
  __shared__ int s_accept, s_conv;

  const long ncol = a.ncol;
  // ---- load the current model and the pixel's scalars
  for (int i = t; i < 12 * nz; i += NT) mod[i] = a.atm[(long)i * ncol + pix];
  if (t < 9) s_sla[t] = a.sla ? a.sla[(long)pix * 9 + t] : 0.0;
  // columns whose absolute sum overflows REAL(4) are zeroed (forward.f90:672-678)
  const float* ecur = a.eqv + o_ecur;
  for (int j = t; j < np; j += NT) {
    float s = 0.f;
    for (int f = 0; f < nf; ++f) s += fabsf(ecur[(long)j * nf + f]);
    izero[j] = (s > 3.40282347e38f) ? 1 : 0;
  }
  __syncthreads();

  const bool ext = a.gres != nullptr;
  // ---- NEW_CALCULATE_CHI2 (get_dmodel.f90:960-1032)
  const double* obs = a.obs + o_obs;
  const float* stk = a.stokes + o_stk;
  const double* sd = a.syn_d ? a.syn_d + o_obs : nullptr;
  // current synthesis of row f
  auto syn = [&](int f) -> double {
    if (sd) return sd[f];
    const int b = f / numw, l = f - b * numw;
    return (double)stk[a.rowstk[b] * numw + l];
  };
  double part = 0.0;
  for (int f = t; f < nf; f += NT) {
    const double d = obs[f] - syn(f);
    part += d * d * isig(f);
  }
  double chi = block_sum<NT>(part, red) / inu;
  if (a.regul) {
    if (t == 0) lm_penalties(a, mod, pen_res, pen_hss, nod, nod + np);
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < np; ++i) s += pen_res[i] * 1.0 * 1.0 * pen_res[i];
    chi = chi + s;
  }
  if (t == 0) s_accept = (ext || chi < s_sla[2]) ? 1 : 0;
  __syncthreads();
  const bool accept = s_accept != 0;
  double* bsyn = ext ? nullptr : a.best_syn + o_obs;
  float* ebest = ext ? nullptr : a.eqv_best + o_ebest;
  if (ext) {
    // the caller keeps the best model; the step starts from atm
  } else if (accept) {
    // best := current (model, equivalent response functions, synthesis)
    for (int i = t; i < 12 * nz; i += NT) a.atm_best[(long)i * ncol + pix] = mod[i];
    for (long i = t; i < (long)np * nf; i += NT) ebest[i] = izero[i / nf] ? 0.f : ecur[i];
    for (int f = t; f < nf; f += NT) bsyn[f] = syn(f);
  } else {
    // current := best
    for (int i = t; i < 12 * nz; i += NT) mod[i] = a.atm_best[(long)i * ncol + pix];
  }
  __syncthreads();
  if (!accept && a.regul) {
    if (t == 0) lm_penalties(a, mod, pen_res, pen_hss, nod, nod + np);
  }
  if (t == 0 && ext) {
    s_sla[1] = chi;
    s_sla[8] = a.ext_lambda;
  } else if (t == 0) {
    // UPDATE_LAMBDA (get_dmodel.f90:917-956) with INV_SLA(2,3,9,1,4), factor 0.01
    double chicur = chi, chipre = s_sla[2], lam = s_sla[8], steps = s_sla[0], precyc = s_sla[3];
    if ((chipre - chicur) > 0) {
      if ((chipre - chicur) > 0.01 * chipre) steps = fmax(0.0, steps - 1.0);
      else steps = steps + 1.0;
      chipre = chicur;
      lam = lam / sqrt(10.0);
    } else {
      if (fabs(precyc - 1.0) < 1.0e-2) {
        precyc = 0.0;
        lam = lam * 3.0;
      } else {
        lam = lam * 10.0;
      }
      steps = steps + 1.0;
    }
    lam = fmax(1.0e-2, lam);
    s_sla[0] = steps;
    s_sla[1] = chicur;
    s_sla[2] = chipre;
    s_sla[3] = precyc;
    s_sla[8] = lam;
  }
  __syncthreads();

  // ---- normal matrix and gradient of the normalised problem
  //      J_n = J sqrt(w) / sqrt(nu) * PARAM_NORM,  r = (obs - best) sqrt(w) / sqrt(nu)
  const float* jsrc = accept ? ecur : ebest;
  const double snu = sqrt(inu);
  const int ntri = np * (np + 1) / 2, nq = ntri + np;
  // sums of the normal matrix (upper triangle, row-major) and of the
  // gradient, accumulated in LDS over the chunks in row order
  for (int q = t; q < nq; q += NT) {
    int i, j;
    if (q < ntri) {
      int r = 0, base = 0;
      while (q >= base + (np - r)) {
        base += np - r;
        ++r;
      }
      i = r;
      j = r + (q - base);
    } else {
      i = q - ntri;
      j = 0xffff;                              // gradient entry
    }
    qtab[q] = (i << 16) | j;
  }
  for (int e = t; e < npe * ld; e += NT) A[e] = 0.0;
  for (int i = t; i < np; i += NT) g[i] = 0.0;
  __syncthreads();
  for (int f0 = 0; f0 < nf; f0 += LM_CH) {
    const int nc = min(LM_CH, nf - f0);
    for (int e = t; e < nc * np; e += NT) {
      const int c = e / np, j = e - c * np;
      const int f = f0 + c;
      double v = (accept && izero[j]) ? 0.0 : (double)jsrc[(long)j * nf + f];
      int ip = 0;
      while (j >= a.j0[ip] + a.nj[ip]) ++ip;
      Jn[c * np + j] = v * sqrt(isig(f)) / snu * (1.0 * c_lm_norm[a.par[ip]]);
    }
    for (int c = t; c < nc; c += NT) {
      const int f = f0 + c;
      rch[c] = (ext ? a.gres[o_obs + f] : obs[f] - bsyn[f]) * sqrt(isig(f)) / snu;
    }
    __syncthreads();
    for (int q = t; q < nq; q += NT) {
      const int i = qtab[q] >> 16, j = qtab[q] & 0xffff;
      if (j != 0xffff) {
        double s = A[i * ld + j];
        for (int c = 0; c < nc; ++c) s += Jn[c * np + i] * Jn[c * np + j];
        A[i * ld + j] = s;
      } else {
        double s = g[i];
        for (int c = 0; c < nc; ++c) s += Jn[c * np + i] * rch[c];
        g[i] = s;
      }
    }
    __syncthreads();
  }
  const double lam = s_sla[8];
  if (ext && a.hscale != 1.0) {
    for (int e = t; e < np * np; e += NT) {
      const int i = e / np, j = e - i * np;
      if (j >= i) A[i * ld + j] *= a.hscale;
    }
    __syncthreads();
  }
  // A = H + REG_HSS, diag += diag * lambda; g = J^T r - REG_PEN
  for (int e = t; e < np * np; e += NT) {
    const int i = e / np, j = e - i * np;
    if (j < i) A[i * ld + j] = A[j * ld + i];
  }
  for (int i = t; i < np; i += NT)
    if (a.regul) g[i] -= pen_res[i];
  __syncthreads();
  if (a.regul)
    for (int e = t; e < np * np; e += NT) {
      const int i = e / np, j = e - i * np;
      A[i * ld + j] += pen_hss[e];
    }
  __syncthreads();
  for (int i = t; i < np; i += NT) A[i * ld + i] = A[i * ld + i] + A[i * ld + i] * lam;
  for (int e = t; e < npe * ld; e += NT) {
    const int i = e / ld, j = e - i * ld;
    V[e] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();

  // ---- damped solve (symmetric eigen-decomposition or CG)
  damped_solve<NT>(a.solver, a.svdtol, np, A, V, g, dv, del, red, rot, nod);
  __syncthreads();
  // NaN filter (get_dmodel.f90:806-822)
  if (t == 0) {
    double s = 0.0;
    for (int i = 0; i < np; ++i) s += del[i];
    if (s != s)
      for (int i = 0; i < np; ++i)
        if (del[i] != del[i]) del[i] = 0.0;
  }
  __syncthreads();

  // ---- LIMIT_PERTURBATION -> REARANGE_DA_MODEL -> SORT_DELTA_TO_ATMOSPHERE
  double* itd = nod;            // [np] scaled / limited node perturbation
  double* avg = nod + np;       // [np] node means of the model
  double* sumv = nod + 2 * np;  // [np]
  for (int ip = 0; ip < a.nparam; ++ip) {
    const int p = a.par[ip], j0 = a.j0[ip], nj = a.nj[ip];
    float* col = mod + a.mcol[ip] * nz;
    const double factor = c_lm_norm[p], inf = c_lm_low[p], sup = c_lm_upp[p];
    const double absolute = c_lm_norm[p] * a.mpert;
    for (int k = t; k < nz; k += NT) bum[k] = col[k];
    for (int j = t; j < nj; j += NT) {
      const double* cf = a.coefs + (long)(j0 + j) * nz;
      double s = 0.0;
      for (int k = 0; k < nz; ++k) s = s + (double)col[k] * cf[k];
      const double av = s / a.colsum[j0 + j];
      double lim;
      if (absolute > 0) lim = fmax(absolute, fabs(a.mpert * av));
      else lim = fabs(a.mpert * av);
      if (av == 0) lim = absolute;
      double x = del[j0 + j] * factor;
      if (fabs(x) > lim) x = x / fabs(x) * lim;
      itd[j] = x;
      avg[j] = av;
      sumv[j] = av + x;
      del[j0 + j] = x;
    }
    __syncthreads();
    for (int k = t; k < nz; k += NT) {
      double s = 0.0;
      for (int j = 0; j < nj; ++j) s = s + sumv[j] * a.coefs[(long)(j0 + j) * nz + k];
      nzp[k] = s;
    }
    __syncthreads();
    if (a.node_expansion == 2) {
      for (int k = t; k < nz; k += NT) {
        float o = (float)nzp[k];
        if ((double)o < inf) o = (float)inf;
        if ((double)o > sup) o = (float)sup;
        col[k] = o - col[k];
      }
    } else if (a.node_expansion == 1) {
      for (int k = t; k < nz; k += NT) {
        nzp[k] = nzp[k] - (double)col[k];
        ftmp[k] = (float)nzp[k];
      }
      __syncthreads();
      for (int j = t; j < nj; j += NT) {
        const double* cf = a.coefs + (long)(j0 + j) * nz;
        double s = 0.0;
        for (int k = 0; k < nz; ++k) s = s + (double)ftmp[k] * cf[k];
        itd[j] = s / a.colsum[j0 + j];
        del[j0 + j] = itd[j];
      }
      __syncthreads();
      for (int k = t; k < nz; k += NT) {
        double s = 0.0;
        for (int j = 0; j < nj; ++j) s = s + itd[j] * a.coefs[(long)(j0 + j) * nz + k];
        const float o = col[k] + (float)s;
        if ((double)o < inf) s = inf - (double)col[k];
        if ((double)o > sup) s = sup - (double)col[k];
        col[k] = (float)s;
      }
    }
    __syncthreads();
    if (p == 0 && t == 0) {
      // temperature tail rule in log tau5 (get_dmodel.f90:1970-1984)
      const float* tau = mod + A_TAU * nz;
      const float tllim = -11.5f, tulim = 0.5f;
      for (int k = nz - 1; k >= 0; --k) {
        if (tau[k] < tllim) {
          const float x = fabsf(tau[k] - tllim) / 0.5f;
          col[k] = expf(-(x * x)) * col[k];
        } else if (tau[k] > tulim) {
          if (k + 1 < nz) col[k] = col[k + 1];
        }
      }
    }
    __syncthreads();
    for (int k = t; k < nz; k += NT) col[k] = col[k] + bum[k];
    __syncthreads();
  }

  // ---- write back
  for (int i = t; i < 12 * nz; i += NT) a.atm[(long)i * ncol + pix] = mod[i];
  if (t < 9 && a.sla) a.sla[(long)pix * 9 + t] = s_sla[t];
  if (a.delta)
    for (int i = t; i < np; i += NT) a.delta[(long)pix * np + i] = del[i];
  if (t == 0 && !ext && a.done && s_sla[0] > a.maxsteps) a.done[pix] = 1;
}

// bytes of lm_kernel's work arrays (LDS, or the device-memory slice of one
// pixel) for np free parameters, nz heights and nt threads
inline size_t lm_lds_bytes(int np, int nz, bool regul, int nt) {
  const int npe = np + (np & 1), ld = np + 1;
  size_t d = (size_t)2 * npe * ld + (size_t)LM_CH * np + LM_CH + 3 * np + nt + npe + nz + 4 * np + np +
             (regul ? (size_t)np * np : 0);
  size_t f = (size_t)12 * nz + 2 * nz + np + (size_t)np * (np + 1) / 2 + np;   // floats + ints
  return d * 8 + f * 4;
}

}  // namespace fz
