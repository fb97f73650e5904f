// GET_PERTURBATION and HESSIAN (lm.f90:30-39, 137-204) as device kernels.
//
//   gram_kernel       : the normal matrix H = J^T J and the gradient J^T DIFF
//                       of one problem in one launch: the augmented matrix
//                       X = [J | DIFF] (NF rows, NP+1 columns) is contracted
//                       with itself on the FP64 matrix cores
//                       (v_mfma_f64_16x16x4_f64), 32x32 output tiles per
//                       workgroup over the upper triangle, NF streamed
//                       through LDS in chunks of 16 rows.
//   pert_solve_kernel : TERM1 = H + REG_HSS with the damped diagonal
//                       (GET_TERM1, lm.f90:43-57), TERM2 = J^T DIFF - REG_PEN,
//                       and the damped solve (truncated pseudo-inverse or
//                       CG, damped_solve) by one workgroup; work arrays in
//                       LDS up to LM_MAXNP free parameters, beyond in device
//                       memory.
// FP64 throughout; the products of the contraction are exact FP64 products
// summed in the matrix core's order (the reference's MATMUL sums in row
// order), so H agrees with a sequential sum to a few ulp.
#pragma once

namespace fz {

constexpr int GR_TILE = 32;      // output tile (rows and columns) per workgroup
constexpr int GR_KC = 16;        // rows of X per LDS chunk
constexpr int GR_THREADS = 256;  // four waves, one 16x16 block each

typedef double gr_dbl4 __attribute__((ext_vector_type(4)));

struct GramArgs {
  int np, nf;
  const double* x;   // column p < np of X at x[p*sp + f*sf]
  long sp, sf;
  const double* y;   // column np of X (DIFF), or null
  double* gram;      // [np][np+1] row-major: G[i][j] for j >= i
};

// v_mfma_f64_16x16x4_f64 operand and result layout (64 lanes, l = lane):
//   A (16x4): lane l holds A[l % 16][l / 16]
//   B (4x16): lane l holds B[l / 16][l % 16]
//   D (16x16): lane l holds D[4 r + l / 16][l % 16] in element r = 0..3
__global__ void __launch_bounds__(GR_THREADS) gram_kernel(GramArgs a) {
  const int ti = blockIdx.y, tj = blockIdx.x;
  if (tj < ti) return;   // lower triangle: mirrored by the reader
  __shared__ double xs[2][GR_TILE][GR_KC + 1];
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  const int i0 = ti * GR_TILE, j0 = tj * GR_TILE;
  const int wi = (w >> 1) * 16, wj = (w & 1) * 16;
  gr_dbl4 acc = {0.0, 0.0, 0.0, 0.0};
  for (int f0 = 0; f0 < a.nf; f0 += GR_KC) {
    // stage rows f0 .. f0+15 of columns i0 .. i0+31 (s = 0) and j0 .. j0+31
    // (s = 1), consecutive threads along X's contiguous direction
    for (int e = t; e < 2 * GR_TILE * GR_KC; e += GR_THREADS) {
      const int s = e / (GR_TILE * GR_KC), r = e - s * (GR_TILE * GR_KC);
      int c, ff;
      if (a.sf == 1) {
        c = r / GR_KC;
        ff = r - c * GR_KC;
      } else {
        ff = r / GR_TILE;
        c = r - ff * GR_TILE;
      }
      const int col = (s == 0 ? i0 : j0) + c, f = f0 + ff;
      double v = 0.0;
      if (f < a.nf) {
        if (col < a.np) v = a.x[(long)col * a.sp + (long)f * a.sf];
        else if (col == a.np && a.y) v = a.y[f];
      }
      xs[s][c][ff] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GR_KC; kk += 4) {
      const double av = xs[0][wi + (l & 15)][kk + (l >> 4)];
      const double bv = xs[1][wj + (l & 15)][kk + (l >> 4)];
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  const int ncolx = a.np + (a.y ? 1 : 0);
  const int j = j0 + wj + (l & 15);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = i0 + wi + 4 * r + (l >> 4);
    if (i < a.np && j < ncolx && j >= i) a.gram[(long)i * (a.np + 1) + j] = acc[r];
  }
}

struct PertArgs {
  int np, dtype;
  double lambda, svdtol;
  const double* gram;     // [np][np+1] from gram_kernel (upper triangle, gradient in column np)
  const double* reg_pen;  // REG_PEN(np), or null
  const double* reg_hss;  // REG_HSS(np,np) column-major, or null
  double* delta;          // DELTA(np)
  double* work;           // work arrays in device memory (more than LM_MAXNP free parameters)
};

// doubles of pert_solve_kernel's work arrays for np free parameters and nt threads
inline size_t pert_work_doubles(int np, int nt) {
  const size_t npe = np + (np & 1), ld = np + 1;
  return 2 * npe * ld + 3 * (size_t)np + nt + npe + 4 * (size_t)np;
}

template <bool BIG>
__global__ void __launch_bounds__(BIG ? LM_BIG_THREADS : LM_THREADS) pert_solve_kernel(PertArgs a) {
  constexpr int NT = BIG ? LM_BIG_THREADS : LM_THREADS;
  extern __shared__ double smem_dyn[];
  double* smem = BIG ? a.work : smem_dyn;
  const int t = threadIdx.x;
  const int np = a.np, ld = np + 1, npe = np + (np & 1);
  double* A = smem;               // [npe][ld]
  double* V = A + npe * ld;       // [npe][ld]
  double* g = V + npe * ld;       // [np]
  double* dv = g + np;            // [np]
  double* del = dv + np;          // [np]
  double* red = del + np;         // [NT]
  double* rot = red + NT;         // [npe]
  double* nod = rot + npe;        // [4][np]
  const double* G = a.gram;
  // TERM1(i,j): H + REG_HSS, diagonal times (1 + lambda) (GET_TERM1)
  auto term1 = [&](int i, int j) -> double {
    double v = G[(long)min(i, j) * ld + max(i, j)];
    if (a.reg_hss) v += a.reg_hss[(long)j * np + i];
    if (i == j) v = v + v * a.lambda;
    return v;
  };
  for (int e = t; e < npe * ld; e += NT) {
    const int i = e / ld, j = e - i * ld;
    double v = 0.0;
    if (i < np && j < np) {
      // solver 1 works on the symmetric part (TERM1 is symmetric up to
      // REG_HSS); the CG products read A[j][i] as TERM1(i, j)
      v = a.dtype == 1 ? 0.5 * (term1(i, j) + term1(j, i)) : term1(j, i);
    }
    A[e] = v;
    V[e] = (i == j) ? 1.0 : 0.0;
  }
  for (int i = t; i < np; i += NT) g[i] = G[(long)i * ld + np] - (a.reg_pen ? a.reg_pen[i] : 0.0);
  __syncthreads();
  damped_solve<NT>(a.dtype, a.svdtol, np, A, V, g, dv, del, red, rot, nod);
  __syncthreads();
  for (int i = t; i < np; i += NT) a.delta[i] = del[i];
}

}  // namespace fz
