// Learned initial model for the inversion: a small fully connected network
// (Stokes profiles -> node values of T, Bz, B_perp cos 2phi, B_perp sin 2phi,
// v_z) with BF16 weights and activations and FP32 accumulation on the MFMA
// units.  The reference has no counterpart: its initial model is the input
// model file (invert3d.f90:57-178); the inversion itself is unchanged.
//
//   mlp_gemm_kernel      C = op(A) op(B)^T on v_mfma_f32_16x16x32_bf16, with
//                        the layer epilogues (bias, ReLU, ReLU' mask, BF16
//                        and FP32 stores)
//   mlp_input_kernel     standardised BF16 inputs (optionally gathered by a
//                        permutation of the training set)
//   mlp_loss_kernel      mean squared error on standardised targets and its
//                        gradient
//   mlp_colsum_kernel    bias gradients
//   mlp_adam_kernel      Adam on FP32 master weights, refreshing the BF16 copy
//   mlp_targets_kernel   training targets from a model cube (least-squares
//                        node values through a pseudo-inverse of the node basis)
//   mlp_decode_kernel    network output -> initial model rows (node expansion)
#pragma once

namespace fz {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int MLP_TM = 64;     // output tile rows per workgroup
constexpr int MLP_TN = 64;     // output tile columns per workgroup
constexpr int MLP_TK = 32;     // K step (one MFMA 16x16x32)
constexpr int MLP_LDK = MLP_TK + 8;  // LDS row pitch in bf16 (80 B, 16-B aligned)
constexpr int MLP_MAXNODE = 8;
constexpr int MLP_NQ = 5;      // quantities per node: T, Bz, Bp cos2phi, Bp sin2phi, vz

struct GemmArgs {
  int M, N, K;
  const bf16* A;     // A(m, k) = A[m * sam + k * sak]
  long sam, sak;
  const bf16* B;     // B(n, k) = B[n * sbn + k * sbk]
  long sbn, sbk;
  const float* bias; // [N] added when non-null
  int relu;          // max(0, .) after the bias
  const bf16* mask;  // when non-null the result is zeroed where mask(m, n) <= 0
  long smm, smn;
  float* Cf;         // FP32 result (optional)
  long scm, scn;
  bf16* Cb;          // BF16 result (optional)
  long sbm2, sbn2;
};

// Tile loader: rows r0..r0+63 of an operand with element (r, k) at
// P[r * sr + k * sk] into LDS [64][MLP_LDK], zero outside rows < R, k < K.
// The thread mapping follows the operand's contiguous direction so that the
// global reads of a wave are consecutive; where eight consecutive elements
// along that direction are in range and 16-byte aligned, a thread moves
// them with one 16-byte load (and, along rows, scatters them down a column
// of the tile).
__device__ __forceinline__ bool mlp_al16(const bf16* p) { return ((uintptr_t)p & 15) == 0; }

__device__ __forceinline__ void mlp_load_tile(const bf16* __restrict__ P, long sr, long sk, int R, int K, int r0, int k0,
                                              bf16* __restrict__ S) {
  const int t = threadIdx.x;
  const bf16 z = (bf16)0.0f;
  if (sk == 1) {
    // 4 threads per row, 8 consecutive k each
    const int r = t >> 2, kk = (t & 3) * 8;
    const int gr = r0 + r, gk0 = k0 + kk;
    const bf16* src = P + (long)gr * sr + gk0;
    if (gr < R && gk0 + 8 <= K && mlp_al16(src)) {
      *reinterpret_cast<uint4*>(&S[r * MLP_LDK + kk]) = *reinterpret_cast<const uint4*>(src);
      return;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int gk = gk0 + i;
      S[r * MLP_LDK + kk + i] = (gr < R && gk < K) ? P[(long)gr * sr + gk] : z;
    }
  } else if (sr == 1) {
    // rows contiguous: 8 threads along the 64 rows (8 rows each), 32 k
    const int rr = (t & 7) * 8, kk = t >> 3;
    const int gr0 = r0 + rr, gk = k0 + kk;
    const bf16* src = P + gr0 + (long)gk * sk;
    if (gk < K && gr0 + 8 <= R && mlp_al16(src)) {
      const bf16x8 v = *reinterpret_cast<const bf16x8*>(src);
#pragma unroll
      for (int i = 0; i < 8; ++i) S[(rr + i) * MLP_LDK + kk] = v[i];
      return;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int gr = gr0 + i;
      S[(rr + i) * MLP_LDK + kk] = (gr < R && gk < K) ? P[gr + (long)gk * sk] : z;
    }
  } else {
    // neither direction contiguous: 64 threads along rows, 4 k-groups of 8
    const int r = t & 63, kk = (t >> 6) * 8;
    const int gr = r0 + r;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int gk = k0 + kk + i;
      S[r * MLP_LDK + kk + i] = (gr < R && gk < K) ? P[(long)gr * sr + (long)gk * sk] : z;
    }
  }
}

// 256 threads = 4 waves in a 2x2 arrangement, each wave 32x32 outputs as
// 2x2 MFMA 16x16 blocks.  MFMA operand layout (16x16x32, BF16): lane l holds
// row l%16 of A and column l%16 of B with 8 consecutive k of the lane group
// l/16; the accumulator holds D[4*(l/16) + i][l%16], i = 0..3.
__global__ void __launch_bounds__(256) mlp_gemm_kernel(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) bf16 As[MLP_TM * MLP_LDK];
  __shared__ __attribute__((aligned(16))) bf16 Bs[MLP_TN * MLP_LDK];
  const int m0 = blockIdx.y * MLP_TM, n0 = blockIdx.x * MLP_TN;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int lr = lane & 15, lg = lane >> 4;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < g.K; k0 += MLP_TK) {
    mlp_load_tile(g.A, g.sam, g.sak, g.M, g.K, m0, k0, As);
    mlp_load_tile(g.B, g.sbn, g.sbk, g.N, g.K, n0, k0, Bs);
    __syncthreads();
    bf16x8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      a[i] = *reinterpret_cast<const bf16x8*>(&As[(wm + 16 * i + lr) * MLP_LDK + lg * 8]);
      b[i] = *reinterpret_cast<const bf16x8*>(&Bs[(wn + 16 * i + lr) * MLP_LDK + lg * 8]);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int m = m0 + wm + 16 * i + 4 * lg + q;
        const int n = n0 + wn + 16 * j + lr;
        if (m >= g.M || n >= g.N) continue;
        float v = acc[i][j][q];
        if (g.bias) v += g.bias[n];
        if (g.relu) v = v > 0.f ? v : 0.f;
        if (g.mask && !((float)g.mask[(long)m * g.smm + (long)n * g.smn] > 0.f)) v = 0.f;
        if (g.Cf) g.Cf[(long)m * g.scm + (long)n * g.scn] = v;
        if (g.Cb) g.Cb[(long)m * g.sbm2 + (long)n * g.sbn2] = (bf16)v;
      }
}

// X[m][i] = bf16((x[row][i] - mean[i]) / std[i]), row = idx ? idx[m] : m;
// x is float (xf) or double (xd).
__global__ void mlp_input_kernel(int n, int nin, const float* __restrict__ xf, const double* __restrict__ xd,
                                 const int32_t* __restrict__ idx, const float* __restrict__ mean,
                                 const float* __restrict__ sdev, bf16* __restrict__ X) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)n * nin) return;
  const int m = (int)(t / nin), i = (int)(t - (long)m * nin);
  const long row = idx ? (long)idx[m] : (long)m;
  const float v = xf ? xf[row * nin + i] : (float)xd[row * nin + i];
  X[t] = (bf16)((v - mean[i]) / sdev[i]);
}

// loss partials and dY = 2 (Y - (T - tmean) / tstd) / (n nout), FP32 and BF16
__global__ void __launch_bounds__(256) mlp_loss_kernel(int n, int nout, const float* __restrict__ Y,
                                                       const float* __restrict__ T, const int32_t* __restrict__ idx,
                                                       const float* __restrict__ tmean, const float* __restrict__ tstd,
                                                       float* __restrict__ dYf, bf16* __restrict__ dYb,
                                                       float* __restrict__ partial) {
  __shared__ float red[256];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float e2 = 0.f;
  if (t < (long)n * nout) {
    const int m = (int)(t / nout), o = (int)(t - (long)m * nout);
    const long row = idx ? (long)idx[m] : (long)m;
    const float tn = (T[row * nout + o] - tmean[o]) / tstd[o];
    const float d = Y[t] - tn;
    e2 = d * d;
    const float gd = 2.f * d / ((float)n * (float)nout);
    dYf[t] = gd;
    dYb[t] = (bf16)gd;
  }
  red[threadIdx.x] = e2;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// out[0] = sum(partial[0..np)) / scale, one workgroup
__global__ void __launch_bounds__(256) mlp_reduce_kernel(int np, const float* __restrict__ partial, float scale,
                                                         float* __restrict__ out) {
  __shared__ float red[256];
  float s = 0.f;
  for (int i = threadIdx.x; i < np; i += 256) s += partial[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0] / scale;
}

// g[o] = sum_m dZ[m][o]  (one thread per output column, batch loop)
__global__ void mlp_colsum_kernel(int n, int nout, const float* __restrict__ dZ, float* __restrict__ g) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nout) return;
  float s = 0.f;
  for (int m = 0; m < n; ++m) s += dZ[(long)m * nout + o];
  g[o] = s;
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters; the weight segments
// are copied to BF16 (wb index = parameter index, biases copied too).
__global__ void mlp_adam_kernel(long np, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1,
                                float* __restrict__ m2, bf16* __restrict__ wb, float lr, float b1, float b2, float eps,
                                float c1, float c2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  const float v = w[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
  w[i] = v;
  wb[i] = (bf16)v;
}

// Per-feature mean and standard deviation of x[n][nf] (FP64 sums); a
// feature with no spread gets a unit scale.
__global__ void mlp_stats_kernel(int n, int nf, const float* __restrict__ x, float* __restrict__ mean,
                                 float* __restrict__ sdev) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  double s = 0.0, s2 = 0.0;
  for (long m = 0; m < n; ++m) {
    const double v = x[m * nf + f];
    s += v;
    s2 += v * v;
  }
  const double mu = s / n;
  const double var = fmax(s2 / n - mu * mu, 0.0);
  const double sd = sqrt(var);
  mean[f] = (float)mu;
  sdev[f] = sd > 1e-6 * fmax(fabs(mu), 1e-30) && sd > 0.0 ? (float)sd : 1.0f;
}

__global__ void mlp_to_bf16_kernel(long n, const float* __restrict__ w, bf16* __restrict__ wb) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) wb[i] = (bf16)w[i];
}

// Node values of a model cube atm[12][nz][ncol]: per quantity the least-
// squares fit x_j = sum_k pinv[j][k] q(z_k) through the hat basis; B_perp and
// its azimuth enter as (B_perp cos 2phi, B_perp sin 2phi), invariant under the
// 180-degree ambiguity of Q and U.  t[col][q * nj + j].
__global__ void mlp_targets_kernel(int ncol, int nz, int nj, const float* __restrict__ atm,
                                   const double* __restrict__ pinv, float* __restrict__ t) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  const long cs = (long)nz * ncol;
  for (int j = 0; j < nj; ++j) {
    double T = 0.0, bx = 0.0, by = 0.0, bz = 0.0, vz = 0.0;
    for (int k = 0; k < nz; ++k) {
      const double p = pinv[j * nz + k];
      const long o = (long)k * ncol + c;
      T += p * atm[0 * cs + o];
      bx += p * atm[5 * cs + o];
      by += p * atm[6 * cs + o];
      bz += p * atm[7 * cs + o];
      vz += p * atm[10 * cs + o];
    }
    const double bp = sqrt(bx * bx + by * by);
    const double ph2 = 2.0 * atan2(by, bx);
    float* o = t + (long)c * MLP_NQ * nj;
    o[0 * nj + j] = (float)T;
    o[1 * nj + j] = (float)bz;
    o[2 * nj + j] = (float)(bp * cos(ph2));
    o[3 * nj + j] = (float)(bp * sin(ph2));
    o[4 * nj + j] = (float)vz;
  }
}

// Network output (standardised) -> node values -> heights through the hat
// basis coefs[k][j]; writes T, Bx, By, Bz, vz of atm (other rows untouched).
// T is kept inside the inversion's bounds (2500-18000 K, get_dmodel.f90
// LOWLIM/UPPLIM).
__global__ void mlp_decode_kernel(int ncol, int nz, int nj, const float* __restrict__ y,
                                  const float* __restrict__ tmean, const float* __restrict__ tstd,
                                  const double* __restrict__ coefs, float* __restrict__ atm) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  float q[MLP_NQ][MLP_MAXNODE];
  const int no = MLP_NQ * nj;
  for (int i = 0; i < no; ++i) q[i / nj][i % nj] = y[(long)c * no + i] * tstd[i] + tmean[i];
  float bx[MLP_MAXNODE], by[MLP_MAXNODE];
  for (int j = 0; j < nj; ++j) {
    const float bp = sqrtf(q[2][j] * q[2][j] + q[3][j] * q[3][j]);
    const float ph = 0.5f * atan2f(q[3][j], q[2][j]);
    bx[j] = bp * cosf(ph);
    by[j] = bp * sinf(ph);
  }
  const long cs = (long)nz * ncol;
  for (int k = 0; k < nz; ++k) {
    double T = 0.0, X = 0.0, Y = 0.0, Z = 0.0, V = 0.0;
    for (int j = 0; j < nj; ++j) {
      const double w = coefs[k * nj + j];
      T += w * q[0][j];
      X += w * bx[j];
      Y += w * by[j];
      Z += w * q[1][j];
      V += w * q[4][j];
    }
    const long o = (long)k * ncol + c;
    atm[0 * cs + o] = (float)fmin(fmax(T, 2500.0), 18000.0);
    atm[5 * cs + o] = (float)X;
    atm[6 * cs + o] = (float)Y;
    atm[7 * cs + o] = (float)Z;
    atm[10 * cs + o] = (float)V;
  }
}

}  // namespace fz
