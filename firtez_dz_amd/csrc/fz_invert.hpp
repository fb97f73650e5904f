// Cycle driver kernels of the FIRTEZ-dz inversion (INVERT3D, invert3d.f90):
//   inv_init_kernel   : per-cycle scalars INV_MAS(1:9) and AM_I_DONE (:286-298)
//   smooth_axis_kernel: one axis of SMOOTHING_B's 1-2-1 filter
//                       (model_spatial_extension.f90:351-491), FP64 like the
//                       reference's DBLE(...) work array
//   stats_kernel      : MAKE_DECISION's sums (:820-859) in one workgroup
//   copy_best_kernel  : UPDATE_BEST_TO_CURRENT / UPDATE_CURRENT_TO_BEST of the
//                       model (forward.f90:310-359)
#pragma once

namespace fz {

__global__ void inv_init_kernel(int ncol, double* __restrict__ sla, int32_t* __restrict__ done) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  double* s = sla + (long)c * 9;
  s[0] = 0.0;      // steps without improvement
  s[1] = 1.0e29;   // current chi^2
  s[2] = 1.0e29;   // previous (best) chi^2
  s[3] = 0.0;      // PRECYC
  s[4] = s[5] = s[6] = s[7] = 0.0;
  s[8] = 1.0e1;    // lambda
  done[c] = 0;
}

// a[k][y][x] (double, col = y + ny*x) along one axis: interior points
// (a[+1] + a[-1] + 2 a) / 4 from the unfiltered values in src.
__global__ void smooth_axis_kernel(const double* __restrict__ src, double* __restrict__ dst, int nz, int ny, int nx,
                                   int axis) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)nz * ny * nx;
  if (i >= n) return;
  const long ncol = (long)ny * nx;
  const int k = (int)(i / ncol);
  const long col = i - (long)k * ncol;
  const int x = (int)(col / ny), y = (int)(col - (long)x * ny);
  double v = src[i];
  if (axis == 0 && x > 0 && x < nx - 1) v = (src[i + ny] + src[i - ny] + src[i] * 2.0) / 4.0;
  if (axis == 1 && y > 0 && y < ny - 1) v = (src[i + 1] + src[i - 1] + src[i] * 2.0) / 4.0;
  if (axis == 2 && k > 0 && k < nz - 1) v = (src[i + ncol] + src[i - ncol] + src[i] * 2.0) / 4.0;
  dst[i] = v;
}

__global__ void to_double_kernel(const float* __restrict__ src, double* __restrict__ dst, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (double)src[i];
}

// REAL(ARRAY) into the current and the best model row
__global__ void from_double_kernel(const double* __restrict__ src, float* __restrict__ cur, float* __restrict__ best,
                                   long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = (float)src[i];
  cur[i] = v;
  best[i] = v;
}

// out[0] = sum of steps, out[1] = pixels done (one workgroup of 256)
__global__ void __launch_bounds__(256) stats_kernel(int ncol, const double* __restrict__ sla,
                                                    const int32_t* __restrict__ done, double* __restrict__ out) {
  __shared__ double r0[256], r1[256];
  const int t = threadIdx.x;
  double s = 0.0, d = 0.0;
  for (int c = t; c < ncol; c += 256) {
    s += sla[(long)c * 9];
    d += done[c] ? 1.0 : 0.0;
  }
  r0[t] = s;
  r1[t] = d;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) {
      r0[t] += r0[t + h];
      r1[t] += r1[t + h];
    }
    __syncthreads();
  }
  if (t == 0) {
    out[0] = r0[0];
    out[1] = r1[0];
  }
}

// Active pixels (AM_I_DONE == 0) in ascending order: idx[0..count), one
// workgroup of 1024 threads, each scanning a contiguous chunk of the flags.
__global__ void __launch_bounds__(1024) compact_kernel(int ncol, const int32_t* __restrict__ done,
                                                       int32_t* __restrict__ idx, int32_t* __restrict__ count) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int chunk = (ncol + 1023) / 1024;
  const int b = t * chunk, e = min(ncol, b + chunk);
  int n = 0;
  for (int i = b; i < e; ++i) n += done[i] == 0;
  part[t] = n;
  __syncthreads();
  // inclusive scan (Hillis-Steele) of the per-thread counts
  for (int off = 1; off < 1024; off <<= 1) {
    const int v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int o = part[t] - n;
  for (int i = b; i < e; ++i)
    if (done[i] == 0) idx[o++] = i;
  if (t == 1023) count[0] = part[1023];
}

// rows x [nrow][ncol] <-> compacted xc [nrow][na] through idx
__global__ void gather_cols_kernel(const float* __restrict__ x, float* __restrict__ xc,
                                   const int32_t* __restrict__ idx, int na, int ncol, int nrow) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)na * nrow) return;
  const int r = (int)(i / na), j = (int)(i - (long)r * na);
  xc[i] = x[(long)r * ncol + idx[j]];
}

__global__ void scatter_cols_kernel(const float* __restrict__ xc, float* __restrict__ x,
                                    const int32_t* __restrict__ idx, int na, int ncol, int nrow) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)na * nrow) return;
  const int r = (int)(i / na), j = (int)(i - (long)r * na);
  x[(long)r * ncol + idx[j]] = xc[i];
}

__global__ void copy_f32_kernel(const float* __restrict__ src, float* __restrict__ dst, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

}  // namespace fz
