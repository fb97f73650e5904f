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

namespace fz {

// SET_TAU's assist initialisers (invert3d.f90:528-700), applied to the start
// model of the first cycle when neither Pg nor P0 is inverted.  They read
// the observations of the first spectral region (the itz = PIXEL_END(1) -
// PIXEL_INI(1) + 1 first entries of each Stokes block of OBS3D):
//   B   (Q, U, V inverted): the map-mean profile I0, its centred differences
//       I1 = dI0, I2 = dI1 (each normalised to unit absolute sum); per pixel
//       V.I1 -> Bz, (Q.I2, U.I2) -> transverse field and azimuth, each
//       normalised by its maximum over the map, 1000 G scale;
//   V   the pixel's first block projected on I1 minus its projection on I0
//       times I1.I0 -> v_z, normalised by the map maximum, 2 km/s scale;
//   T   the first block's mean -> brightness temperature of the Planck
//       function normalised at 6390 K; T shifted so that T at height
//       ceiling(nz/3) takes it;
//   P   Pg scaled so that Pg at height ceiling(nz/3) is 1.333e5.
// FP64 sums; the model rows keep the reference's REAL(4) arithmetic.
enum { ASSIST_B = 1, ASSIST_V = 2, ASSIST_T = 4, ASSIST_P = 8 };

struct AssistArgs {
  int ncol, nz, nfreq, numw, itz, nstk, flags, kz;   // kz: height ceiling(nz/3) - 1
  double lk_hc;                                      // lambda0 k / (h c) of the first region's line
  const double* obs;                                 // [ncol][nfreq]
  float* atm;                                        // [12][nz][ncol]
  double* vec;                                       // [3][itz] + [1] (I0, I1, I2, I1.I0)
  double* pix;                                       // [ncol][5]: V.I1, Q.I2, U.I2, V-projection, mean I
  double* mx;                                        // [3]: max |V.I1|, max transverse, max |V-projection|
};

// I0[w] = sum over the map of the first block (one workgroup per wavelength)
__global__ void __launch_bounds__(256) assist_sum_kernel(AssistArgs a) {
  __shared__ double red[256];
  const int w = blockIdx.x, t = threadIdx.x;
  double s = 0.0;
  for (int p = t; p < a.ncol; p += 256) s += a.obs[(long)p * a.nfreq + w];
  red[t] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) red[t] += red[t + h];
    __syncthreads();
  }
  if (t == 0) a.vec[w] = red[0];
}

// normalisations and centred differences of I0 (one thread: itz is a few hundred)
__global__ void assist_vec_kernel(AssistArgs a) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int n = a.itz;
  double* v0 = a.vec;
  double* v1 = v0 + n;
  double* v2 = v1 + n;
  auto normalise = [&](double* v) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += fabs(v[i]);
    for (int i = 0; i < n; ++i) v[i] = v[i] / s;
  };
  auto diff = [&](const double* v, double* d) {
    for (int i = 1; i < n - 1; ++i) d[i] = (v[i + 1] - v[i - 1]) / 2.0;
    d[0] = d[1];
    d[n - 1] = d[n - 2];
  };
  normalise(v0);
  diff(v0, v1);
  normalise(v1);
  diff(v1, v2);
  normalise(v2);
  double kol = 0.0;
  for (int i = 0; i < n; ++i) kol += v1[i] * v0[i];
  v2[n] = kol;
}

__global__ void assist_pix_kernel(AssistArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.ncol) return;
  const int n = a.itz;
  const double* v0 = a.vec;
  const double* v1 = v0 + n;
  const double* v2 = v1 + n;
  const double* o = a.obs + (long)p * a.nfreq;
  double s0 = 0.0, s1 = 0.0, sq = 0.0, su = 0.0, sv = 0.0, si = 0.0;
  for (int w = 0; w < n; ++w) {
    s0 += o[w] * v0[w];
    s1 += o[w] * v1[w];
    si += o[w];
    if (a.nstk == 4) {
      sq += o[a.numw + w] * v2[w];
      su += o[2 * a.numw + w] * v2[w];
      sv += o[3 * a.numw + w] * v1[w];
    }
  }
  double* q = a.pix + (long)p * 5;
  q[0] = sv;
  q[1] = sq;
  q[2] = su;
  q[3] = s1 - s0 * v2[n];
  q[4] = si / (double)n;
}

// the three map maxima (one workgroup)
__global__ void __launch_bounds__(256) assist_max_kernel(AssistArgs a) {
  __shared__ double red[3][256];
  const int t = threadIdx.x;
  double m0 = 0.0, m1 = 0.0, m2 = 0.0;
  for (int p = t; p < a.ncol; p += 256) {
    const double* q = a.pix + (long)p * 5;
    m0 = fmax(m0, fabs(q[0]));
    m1 = fmax(m1, sqrt(q[1] * q[1] + q[2] * q[2]));
    m2 = fmax(m2, fabs(q[3]));
  }
  red[0][t] = m0;
  red[1][t] = m1;
  red[2][t] = m2;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h)
      for (int i = 0; i < 3; ++i) red[i][t] = fmax(red[i][t], red[i][t + h]);
    __syncthreads();
  }
  if (t < 3) a.mx[t] = red[t][0];
}

__global__ void assist_apply_kernel(AssistArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.ncol) return;
  const long cs = a.ncol, ps = (long)a.nz * a.ncol;
  float* col = a.atm + p;
  const double* q = a.pix + (long)p * 5;
  if ((a.flags & ASSIST_B) && a.nstk == 4) {
    // a map without polarisation leaves 0 (the reference divides 0 by 0)
    const double t3 = a.mx[0] > 0.0 ? q[0] / a.mx[0] : 0.0;
    const double t4 = a.mx[1] > 0.0 ? sqrt(q[1] * q[1] + q[2] * q[2]) / a.mx[1] : 0.0;
    const double atn2 = atan2(q[2], q[1]) / 2.0;
    // -REAL(SIN(ATN2))*1000.E0*TAU1TEM: single-precision product, then times the FP64 factor
    const float bx = (float)((double)(-(float)sin(atn2) * 1000.0f) * t4);
    const float by = (float)((double)((float)cos(atn2) * 1000.0f) * t4);
    const float bz = (float)t3 * (-1000.0f);
    for (int k = 0; k < a.nz; ++k) {
      col[A_BX * ps + k * cs] = bx;
      col[A_BY * ps + k * cs] = by;
      col[A_BZ * ps + k * cs] = bz;
    }
  }
  if (a.flags & ASSIST_V) {
    const float vz = (float)(a.mx[2] > 0.0 ? q[3] / a.mx[2] : 0.0) * (-2.0e5f);
    for (int k = 0; k < a.nz; ++k) col[A_VZ * ps + k * cs] = vz;
  }
  if (a.flags & (ASSIST_T | ASSIST_P)) {
    for (int k = 0; k < a.nz; ++k) col[A_TAU * ps + k * cs] = (k == a.kz) ? 0.0f : 1.0f;
  }
  if (a.flags & ASSIST_T) {
    // brightness temperature of the mean first-block intensity, Planck
    // function normalised at 6390 K
    const double tb = 1.0 / (a.lk_hc * log(1.0 + 1.0 / q[4] * (exp(1.0 / 6.39e3 / a.lk_hc) - 1.0)));
    const double dt = tb - (double)col[A_T * ps + a.kz * cs];
    for (int k = 0; k < a.nz; ++k) col[A_T * ps + k * cs] = (float)((double)col[A_T * ps + k * cs] + dt);
  }
  if (a.flags & ASSIST_P) {
    const float pz = col[A_PG * ps + a.kz * cs];
    for (int k = 0; k < a.nz; ++k) col[A_PG * ps + k * cs] = col[A_PG * ps + k * cs] * 1.333e5f / pz;
  }
}

}  // namespace fz
