// Spatial PSF of the coupled inversion: a periodic 2-D convolution of maps on
// the (x, y) grid, and its adjoint.
//
// The reference reads the PSF (READ_PSF, pre_post_duties.f90:488-555), embeds
// the central (2 YS + 1) x (2 XS + 1) part in an NY x NX array centred at
// (NY/2 + 1, NX/2 + 1), shifts the centre to (1, 1) and multiplies Fourier
// transforms (SPATIAL_CONV_INIT :559-588, SMOOTH_MODEL invert3d.f90:412-488):
// a circular convolution,
//   out(y, x) = sum_{dy, dx} PSF(Y0 + dy, X0 + dx) in(y - dy mod NY, x - dx mod NX).
// Here the same sum is formed directly (a 31 x 31 PSF is 961 taps per
// output; 1.04 ms for 448 maps of 256 x 256 on MI355X), which needs no FFT
// library and keeps the result exact to FP32 summation.  The adjoint (the transpose
// of the convolution, used for the gradient of the coupled chi^2) is the
// correlation out(y, x) = sum PSF(Y0 + dy, X0 + dx) in(y + dy, x + dx).
//
// Layout: maps channel-major, map m at in[m * npix + col] with
// col = y + ny * x (the cube's column order), so a map is an nx x ny image
// with y contiguous.  One workgroup of four waves computes a 32 (x) x 64 (y)
// output tile of one map: the lanes of a wave take 64 consecutive y, and
// each lane 8 consecutive x.  The (32 + NTX - 1) x (64 + 2 ry) input tile is
// staged in LDS with the periodic wrap applied at load.  For each of the
// 2 ry + 1 tap rows along y, a lane reads the strip of 8 + NTX - 1 inputs
// along x that its 8 outputs need into registers (conflict-free: the 64
// lanes read 64 consecutive words) and applies the NTX taps of that row,
// wave-uniform scalar loads, to all 8 outputs: 8 NTX FMAs per
// 8 + NTX - 1 LDS reads, which keeps the FP32 VALU, not LDS bandwidth, the
// bound.  NTX is the x-tap count 2 rx + 1 rounded up to a multiple of 8;
// the padding taps are zero and the padding rows of the tile hold wrapped
// (finite) inputs.
#pragma once

namespace fz {

constexpr int PSF_OX = 8;                  // outputs per lane along x
constexpr int PSF_TX = 4 * PSF_OX;         // output tile along x (four waves)
constexpr int PSF_TY = 64;                 // output tile along y (the lanes of a wave)
constexpr int PSF_RMAX = 31;               // largest tap radius (63 x 63 PSF)

struct PsfArgs {
  int nx, ny, nmaps, rx, ry;
  long npix;
  // [2 ry + 1][NTX] taps in the tile's order: taps[u * NTX + t] multiplies
  // tile row (x) offset t and column (y) offset u of the output's window
  // (the PSF reversed for the convolution, as given for the adjoint)
  const float* taps;
  const float* in;
  float* out;
};

template <int NTX>
__global__ void __launch_bounds__(256) psf_conv_kernel(PsfArgs a) {
  extern __shared__ float tile[];
  constexpr int NW = PSF_OX + NTX - 1;     // register strip along x
  constexpr int ROWS = PSF_TX + NTX - 1;   // tile rows (x)
  const int m = blockIdx.z;
  const int x0 = blockIdx.y * PSF_TX, y0 = blockIdx.x * PSF_TY;
  const int cols = PSF_TY + 2 * a.ry;      // tile columns (y)
  const float* src = a.in + (long)m * a.npix;
  // tile[i][j] = in(x0 - rx + i, y0 - ry + j), periodic in x and y
  for (int e = threadIdx.x; e < ROWS * cols; e += 256) {
    const int i = e / cols, j = e - i * cols;
    int x = (x0 - a.rx + i) % a.nx;
    if (x < 0) x += a.nx;
    int y = (y0 - a.ry + j) % a.ny;
    if (y < 0) y += a.ny;
    tile[e] = src[(long)x * a.ny + y];
  }
  __syncthreads();
  const int ty = threadIdx.x & 63, xw = (threadIdx.x >> 6) * PSF_OX;
  float acc[PSF_OX];
#pragma unroll
  for (int i = 0; i < PSF_OX; ++i) acc[i] = 0.f;
  const int nu = 2 * a.ry + 1;
  for (int u = 0; u < nu; ++u) {
    // output (xw + i, ty) reads tile row xw + i + t, column ty + u
    const float* colp = tile + xw * cols + ty + u;
    float v[NW];
#pragma unroll
    for (int j = 0; j < NW; ++j) v[j] = colp[j * cols];
    const float* k = a.taps + u * NTX;
#pragma unroll
    for (int t = 0; t < NTX; ++t) {
      const float kt = k[t];
#pragma unroll
      for (int i = 0; i < PSF_OX; ++i) acc[i] = fmaf(kt, v[i + t], acc[i]);
    }
  }
  const int y = y0 + ty;
  if (y >= a.ny) return;
  float* dst = a.out + (long)m * a.npix + y;
#pragma unroll
  for (int i = 0; i < PSF_OX; ++i) {
    const int x = x0 + xw + i;
    if (x < a.nx) dst[(long)x * a.ny] = acc[i];
  }
}

// 32 x 32 transpose through LDS: out[c][r] = in[r][c] for an R x C matrix
// (in row stride C, out row stride R), converting the element type
template <class Tin, class Tout>
__global__ void __launch_bounds__(256) transpose_kernel(const Tin* __restrict__ in, Tout* __restrict__ out, long R,
                                                        long C) {
  __shared__ Tin t[32][33];
  const long r0 = (long)blockIdx.y * 32, c0 = (long)blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8 threads
  for (int k = ty; k < 32; k += 8) {
    const long r = r0 + k, c = c0 + tx;
    if (r < R && c < C) t[k][tx] = in[r * C + c];
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const long c = c0 + k, r = r0 + tx;
    if (r < R && c < C) out[c * R + r] = (Tout)t[tx][k];
  }
}

}  // namespace fz

namespace fz {

// The coupled residual: from the PSF-convolved synthesis maps conv
// ([4 numw][npix], channel-major) take the inverted Stokes rows, and per
// pixel write syn_c and r = obs - syn_c ([npix][nfreq]) and its chi^2 (the
// weighted sum over the rows / inu).  One thread per pixel.
struct CoupledArgs {
  int npix, nfreq, numw, nstk;
  int rowstk[4];
  double inu;
  const float* conv;
  const double* obs;
  const double* isig;   // [nfreq]
  double* syn_c;
  double* res;
  double* chi;          // [npix]
};

__global__ void __launch_bounds__(256) coupled_resid_kernel(CoupledArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.npix) return;
  double s = 0.0;
  for (int f = 0; f < a.nfreq; ++f) {
    const int b = f / a.numw, l = f - b * a.numw;
    const double sy = (double)a.conv[(long)(a.rowstk[b] * a.numw + l) * a.npix + p];
    const double r = a.obs[(long)p * a.nfreq + f] - sy;
    a.syn_c[(long)p * a.nfreq + f] = sy;
    a.res[(long)p * a.nfreq + f] = r;
    s += r * r * a.isig[f];
  }
  a.chi[p] = s / a.inu;
}

// sum of n doubles by one workgroup (fixed order for a given n)
__global__ void __launch_bounds__(256) sum_kernel(const double* __restrict__ x, int n, double* __restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += x[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

}  // namespace fz
