// FP32 per-(column, wavelength) radiative transfer pieces: Faraday-Voigt
// profiles, the Zeeman-split propagation matrix and its derivatives, the
// analytic evolution operator and its directional derivatives.
#pragma once
#include <hip/hip_runtime.h>

namespace fz {

struct v3 {
  float x, y, z;
};
__device__ __forceinline__ v3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ v3 operator*(float s, v3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ v3 operator-(v3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ v3 cross(v3 a, v3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Voigt H(a,u) and Faraday-Voigt L(a,u) = Re, Im of w(u + i a) with the
// Hui, Armstrong & Wray (1978) 7/7 rational form in z = a - i|u|; L is odd
// in u (voigt.f:17-98).  a > 0 always here (radiative damping is positive).
// Both polynomials have real coefficients, so they are evaluated at the
// complex point with Knuth's two-FMA recurrence on the quadratic
// z^2 - 2 Re(z) z + |z|^2 (TAOCP vol. 2, 4.6.4): P(z) = p1 z + p0 after
// p1 <- p0 + r p1, p0 <- c_k - s p1 with r = 2 Re z, s = |z|^2.
__device__ __forceinline__ void voigt(float a, float u, float& H, float& F) {
  const float A0 = 122.607931777104326f, A1 = 214.382388694706425f, A2 = 181.928533092181549f,
              A3 = 93.155580458138441f, A4 = 30.180142196210589f, A5 = 5.912626209773153f,
              A6 = .564189583562615f;
  const float B0 = 122.60793177387535f, B1 = 352.730625110963558f, B2 = 457.334478783897737f,
              B3 = 348.703917719495792f, B4 = 170.354001821091472f, B5 = 53.992906912940207f,
              B6 = 10.479857114260399f;
  const float v = fabsf(u);
  const float x = a, y = -v;
  const float r = 2.f * x, s = x * x + y * y;
  float p1, p0, t;
  // numerator A6 z^6 + ... + A0
  p1 = A6;
  p0 = A5;
  t = p1; p1 = fmaf(r, t, p0); p0 = fmaf(-s, t, A4);
  t = p1; p1 = fmaf(r, t, p0); p0 = fmaf(-s, t, A3);
  t = p1; p1 = fmaf(r, t, p0); p0 = fmaf(-s, t, A2);
  t = p1; p1 = fmaf(r, t, p0); p0 = fmaf(-s, t, A1);
  t = p1; p1 = fmaf(r, t, p0); p0 = fmaf(-s, t, A0);
  const float nr = fmaf(p1, x, p0), ni = p1 * y;
  // denominator z^7 + B6 z^6 + ... + B0
  float q1 = 1.f, q0 = B6;
  t = q1; q1 = fmaf(r, t, q0); q0 = fmaf(-s, t, B5);
  t = q1; q1 = fmaf(r, t, q0); q0 = fmaf(-s, t, B4);
  t = q1; q1 = fmaf(r, t, q0); q0 = fmaf(-s, t, B3);
  t = q1; q1 = fmaf(r, t, q0); q0 = fmaf(-s, t, B2);
  t = q1; q1 = fmaf(r, t, q0); q0 = fmaf(-s, t, B1);
  t = q1; q1 = fmaf(r, t, q0); q0 = fmaf(-s, t, B0);
  const float dr = fmaf(q1, x, q0), di = q1 * y;
  const float inv = __builtin_amdgcn_rcpf(fmaf(dr, dr, di * di));
  H = fmaf(nr, dr, ni * di) * inv;
  const float im = fmaf(ni, dr, -nr * di) * inv;
  F = (u < 0.f) ? -im : im;
}

// Field geometry of one height (absorption_matrix.f90:1800-1871): |B|,
// cos and sin of the inclination, 1 +- cos^2, cos 2phi, sin 2phi.  Formed
// once per (height, column) with the opacities and stored in the per-height
// record, so the radiative transfer reads it instead of forming it in
// every lane of the column.
struct Geo {
  float B, cg, sg, sgf, cgf, c2p, s2p;
};

__device__ __forceinline__ Geo geometry(float bx, float by, float bz) {
  // absorption_matrix.f90:1800-1871
  Geo g;
  const float bt2 = bx * bx + by * by, b2 = bt2 + bz * bz;
  g.B = sqrtf(b2);
  if (g.B > 0.f) {
    g.cg = bz / g.B;
    g.sg = sqrtf(bt2) / g.B;
    g.cgf = 1.f + bz * bz / b2;
    g.sgf = 1.f - bz * bz / b2;
  } else {
    g.cg = 0.f; g.sg = 0.f; g.cgf = 1.f; g.sgf = 1.f;
  }
  if (bt2 > 0.f) {
    g.c2p = (bx * bx - by * by) / bt2;
    g.s2p = 2.f * bx * by / bt2;
  } else {
    g.c2p = 0.f; g.s2p = 0.f;
  }
  return g;
}

// 4-vector helpers
struct V4 {
  float v[4];
};

// (P(a,w) v) for the matrix [[0,a1,a2,a3],[a1,0,w3,-w2],[a2,-w3,0,w1],[a3,w2,-w1,0]]
__device__ __forceinline__ void pmul(v3 a, v3 w, const float* v, float* o) {
  o[0] = a.x * v[1] + a.y * v[2] + a.z * v[3];
  o[1] = a.x * v[0] + w.z * v[2] - w.y * v[3];
  o[2] = a.y * v[0] - w.z * v[1] + w.x * v[3];
  o[3] = a.z * v[0] + w.y * v[1] - w.x * v[2];
}

// Evolution operator of a constant-property layer, O = exp(-K dz), in the
// closed form of Landi Degl'Innocenti & Landi Degl'Innocenti (1985)
// (rtesolver.f90:269-473).  Written for FP32: all propagation-matrix elements
// are scaled by dz first (optical thicknesses), the two eigen-parameters are
// formed without cancellation (lambda1*lambda2 = |eta.rho|), and the
// exp(-eta_I dz) factor is folded into cosh/sinh so nothing overflows.
//   O = A*1 + P(alpha, omega) + D4 * M4raw
//
// Near-degenerate polarisation eigenvalues.  The closed form divides by
// th = sqrt(d^2 + 4 (e.r)^2) and by l1, l2; when th is small the terms of
// the directional derivatives grow like 1/th^1.5 and cancel, and in FP32 the
// error grows like eps (s0/th)^2 (s0 = (e.e + r.r)/2).  The four coefficient
// functions the operator needs are entire functions of d and q = (e.r)^2:
// with x, y the roots of z^2 - d z - q (x = l1^2, y = -l2^2),
//   A  e^ei = (cosh l1 + cos l2)/2            = 1/2 sum_n P_n / (2n)!
//   D4 e^ei = (cosh l1 - cos l2)/th           = sum_{n>=1} H_{n-1} / (2n)!
//   ga e^ei = (l1 sinh l1 + l2 sin l2)/th     = sum_n H_n / (2n+1)!
//   gb e^ei = e.r (sinh l1/l1 - sin l2/l2)/th = e.r sum_{n>=1} H_{n-1} / (2n+1)!
// where P_n = x^n + y^n and H_n = sum_{i<=n} x^i y^(n-i) both follow
// F_n = d F_{n-1} + q F_{n-2} (P_0 = 2, H_0 = 1, F_1 = d), so the series and
// their d- and q-derivatives need no square root and no division.  For
// th < TH_SER the operator and its derivatives use these series (NSER terms,
// truncation below 1e-11 at th = 2); the closed form is kept above, where
// th is not small against s0 unless the layer is optically thick and
// e^-ei removes it.
constexpr float TH_SER = 2.0f;
constexpr int NSER = 9;

// Fields of one branch of the operator in a shared array (a lane takes one
// branch; sharing keeps the register footprint of the struct that of one):
// closed form: 1/l1 or 1/l2, 1/th, l1, l2, sign(e.r), C2, B3, and the
// e^-ei cosh/sinh/cos/sin products; series: the d- and q-derivatives of the
// A, D4, ga and gb series, the gb series and e^-ei.
enum { W_IL = 0, W_ITH, W_L1, W_L2, W_SIG, W_C2, W_B3, W_CHE, W_SHE, W_CE, W_SE, NW };
enum { W_AD = 0, W_AQ, W_DD, W_DQ, W_GD, W_GQ, W_BD, W_BQ, W_SB, W_EX };

struct Evol {
  float O[4][4];
  v3 e, r;
  float ei, th, d, er;
  float A, D4, s0;
  // al = -(ga e + gb r) and om = -(gb e - ga r)
  float ga, gb;
  bool ser;
  float w[NW];
};

// the four coefficient series of the comment above (times e^ei) and, with
// DER, their partial derivatives in d and q
template <bool DER>
__device__ __forceinline__ void evol_series(float d, float q, float& A0, float& D0, float& G0, float& B0, Evol& E) {
  // 1/k! for k = 0 .. 2 NSER
  constexpr float IF[2 * NSER + 1] = {1.f, 1.f, 1.f / 2.f, 1.f / 6.f, 1.f / 24.f, 1.f / 120.f, 1.f / 720.f,
                                      1.f / 5040.f, 1.f / 40320.f, 1.f / 362880.f, 1.f / 3628800.f,
                                      1.f / 39916800.f, 1.f / 479001600.f, 1.f / 6227020800.f,
                                      1.f / 87178291200.f, 1.f / 1307674368000.f, 1.f / 20922789888000.f,
                                      1.f / 355687428096000.f, 1.f / 6402373705728000.f};
  // F_n and its d-, q-derivatives for H (h*) and P (p*), n = 0 .. NSER - 1
  float h0 = 1.f, h1 = d, hd0 = 0.f, hd1 = 1.f, hq0 = 0.f, hq1 = 0.f;
  float p0 = 2.f, p1 = d, pd0 = 0.f, pd1 = 1.f, pq0 = 0.f, pq1 = 0.f;
  A0 = 0.5f * p0;
  D0 = 0.f;
  G0 = h0 * IF[1];
  B0 = 0.f;
  float Ad = 0.f, Aq = 0.f, Dd = 0.f, Dq = 0.f, Gd = 0.f, Gq = 0.f, Bd = 0.f, Bq = 0.f;
#pragma unroll
  for (int n = 1; n < NSER; ++n) {
    // terms with H_{n-1} (held in h0 ...) and P_n, H_n (held in p1, h1 ...)
    A0 += 0.5f * p1 * IF[2 * n];
    D0 += h0 * IF[2 * n];
    G0 += h1 * IF[2 * n + 1];
    B0 += h0 * IF[2 * n + 1];
    if (DER) {
      Ad += 0.5f * pd1 * IF[2 * n];
      Aq += 0.5f * pq1 * IF[2 * n];
      Dd += hd0 * IF[2 * n];
      Dq += hq0 * IF[2 * n];
      Gd += hd1 * IF[2 * n + 1];
      Gq += hq1 * IF[2 * n + 1];
      Bd += hd0 * IF[2 * n + 1];
      Bq += hq0 * IF[2 * n + 1];
    }
    // advance to n + 1
    const float h2 = d * h1 + q * h0, p2 = d * p1 + q * p0;
    if (DER) {
      const float hd2 = h1 + d * hd1 + q * hd0, hq2 = d * hq1 + h0 + q * hq0;
      const float pd2 = p1 + d * pd1 + q * pd0, pq2 = d * pq1 + p0 + q * pq0;
      hd0 = hd1;
      hd1 = hd2;
      hq0 = hq1;
      hq1 = hq2;
      pd0 = pd1;
      pd1 = pd2;
      pq0 = pq1;
      pq1 = pq2;
    }
    h0 = h1;
    h1 = h2;
    p0 = p1;
    p1 = p2;
  }
  if (DER) {
    E.w[W_AD] = Ad;
    E.w[W_AQ] = Aq;
    E.w[W_DD] = Dd;
    E.w[W_DQ] = Dq;
    E.w[W_GD] = Gd;
    E.w[W_GQ] = Gq;
    E.w[W_BD] = Bd;
    E.w[W_BQ] = Bq;
  }
}

template <bool DER>
__device__ __forceinline__ void evol_build(const float x7[7], float dz, Evol& E) {
  E.ei = x7[0] * dz;
  E.e = mk(x7[1] * dz, x7[2] * dz, x7[3] * dz);
  E.r = mk(x7[4] * dz, x7[5] * dz, x7[6] * dz);
  const float e2 = dot(E.e, E.e), r2 = dot(E.r, E.r);
  E.er = dot(E.e, E.r);
  E.d = e2 - r2;
  // hypot without its scaling unless the squares could overflow
  E.th = (fabsf(E.d) < 1.0e18f && fabsf(E.er) < 1.0e18f) ? sqrtf(E.d * E.d + 4.0f * E.er * E.er)
                                                          : hypotf(E.d, 2.0f * E.er);
  E.ser = E.th < TH_SER;
  if (E.ser) {
    float A0, D0, G0, B0;
    evol_series<DER>(E.d, E.er * E.er, A0, D0, G0, B0, E);
    E.w[W_EX] = expf(-E.ei);
    E.w[W_SB] = B0;
    E.A = E.w[W_EX] * A0;
    E.D4 = E.w[W_EX] * D0;
    E.ga = E.w[W_EX] * G0;
    E.gb = E.w[W_EX] * (E.er * B0);
  } else {
    float l1, l2;
    if (E.d >= 0.f) {
      l1 = sqrtf(0.5f * (E.th + E.d));
      E.w[W_IL] = l1 > 0.f ? 1.0f / l1 : 0.f;
      l2 = fabsf(E.er) * E.w[W_IL];
    } else {
      l2 = sqrtf(0.5f * (E.th - E.d));
      E.w[W_IL] = 1.0f / l2;
      l1 = fabsf(E.er) * E.w[W_IL];
    }
    E.w[W_L1] = l1;
    E.w[W_L2] = l2;
    E.w[W_SIG] = E.er > 0.f ? 1.f : (E.er < 0.f ? -1.f : 0.f);
    if (l1 < 40.f) {
      // cosh and sinh from one exponential; sinh by its Taylor series below
      // 0.5 (truncation 1e-11 relative), where E - 1/E would cancel
      const float ex = expf(-E.ei);
      const float El = expf(l1), iEl = 1.0f / El;
      const float x2 = l1 * l1;
      const float shs = l1 * (1.f + x2 * (1.f / 6.f) * (1.f + x2 * (1.f / 20.f) * (1.f + x2 * (1.f / 42.f) * (1.f + x2 * (1.f / 72.f)))));
      E.w[W_CHE] = ex * (0.5f * (El + iEl));
      E.w[W_SHE] = ex * (l1 < 0.5f ? shs : 0.5f * (El - iEl));
    } else {
      const float p = expf(l1 - E.ei), m = expf(-l1 - E.ei);
      E.w[W_CHE] = 0.5f * (p + m);
      E.w[W_SHE] = 0.5f * (p - m);
    }
    float sn, cs;
    sincosf(l2, &sn, &cs);
    const float ex = expf(-E.ei);
    E.w[W_CE] = ex * cs;
    E.w[W_SE] = ex * sn;
    E.A = 0.5f * (E.w[W_CHE] + E.w[W_CE]);
    const float ith = E.th > 0.f ? 1.0f / E.th : 0.f;
    E.w[W_ITH] = ith;
    E.w[W_C2] = E.w[W_SE] * ith;
    E.w[W_B3] = E.w[W_SHE] * ith;
    E.D4 = (E.w[W_CHE] - E.w[W_CE]) * ith;
    // With a2 = l2 e - s l1 r, a3 = l1 e + s l2 r, w2 = -(s l1 e + l2 r) and
    // w3 = s l2 e - l1 r, the vectors al = -(C2 a2 + B3 a3) and
    // om = -(C2 w2 + B3 w3) take two scalars: ga = C2 l2 + B3 l1 and
    // gb = s (B3 l2 - C2 l1).
    const float s = E.w[W_SIG];
    E.ga = E.w[W_C2] * l2 + E.w[W_B3] * l1;
    E.gb = s * (E.w[W_B3] * l2 - E.w[W_C2] * l1);
  }
  const v3 al = -(E.ga * E.e + E.gb * E.r);
  const v3 om = -(E.gb * E.e - E.ga * E.r);
  E.s0 = 0.5f * (e2 + r2);
  const v3 c = cross(E.e, E.r);
  const float D4 = E.D4;
  const float ev[3] = {E.e.x, E.e.y, E.e.z}, rv[3] = {E.r.x, E.r.y, E.r.z};
  const float alv[3] = {al.x, al.y, al.z}, cv[3] = {c.x, c.y, c.z};
  E.O[0][0] = E.A + D4 * E.s0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    E.O[0][k + 1] = alv[k] + D4 * cv[k];
    E.O[k + 1][0] = alv[k] - D4 * cv[k];
    E.O[k + 1][k + 1] = E.A + D4 * (ev[k] * ev[k] + rv[k] * rv[k] - E.s0);
  }
  E.O[1][2] = om.z + D4 * (ev[0] * ev[1] + rv[0] * rv[1]);
  E.O[2][1] = -om.z + D4 * (ev[0] * ev[1] + rv[0] * rv[1]);
  E.O[1][3] = -om.y + D4 * (ev[0] * ev[2] + rv[0] * rv[2]);
  E.O[3][1] = om.y + D4 * (ev[0] * ev[2] + rv[0] * rv[2]);
  E.O[2][3] = om.x + D4 * (ev[1] * ev[2] + rv[1] * rv[2]);
  E.O[3][2] = -om.x + D4 * (ev[1] * ev[2] + rv[1] * rv[2]);
}

// M4raw v and its pieces (c = e x r)
__device__ __forceinline__ void m4mul(v3 e, v3 r, float s0, const float* v, float* o) {
  const v3 c = cross(e, r);
  const v3 vp = mk(v[1], v[2], v[3]);
  const float ev = dot(e, vp), rv = dot(r, vp);
  o[0] = s0 * v[0] + dot(c, vp);
  o[1] = -c.x * v[0] + e.x * ev + r.x * rv - s0 * vp.x;
  o[2] = -c.y * v[0] + e.y * ev + r.y * rv - s0 * vp.y;
  o[3] = -c.z * v[0] + e.z * ev + r.z * rv - s0 * vp.z;
}

// Directional derivative (dO) v for a change (dx7 * dz) of the propagation
// matrix; the reference forms the six partial derivatives of O explicitly
// (rtesolver.f90:476-1002); here one direction costs one pass of this
// function.
__device__ __forceinline__ void evol_dir(const Evol& E, const float dx7[7], float dz, const float* v,
                                         const float* m4v, float* out) {
  const float dei = dx7[0] * dz;
  const v3 de = mk(dx7[1] * dz, dx7[2] * dz, dx7[3] * dz);
  const v3 dr = mk(dx7[4] * dz, dx7[5] * dz, dx7[6] * dz);
  const float dd = 2.f * (dot(E.e, de) - dot(E.r, dr));
  const float der = dot(de, E.r) + dot(E.e, dr);
  float dA, dD4, dga, dgb;
  if (E.ser) {
    const float dq = 2.f * E.er * der;
    dA = -dei * E.A + E.w[W_EX] * (E.w[W_AD] * dd + E.w[W_AQ] * dq);
    dD4 = -dei * E.D4 + E.w[W_EX] * (E.w[W_DD] * dd + E.w[W_DQ] * dq);
    dga = -dei * E.ga + E.w[W_EX] * (E.w[W_GD] * dd + E.w[W_GQ] * dq);
    dgb = -dei * E.gb + E.w[W_EX] * (der * E.w[W_SB] + E.er * (E.w[W_BD] * dd + E.w[W_BQ] * dq));
  } else {
    float dth = 0.f, dl1 = 0.f, dl2 = 0.f, ith = 0.f;
    if (E.th > 0.f) {
      ith = E.w[W_ITH];
      dth = (E.d * dd + 4.f * E.er * der) * ith;
      if (E.d >= 0.f) {
        dl1 = (dth + dd) * (0.25f * E.w[W_IL]);
        dl2 = (E.w[W_SIG] * der - E.w[W_L2] * dl1) * E.w[W_IL];
      } else {
        dl2 = (dth - dd) * (0.25f * E.w[W_IL]);
        dl1 = (E.w[W_SIG] * der - E.w[W_L1] * dl2) * E.w[W_IL];
      }
    }
    dA = -dei * E.A + 0.5f * (E.w[W_SHE] * dl1 - E.w[W_SE] * dl2);
    const float dC2 = -dei * E.w[W_C2] + (E.w[W_CE] * dl2 - E.w[W_C2] * dth) * ith;
    const float dB3 = -dei * E.w[W_B3] + (E.w[W_CHE] * dl1 - E.w[W_B3] * dth) * ith;
    dD4 = -dei * E.D4 + (E.w[W_SHE] * dl1 + E.w[W_SE] * dl2 - E.D4 * dth) * ith;
    dga = dC2 * E.w[W_L2] + E.w[W_C2] * dl2 + dB3 * E.w[W_L1] + E.w[W_B3] * dl1;
    dgb = E.w[W_SIG] * (dB3 * E.w[W_L2] + E.w[W_B3] * dl2 - dC2 * E.w[W_L1] - E.w[W_C2] * dl1);
  }
  // d(al) = -(dga e + ga de + dgb r + gb dr), d(om) = -(dgb e + gb de - dga r - ga dr)
  const v3 dal = -(dga * E.e + E.ga * de + dgb * E.r + E.gb * dr);
  const v3 dom = -(dgb * E.e + E.gb * de - dga * E.r - E.ga * dr);
  float p[4];
  pmul(dal, dom, v, p);
  // d(M4raw) v
  const v3 vp = mk(v[1], v[2], v[3]);
  const float ds0 = dot(E.e, de) + dot(E.r, dr);
  const v3 dc = cross(de, E.r) + cross(E.e, dr);
  const float ev = dot(E.e, vp), dev = dot(de, vp), rv = dot(E.r, vp), drv = dot(dr, vp);
  float q[4];
  q[0] = ds0 * v[0] + dot(dc, vp);
  q[1] = -dc.x * v[0] + de.x * ev + E.e.x * dev + dr.x * rv + E.r.x * drv - ds0 * vp.x;
  q[2] = -dc.y * v[0] + de.y * ev + E.e.y * dev + dr.y * rv + E.r.y * drv - ds0 * vp.y;
  q[3] = -dc.z * v[0] + de.z * ev + E.e.z * dev + dr.z * rv + E.r.z * drv - ds0 * vp.z;
#pragma unroll
  for (int i = 0; i < 4; ++i) out[i] = dA * v[i] + p[i] + dD4 * m4v[i] + E.D4 * q[i];
}

}  // namespace fz
