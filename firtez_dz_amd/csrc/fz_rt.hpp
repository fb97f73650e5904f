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
struct Evol {
  float O[4][4];
  v3 e, r;
  float ei, th, d, er, l1, l2, sig;
  float A, C2, B3, D4, che, she, ce, se, s0;
  // al = -(ga e + gb r) and om = -(gb e - ga r)
  float ga, gb;
};

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
  float l1, l2;
  if (E.d >= 0.f) {
    l1 = sqrtf(0.5f * (E.th + E.d));
    l2 = l1 > 0.f ? fabsf(E.er) / l1 : 0.f;
  } else {
    l2 = sqrtf(0.5f * (E.th - E.d));
    l1 = fabsf(E.er) / l2;
  }
  E.l1 = l1;
  E.l2 = l2;
  E.sig = E.er > 0.f ? 1.f : (E.er < 0.f ? -1.f : 0.f);
  if (l1 < 40.f) {
    // cosh and sinh from one exponential; sinh by its Taylor series below
    // 0.5 (truncation 1e-11 relative), where E - 1/E would cancel
    const float ex = expf(-E.ei);
    const float El = expf(l1), iEl = 1.0f / El;
    const float x2 = l1 * l1;
    const float shs = l1 * (1.f + x2 * (1.f / 6.f) * (1.f + x2 * (1.f / 20.f) * (1.f + x2 * (1.f / 42.f) * (1.f + x2 * (1.f / 72.f)))));
    E.che = ex * (0.5f * (El + iEl));
    E.she = ex * (l1 < 0.5f ? shs : 0.5f * (El - iEl));
  } else {
    const float p = expf(l1 - E.ei), m = expf(-l1 - E.ei);
    E.che = 0.5f * (p + m);
    E.she = 0.5f * (p - m);
  }
  float sn, cs;
  sincosf(l2, &sn, &cs);
  const float ex = expf(-E.ei);
  E.ce = ex * cs;
  E.se = ex * sn;
  E.A = 0.5f * (E.che + E.ce);
  const float ith = E.th > 0.f ? 1.0f / E.th : 0.f;
  E.C2 = E.se * ith;
  E.B3 = E.she * ith;
  E.D4 = (E.che - E.ce) * ith;
  // With a2 = l2 e - s l1 r, a3 = l1 e + s l2 r, w2 = -(s l1 e + l2 r) and
  // w3 = s l2 e - l1 r, the vectors al = -(C2 a2 + B3 a3) and
  // om = -(C2 w2 + B3 w3) take two scalars: ga = C2 l2 + B3 l1 and
  // gb = s (B3 l2 - C2 l1).
  const float s = E.sig;
  E.ga = E.C2 * l2 + E.B3 * l1;
  E.gb = s * (E.B3 * l2 - E.C2 * l1);
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
  float dth = 0.f, dl1 = 0.f, dl2 = 0.f, ith = 0.f;
  if (E.th > 0.f) {
    ith = 1.0f / E.th;
    dth = (E.d * dd + 4.f * E.er * der) * ith;
    if (E.d >= 0.f) {
      dl1 = (dth + dd) / (4.f * E.l1);
      dl2 = (E.sig * der - E.l2 * dl1) / E.l1;
    } else {
      dl2 = (dth - dd) / (4.f * E.l2);
      dl1 = (E.sig * der - E.l1 * dl2) / E.l2;
    }
  }
  const float dA = -dei * E.A + 0.5f * (E.she * dl1 - E.se * dl2);
  const float dC2 = -dei * E.C2 + (E.ce * dl2 - E.C2 * dth) * ith;
  const float dB3 = -dei * E.B3 + (E.che * dl1 - E.B3 * dth) * ith;
  const float dD4 = -dei * E.D4 + (E.she * dl1 + E.se * dl2 - E.D4 * dth) * ith;
  // d(al) = -(dga e + ga de + dgb r + gb dr), d(om) = -(dgb e + gb de - dga r - ga dr)
  const float dga = dC2 * E.l2 + E.C2 * dl2 + dB3 * E.l1 + E.B3 * dl1;
  const float dgb = E.sig * (dB3 * E.l2 + E.B3 * dl2 - dC2 * E.l1 - E.C2 * dl1);
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
