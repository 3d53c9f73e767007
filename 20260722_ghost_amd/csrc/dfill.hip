// Analytic shape derivative of the fill, contracted with the adjoint weights.
//
// The adjoint of a real far-field objective J needs, for every generatrix node
// radius rho_q, dJ/drho_q = Re sum_m sum Q_m (.) dA_m/drho_q (+ the right-hand
// side and far-field terms), with Q_m = [-Lambda X^T | Lambda] fixed.  The
// fill is linear in its azimuth samples: every far block is
//   A_m(e, f) = sum_{p in e, q in f} sum_xi c(xi) k(p, q, xi) w_m(p, q, xi)
// (bor_solver.py:130-154, 293-345, 348-370; the modal transforms of
// bor_kernels.py:132-167, 273-331), so the sum over modes of Q_m against the
// blocks moves into the azimuth domain:
//   sum_m Q_m A_m = sum_{p, q, xi} c(xi) k(p, q, xi) Phi(p, q, xi),
// with Phi built from the mode series sum_m Q_m {cos, sin}(m xi) of the
// element pair's 2 x 2 node blocks.  The derivative of that sum along the
// four radii of an element pair (test nodes e, e + 1, source nodes f, f + 1)
// is carried by forward-mode dual numbers (D4: a value and four partials)
// through the point positions, weights, tangents, basis derivatives, the
// samples k and, on near pairs, the point pair's own sinh-graded azimuth rule
// (bor_kernels.py:216-246), so the result is the exact derivative of the
// discrete operator the forward fill builds: no finite step, no cancellation.
//
// Far pairs (|e - f| >= 2): one work-group per ordered pair.  The Q series
// are evaluated once per azimuth sample for the pair (shared by its 16 point
// pairs) and staged in LDS.  Near pairs (|e - f| <= 1): per cell point pair
// and rule node, with the series and their xi derivatives (the rule nodes
// move with the geometry).  The mass term rides on the diagonal near pairs.
// The right-hand-side and far-field projection terms differentiate the
// Bessel terms J_n(k rho sin theta) through J_n' = (J_{n-1} - J_{n+1}) / 2.
#include "bor_common.hpp"

namespace ghost {

// ─────────────────────────────────────────────────────────────────────────────
// dual numbers: value + 4 directional derivatives
// ─────────────────────────────────────────────────────────────────────────────

struct D4 {
  double v, d[4];
};
__device__ __forceinline__ D4 dconst(double v) { return D4{v, {0.0, 0.0, 0.0, 0.0}}; }
__device__ __forceinline__ D4 dvar(double v, double d0, double d1, double d2, double d3) {
  return D4{v, {d0, d1, d2, d3}};
}
__device__ __forceinline__ D4 operator+(const D4& a, const D4& b) {
  D4 r; r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < 4; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
__device__ __forceinline__ D4 operator-(const D4& a, const D4& b) {
  D4 r; r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < 4; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
__device__ __forceinline__ D4 operator-(const D4& a) {
  D4 r; r.v = -a.v;
#pragma unroll
  for (int i = 0; i < 4; ++i) r.d[i] = -a.d[i];
  return r;
}
__device__ __forceinline__ D4 operator*(const D4& a, const D4& b) {
  D4 r; r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < 4; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
__device__ __forceinline__ D4 operator*(double s, const D4& a) {
  D4 r; r.v = s * a.v;
#pragma unroll
  for (int i = 0; i < 4; ++i) r.d[i] = s * a.d[i];
  return r;
}
__device__ __forceinline__ D4 operator*(const D4& a, double s) { return s * a; }
__device__ __forceinline__ D4 operator+(const D4& a, double s) { D4 r = a; r.v += s; return r; }
__device__ __forceinline__ D4 operator+(double s, const D4& a) { D4 r = a; r.v += s; return r; }
__device__ __forceinline__ D4 operator-(double s, const D4& a) { D4 r = -a; r.v += s; return r; }
__device__ __forceinline__ D4 operator-(const D4& a, double s) { D4 r = a; r.v -= s; return r; }
// f(a) with f(a.v) = fv and f'(a.v) = fp
__device__ __forceinline__ D4 dchain(const D4& a, double fv, double fp) {
  D4 r; r.v = fv;
#pragma unroll
  for (int i = 0; i < 4; ++i) r.d[i] = fp * a.d[i];
  return r;
}
__device__ __forceinline__ D4 drcp(const D4& a) { const double r = 1.0 / a.v; return dchain(a, r, -r * r); }
__device__ __forceinline__ D4 operator/(const D4& a, const D4& b) { return a * drcp(b); }
__device__ __forceinline__ D4 dsqrt(const D4& a) { const double s = sqrt(a.v); return dchain(a, s, 0.5 / s); }
__device__ __forceinline__ D4 dsinh(const D4& a) { return dchain(a, sinh(a.v), cosh(a.v)); }
__device__ __forceinline__ D4 dcosh(const D4& a) { return dchain(a, cosh(a.v), sinh(a.v)); }
__device__ __forceinline__ D4 dasinh(const D4& a) { return dchain(a, asinh(a.v), 1.0 / sqrt(1.0 + a.v * a.v)); }
__device__ __forceinline__ D4 dasin(const D4& a) { return dchain(a, asin(a.v), 1.0 / sqrt(1.0 - a.v * a.v)); }
__device__ __forceinline__ void dsincos(const D4& a, D4* s, D4* c) {
  double sv, cv;
  sincos(a.v, &sv, &cv);
  *s = dchain(a, sv, cv);
  *c = dchain(a, cv, -sv);
}
// the smaller of two duals (the branch the value takes)
__device__ __forceinline__ D4 dmin(const D4& a, const D4& b) { return a.v <= b.v ? a : b; }
__device__ __forceinline__ D4 dmin(const D4& a, double b) { return a.v <= b ? a : dconst(b); }
__device__ __forceinline__ D4 dmax(const D4& a, double b) { return a.v >= b ? a : dconst(b); }

// complex dual
struct C4 {
  D4 re, im;
};
__device__ __forceinline__ C4 cdz() { return C4{dconst(0.0), dconst(0.0)}; }
// dual real x complex constant
__device__ __forceinline__ C4 rc(const D4& a, double2 c) { return C4{c.x * a, c.y * a}; }
__device__ __forceinline__ C4 operator+(const C4& a, const C4& b) { return C4{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ C4 operator-(const C4& a, const C4& b) { return C4{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ C4 operator*(const D4& s, const C4& a) { return C4{s * a.re, s * a.im}; }
__device__ __forceinline__ C4 operator*(double s, const C4& a) { return C4{s * a.re, s * a.im}; }
__device__ __forceinline__ C4 operator*(const C4& a, const C4& b) {
  return C4{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// j a
__device__ __forceinline__ C4 cj(const C4& a) { return C4{-a.im, a.re}; }
// a + s c (dual real s, complex constant c)
__device__ __forceinline__ C4 cadd_rc(const C4& a, const D4& s, double2 c) { return C4{a.re + c.x * s, a.im + c.y * s}; }
// complex constant as a dual
__device__ __forceinline__ C4 cconst(double2 c) { return C4{dconst(c.x), dconst(c.y)}; }

__device__ __forceinline__ double2 c2add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 c2scale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double2 c2mj(double2 a) { return make_double2(a.y, -a.x); }   // -j a

// ─────────────────────────────────────────────────────────────────────────────
// element-pair geometry as duals
// ─────────────────────────────────────────────────────────────────────────────

// Element e of the generatrix with its two node radii as the dual directions
// sa and sa + 1 (0/1: the test element, 2/3: the source element).
struct DElem {
  D4 r0, dr, L, tr, tz;
  double z0, dz;
};
__device__ __forceinline__ DElem delem(const double* nodes, int e, int sa) {
  DElem E;
  const double r0 = nodes[2 * e], z0 = nodes[2 * e + 1], r1 = nodes[2 * e + 2], z1 = nodes[2 * e + 3];
  E.r0 = dconst(r0);
  E.r0.d[sa] = 1.0;
  D4 r1d = dconst(r1);
  r1d.d[sa + 1] = 1.0;
  E.dr = r1d - E.r0;
  E.z0 = z0;
  E.dz = z1 - z0;
  E.L = dsqrt(E.dr * E.dr + E.dz * E.dz);
  const D4 iL = drcp(E.L);
  E.tr = E.dr * iL;
  E.tz = E.dz * iL;
  return E;
}

// A point at local coordinate s of an element: rho, z, the two basis
// functions T (constant in s) and their derivatives D = d(rho T)/dt.
struct DPoint {
  D4 rho, D0, D1;
  double z, T0, T1;
};
__device__ __forceinline__ DPoint dpoint(const DElem& E, double s) {
  DPoint P;
  P.rho = E.r0 + s * E.dr;
  P.z = E.z0 + s * E.dz;
  P.T0 = 1.0 - s;
  P.T1 = s;
  const D4 iL = drcp(E.L);
  P.D0 = ((1.0 - s) * E.dr - P.rho) * iL;
  P.D1 = (s * E.dr + P.rho) * iL;
  return P;
}

// ─────────────────────────────────────────────────────────────────────────────
// Q blocks of an element pair and their azimuth series
// ─────────────────────────────────────────────────────────────────────────────

// The adjoint weights Q in the augmented system layout: FP64 (FP64 handles)
// or FP32 (mixed handles, whose Q comes from the FP32 GEMM).
struct QView {
  const double2* q64;
  const float2* q32;
  __device__ __forceinline__ double2 operator[](size_t i) const {
    if (q64) return q64[i];
    const float2 v = q32[i];
    return make_double2((double)v.x, (double)v.y);
  }
};

// Row of node a's basis of component c (0: J_t, 1: J_phi) in a system, or -1.
__device__ __forceinline__ int sys_row(const SysInfo& s, int c, int a) {
  if (c == 0) { const int r = a - s.t0; return (r >= 0 && r < s.nt) ? r : -1; }
  const int r = a - s.f0;
  return (r >= 0 && r < s.n - s.nt) ? s.nt + r : -1;
}

// Stage the element pair's Q blocks as mode series coefficients:
// Pq[4 c + 2 a + b][m], m = 0..M, for block c = tt, tf, ft, ff (rows J_t /
// J_phi of test node e + a, columns J_t / J_phi of source node f + b):
// tt and ff keep P+_m = Q_m + Q_-m (P+_0 = Q_0), tf and ft P-_m = Q_m - Q_-m.
// Then sum_m Q_m cos(m xi) = sum_m P+_m cos(m xi), sum_m Q_m sin(m xi) =
// sum_m P-_m sin(m xi), sum_m m Q_m cos(m xi) = sum_m m P-_m cos(m xi) and
// sum_m m^2 Q_m cos(m xi) = sum_m m^2 P+_m cos(m xi).
__device__ void stage_q(const SysInfo* sys, QView Q, int M, int e, int f, double2* Pq) {
  const int mp1 = M + 1;
  for (int t = threadIdx.x; t < 16 * mp1; t += blockDim.x) {
    const int ent = t / mp1, m = t % mp1;
    const int c = ent >> 2, a = (ent >> 1) & 1, b = ent & 1;
    const int rc = c >> 1, cc = c & 1;   // row / column component
    auto qv = [&](int mm) -> double2 {
      const SysInfo s = sys[mm + M];
      const int r = sys_row(s, rc, e + a), col = sys_row(s, cc, f + b);
      if (r < 0 || col < 0) return make_double2(0.0, 0.0);
      return Q[s.a_off + (size_t)r * s.ld + col];
    };
    double2 v;
    if (m == 0) {
      v = (c == 0 || c == 3) ? qv(0) : make_double2(0.0, 0.0);
    } else {
      const double2 qp = qv(m), qm = qv(-m);
      v = (c == 0 || c == 3) ? make_double2(qp.x + qm.x, qp.y + qm.y) : make_double2(qp.x - qm.x, qp.y - qm.y);
    }
    Pq[t] = v;
  }
}

// The seven series of one entry ab at one xi (constant), indices:
// 0 Ct = sum P+tt cos, 1 Sf = sum P-tf sin, 2 Mf = sum m P-tf cos,
// 3 Sg = sum P-ft sin, 4 Mg = sum m P-ft cos, 5 Cf = sum P+ff cos,
// 6 Nf = sum m^2 P+ff cos.  Only 0, 1, 3, 5 when efie is false.
template <bool EFIE>
__device__ __forceinline__ void qseries(const double2* Pq, int M, int ab, double cx, double sx, double2 (&F)[7]) {
#pragma unroll
  for (int i = 0; i < 7; ++i) F[i] = make_double2(0.0, 0.0);
  const double2* ptt = Pq + (0 * 4 + ab) * (M + 1);
  const double2* ptf = Pq + (1 * 4 + ab) * (M + 1);
  const double2* pft = Pq + (2 * 4 + ab) * (M + 1);
  const double2* pff = Pq + (3 * 4 + ab) * (M + 1);
  double c = 1.0, s = 0.0, cp = cx, sp = -sx;   // cos / sin of m xi and (m - 1) xi
  const double tc = 2.0 * cx;
  // two orders per step: the next order's LDS loads are in flight while this
  // order's sums are formed
#pragma unroll 2
  for (int m = 0; m <= M; ++m) {
    const double2 a = ptt[m], b = ptf[m], g = pft[m], d = pff[m];
    F[0].x += a.x * c; F[0].y += a.y * c;
    F[1].x += b.x * s; F[1].y += b.y * s;
    F[3].x += g.x * s; F[3].y += g.y * s;
    F[5].x += d.x * c; F[5].y += d.y * c;
    if (EFIE) {
      const double mc = (double)m * c;
      F[2].x += b.x * mc; F[2].y += b.y * mc;
      F[4].x += g.x * mc; F[4].y += g.y * mc;
      const double m2c = (double)m * mc;
      F[6].x += d.x * m2c; F[6].y += d.y * m2c;
    }
    const double cn = tc * c - cp, sn = tc * s - sp;
    cp = c; sp = s;
    c = cn; s = sn;
  }
}

// ─────────────────────────────────────────────────────────────────────────────
// the contraction at one azimuth sample
// ─────────────────────────────────────────────────────────────────────────────

// Per point pair: duals of the test point p and source point q and of their
// elements' tangents, and the pair weights.
struct DPair {
  D4 rp, rq, tpr, tpz, tqr, tqz;
  D4 Dp0, Dp1, Dq0, Dq1;
  D4 rr, ww;          // rho_p w_p rho_q w_q, w_p w_q
  double dz, Tp0, Tp1, Tq0, Tq1;
};

// Series values at one sample, contracted over the entries with the basis
// weights: constants TT(F) = sum_ab Tp_a Tq_b F_ab for F = Ct, Sf, Sg, Cf, Nf,
// duals DD(Ct) = sum Dp_a Dq_b Ct_ab, DT(Mf) = sum Dp_a Tq_b Mf_ab,
// TD(Mg) = sum Tp_a Dq_b Mg_ab.
struct QAtXi {
  double2 Ct[4], Sf[4], Mf[4], Sg[4], Mg[4], Cf[4], Nf[4];
};

// EFIE weight Phi (no sample): with A1 = tr_p tr_q cos xi + tz_p tz_q,
// Phi = rr [A1 TT(Ct) + cos xi TT(Cf) - j sin xi (tr_p TT(Sf) - tr_q TT(Sg))]
//     - ww / k^2 [DD(Ct) + TT(Nf) + j (DT(Mf) - TD(Mg))]
// (the modal EFIE blocks ztt, ztf, zft, zff of bor_solver.py:130-154 with
// Gc_m = sum c g cos(m xi) cos xi, Gs_m = -j sum c g sin(m xi) sin xi).
__device__ __forceinline__ C4 efie_phi(const DPair& P, const double2* Ct, const double2* Sf, const double2* Mf,
                                       const double2* Sg, const double2* Mg, const double2* Cf, const double2* Nf,
                                       double cx, double sx, double ik2) {
  const double Tp[2] = {P.Tp0, P.Tp1}, Tq[2] = {P.Tq0, P.Tq1};
  const D4* Dp[2] = {&P.Dp0, &P.Dp1};
  const D4* Dq[2] = {&P.Dq0, &P.Dq1};
  double2 TTc = make_double2(0, 0), TTs = TTc, TTg = TTc, TTf = TTc, TTn = TTc;
  C4 DD = cdz(), DT = cdz(), TD = cdz();
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int ab = 2 * a + b;
      const double tt = Tp[a] * Tq[b];
      TTc = c2add(TTc, c2scale(Ct[ab], tt));
      TTs = c2add(TTs, c2scale(Sf[ab], tt));
      TTg = c2add(TTg, c2scale(Sg[ab], tt));
      TTf = c2add(TTf, c2scale(Cf[ab], tt));
      TTn = c2add(TTn, c2scale(Nf[ab], tt));
      DD = cadd_rc(DD, (*Dp[a]) * (*Dq[b]), Ct[ab]);
      DT = cadd_rc(DT, Tq[b] * (*Dp[a]), Mf[ab]);
      TD = cadd_rc(TD, Tp[a] * (*Dq[b]), Mg[ab]);
    }
  const D4 A1 = cx * (P.tpr * P.tqr) + P.tpz * P.tqz;
  // rr part: A1 TTc + cx TTf - j sx (tpr TTs - tqr TTg)
  C4 u = rc(A1, TTc);
  u.re = u.re + dconst(cx * TTf.x);
  u.im = u.im + dconst(cx * TTf.y);
  const C4 w = rc(P.tpr, TTs) - rc(P.tqr, TTg);   // -j sx w
  u.re = u.re + sx * w.im;
  u.im = u.im - sx * w.re;
  // ww part: DD + TTn + j (DT - TD)
  C4 v = DD;
  v.re = v.re + TTn.x;
  v.im = v.im + TTn.y;
  const C4 x = DT - TD;
  v.re = v.re - x.im;
  v.im = v.im + x.re;
  return P.rr * u - (ik2 * P.ww) * v;
}

// e^{-jkR} / (4 pi R) as a complex dual of R (s, c = sin, cos of k R)
__device__ __forceinline__ C4 green(const D4& R, double k, double s, double c) {
  const double i4 = 1.0 / (4.0 * GH_PI * R.v);
  const double gr = c * i4, gi = -s * i4;
  // dg/dR = g (-jk - 1/R)
  const double ir = 1.0 / R.v;
  const double pr = -gr * ir + gi * k, pi = -gi * ir - gr * k;
  return C4{dchain(R, gr, pr), dchain(R, gi, pi)};
}
__device__ __forceinline__ C4 green(const D4& R, double k) {
  double s, c;
  sincos(k * R.v, &s, &c);
  return green(R, k, s, c);
}

// p(R) = (1 + jkR) e^{-jkR} / (4 pi R^3) as a complex dual of R;
// p'(R) = [k^2 R e^{-jkR} - 3 (1 + jkR) e^{-jkR} / R] / (4 pi R^3)
__device__ __forceinline__ C4 pkern(const D4& R, double k, double s, double c) {
  const double r = R.v;
  const double i4 = 1.0 / (4.0 * GH_PI * r * r * r);
  // h = (1 + jkr)(c - j s) = (c + k r s) + j (k r c - s)
  const double hr = c + k * r * s, hi = k * r * c - s;
  const double pr = hr * i4, pi = hi * i4;
  const double dr = (k * k * r * c - 3.0 * hr / r) * i4, di = (-k * k * r * s - 3.0 * hi / r) * i4;
  return C4{dchain(R, pr, dr), dchain(R, pi, di)};
}
__device__ __forceinline__ C4 pkern(const D4& R, double k) {
  double s, c;
  sincos(k * R.v, &s, &c);
  return pkern(R, k, s, c);
}

// The four MFIE brackets (without p(R)) at one xi (bor_kernels.py:273-306):
// b_tt = -(WtR n_t - Wt_t nR), b_tf = -(WtR n_f - Wt_f nR),
// b_ft = -(Ry n_t - Wf_t nR), b_ff = -(Ry n_f - cos xi nR); also R.
__device__ __forceinline__ void brackets(const DPair& P, const D4& cx, const D4& sx, D4* b, D4* R) {
  const D4 Rx = P.rp - P.rq * cx, Ry = P.rq * sx;
  const double Rz = P.dz;
  *R = dsqrt(Rx * Rx + Ry * Ry + Rz * Rz);
  const D4 WtR = P.tpr * Rx + Rz * P.tpz, nR = Rz * P.tpr - P.tpz * Rx;
  const D4 n_t = P.tpr * P.tqz - P.tpz * P.tqr * cx, n_f = -(P.tpz * sx);
  const D4 Wt_t = P.tpr * P.tqr * cx + P.tpz * P.tqz, Wt_f = P.tpr * sx, Wf_t = -(P.tqr * sx);
  b[0] = -(WtR * n_t - Wt_t * nR);
  b[1] = -(WtR * n_f - Wt_f * nR);
  b[2] = -(Ry * n_t - Wf_t * nR);
  b[3] = -(Ry * n_f - cx * nR);
}

// ─────────────────────────────────────────────────────────────────────────────
// far element pairs
// ─────────────────────────────────────────────────────────────────────────────

struct AdjFarArgs {
  const double* nodes;   // [Nn][2]
  const double* gs;      // Gauss points s on [0, 1] [4]
  const double* gwq;     // their weights [4]
  const SysInfo* sys;    // the problem's 2M+1 systems, mode m at index M + m
  QView Q;
  int Ne, M, nE, nB;
  double k;
  double ce_re, ce_im;   // EFIE coefficient (cE j k eta0 2 pi)
  double cm;             // far bracket coefficient (-2 pi cM)
  double* out;           // [Ne * Ne][4]
};

constexpr int ADJ_CHUNK = 64;   // azimuth samples per staging round

__device__ DPair make_pair(const DElem& Ee, const DElem& Ef, const DPoint& p, const DPoint& q, double W) {
  DPair P;
  P.rp = p.rho; P.rq = q.rho;
  P.tpr = Ee.tr; P.tpz = Ee.tz; P.tqr = Ef.tr; P.tqz = Ef.tz;
  P.Dp0 = p.D0; P.Dp1 = p.D1; P.Dq0 = q.D0; P.Dq1 = q.D1;
  P.ww = W * (Ee.L * Ef.L);
  P.rr = P.ww * (p.rho * q.rho);
  P.dz = p.z - q.z;
  P.Tp0 = p.T0; P.Tp1 = p.T1; P.Tq0 = q.T0; P.Tq1 = q.T1;
  return P;
}

// the work-group's four partials to out[0..3] (deterministic order)
__device__ void block_sum4(double (&acc)[4], double* out) {
  __shared__ double red[4][256];
#pragma unroll
  for (int i = 0; i < 4; ++i) red[i][threadIdx.x] = acc[i];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int i = 0; i < 4; ++i) red[i][threadIdx.x] += red[i][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) out[threadIdx.x] = red[threadIdx.x][0];
}

// One 256-thread work-group per ordered element pair (e, f), blockIdx.x =
// e * Ne + f; near pairs return at once.
__global__ __launch_bounds__(256) void k_adj_far(AdjFarArgs a) {
  const int e = blockIdx.x / a.Ne, f = blockIdx.x % a.Ne;
  if (abs(e - f) <= 1) {   // near pairs: k_adj_near
    if (threadIdx.x < 4) a.out[(size_t)blockIdx.x * 4 + threadIdx.x] = 0.0;
    return;
  }
  extern __shared__ double2 lds[];
  double2* Pq = lds;                                   // [16][M+1]
  double2* Fx = lds + 16 * (a.M + 1);                  // [ADJ_CHUNK][7][4]
  stage_q(a.sys, a.Q, a.M, e, f, Pq);
  const int tid = threadIdx.x;
  const int pp = tid & 15;
  // the 16 point pairs' duals live in LDS (read per sample), which keeps
  // them out of the registers the sample loops need
  __shared__ DPair sP[16];
  if (tid < 16) {
    const DElem Ee = delem(a.nodes, e, 0), Ef = delem(a.nodes, f, 2);
    const int pi = tid >> 2, qi = tid & 3;
    const DPoint p = dpoint(Ee, a.gs[pi]), q = dpoint(Ef, a.gs[qi]);
    sP[tid] = make_pair(Ee, Ef, p, q, a.gwq[pi] * a.gwq[qi]);
  }
  const DPair& P = sP[pp];
  const double k = a.k, ik2 = 1.0 / (k * k);
  C4 accE = cdz();
  D4 accK = dconst(0.0);
  __syncthreads();
  for (int grid = 0; grid < 2; ++grid) {
    const bool efie = grid == 0;
    const int N = efie ? a.nE : a.nB;
    if (efie && a.ce_re == 0.0 && a.ce_im == 0.0) continue;
    if (!efie && a.cm == 0.0) continue;
    const int J = N / 2 + 1;
    for (int j0 = 0; j0 < J; j0 += ADJ_CHUNK) {
      // stage the series of ADJ_CHUNK samples: thread = (sample, entry)
      {
        const int js = j0 + (tid >> 2), ab = tid & 3;
        if (js < J) {
          double sx, cx;
          sincos(2.0 * GH_PI * js / N, &sx, &cx);
          double2 F[7];
          if (efie) qseries<true>(Pq, a.M, ab, cx, sx, F);
          else qseries<false>(Pq, a.M, ab, cx, sx, F);
          double2* o = Fx + (size_t)(tid >> 2) * 28;
#pragma unroll
          for (int i = 0; i < 7; ++i) o[i * 4 + ab] = F[i];
        }
      }
      __syncthreads();
#pragma unroll 2
      for (int sl = tid >> 4; sl < ADJ_CHUNK; sl += 16) {
        const int js = j0 + sl;
        if (js >= J) break;
        double sx, cx;
        sincos(2.0 * GH_PI * js / N, &sx, &cx);
        const double wj = (2.0 * GH_PI / N) * ((js == 0 || 2 * js == N) ? 1.0 : 2.0);
        const double2* F = Fx + (size_t)sl * 28;
        if (efie) {
          // g(R), R^2 = d^2 + 4 rho_p rho_q sin^2(xi / 2)
          const D4 dr = P.rp - P.rq;
          const D4 R = dsqrt(dr * dr + P.dz * P.dz + (2.0 * (1.0 - cx)) * (P.rp * P.rq));
          const C4 g = green(R, k);
          const C4 phi = efie_phi(P, F + 0, F + 4, F + 8, F + 12, F + 16, F + 20, F + 24, cx, sx, ik2);
          accE = accE + wj * (g * phi);
        } else {
          D4 b[4], R;
          brackets(P, dconst(cx), dconst(sx), b, &R);
          const C4 pk = pkern(R, k);
          // U = b_tt TT(Ct) - j b_tf TT(Sf) - j b_ft TT(Sg) + b_ff TT(Cf)
          double2 tc = make_double2(0, 0), ts = tc, tg = tc, tf = tc;
          const double Tp[2] = {P.Tp0, P.Tp1}, Tq[2] = {P.Tq0, P.Tq1};
#pragma unroll
          for (int ab = 0; ab < 4; ++ab) {
            const double tt = Tp[ab >> 1] * Tq[ab & 1];
            tc = c2add(tc, c2scale(F[0 + ab], tt));
            ts = c2add(ts, c2scale(F[4 + ab], tt));
            tg = c2add(tg, c2scale(F[12 + ab], tt));
            tf = c2add(tf, c2scale(F[20 + ab], tt));
          }
          const C4 U = rc(b[0], tc) + rc(b[1], c2mj(ts)) + rc(b[2], c2mj(tg)) + rc(b[3], tf);
          const C4 pu = pk * U;
          accK = accK + wj * (P.rr * pu.re);
        }
      }
      __syncthreads();
    }
  }
  double acc[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) acc[i] = a.ce_re * accE.re.d[i] - a.ce_im * accE.im.d[i] + a.cm * accK.d[i];
  block_sum4(acc, a.out + (size_t)blockIdx.x * 4);
}

// ─────────────────────────────────────────────────────────────────────────────
// near element pairs
// ─────────────────────────────────────────────────────────────────────────────

struct AdjNearPair {
  int e, f, kind, pp_off, n_tail, tail_off;
};

struct AdjNearArgs {
  const double* nodes;
  const double* cell_s;
  const double* cell_sp;
  const double* cell_w;
  int cell_off[8], cell_cnt[8];
  const double* core_u;   // [48]
  const double* core_w;
  const double* tail_u;
  const double* tail_w;
  const AdjNearPair* pairs;
  const int2* batches;    // (pair slot, first cell point) per work-group of 32 point pairs
  const SysInfo* sys;
  QView Q;
  const double* gs;       // element Gauss rule (mass term)
  const double* gwq;
  int Ne, M;
  double k;
  double ce_re, ce_im;    // EFIE coefficient
  double cm;              // near bracket coefficient (-2 pi cM)
  double cmass;           // mass coefficient (pi cM: (1/2) cM x 2 pi)
  double* out;            // [n batches][4]
};

// One 256-thread work-group per (near pair, up to 32 cell point pairs).  The
// point pairs are taken one at a time.  For each, the series coefficients are
// first contracted over the four basis entries ab with the pair's weights
// (T_a T_b for the TT sums; value and four partials of D_a D_b, D_a T_b and
// T_a D_b for the DD, DT and TD sums) into 19 arrays over m in LDS; then one
// thread per azimuth node (the pair's 48 core and n_tail tail nodes) evaluates
// the 28 series it needs from those arrays (values and xi slopes), instead of
// four series evaluations per node, one per entry.
__global__ __launch_bounds__(256) void k_adj_near(AdjNearArgs a) {
  const int2 bt = a.batches[blockIdx.x];
  const AdjNearPair pr = a.pairs[bt.x];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  extern __shared__ double2 lds[];
  const int M1 = a.M + 1;
  double2* Pq = lds;             // [16][M+1]
  double2* Cq = lds + 16 * M1;   // [19][M+1], the current point pair's contracted coefficients
  if (pr.e >= 0) stage_q(a.sys, a.Q, a.M, pr.e, pr.f, Pq);
  __syncthreads();
  const int tid = threadIdx.x;
  if (pr.e >= 0) {
    const int cnt = a.cell_cnt[pr.kind];
    const int npp = min(32, cnt - bt.y);
    const double k = a.k, ik2 = 1.0 / (k * k);
    const DElem Ee = delem(a.nodes, pr.e, 0), Ef = delem(a.nodes, pr.f, 2);
    // the point pairs' duals and their azimuth rules, one thread each
    struct DRule {
      D4 d, aa, doa, vmax, xi0, span;
      int ok;
    };
    __shared__ DPair sP[32];
    __shared__ DRule sR[32];
    if (tid < npp) {
      const int qi = a.cell_off[pr.kind] + bt.y + tid;
      const DPoint p = dpoint(Ee, a.cell_s[qi]), qq = dpoint(Ef, a.cell_sp[qi]);
      const DPair P = make_pair(Ee, Ef, p, qq, a.cell_w[qi]);
      sP[tid] = P;
      const D4 dr = P.rp - P.rq;
      const D4 d2 = dr * dr + P.dz * P.dz;
      const D4 rr4 = 4.0 * (P.rp * P.rq);
      DRule r;
      r.ok = rr4.v > 1e-30;   // axis pairs (closed form, order 0 only) do not occur off the axis
      r.d = dsqrt(dmax(d2, 1e-300));
      r.aa = dsqrt(dmax(rr4, 1e-300));
      r.doa = r.d / r.aa;
      const D4 s0 = dmin(20.0 * r.doa, 0.25);
      r.vmax = dasinh(s0 / r.doa);
      r.xi0 = 2.0 * dasin(dmin(s0, 1.0));
      r.span = GH_PI - r.xi0;
      sR[tid] = r;
    }
    __syncthreads();
    C4 accE = cdz();
    D4 accK = dconst(0.0);
    const int nnodes = 48 + pr.n_tail;
    for (int sl = 0; sl < npp; ++sl) {
      if (!sR[sl].ok) continue;   // the same for every thread: the barriers below stay matched
      const DPair& P = sP[sl];
      // contracted coefficients: arrays 0-3 the tt, tf, ft, ff blocks with T_a T_b;
      // 4-8 tt with D_a D_b, 9-13 tf with D_a T_b, 14-18 ft with T_a D_b (value, partials)
      for (int i = tid; i < 19 * M1; i += 256) {
        const int arr = i / M1, m = i - arr * M1;
        const double Tp[2] = {P.Tp0, P.Tp1}, Tq[2] = {P.Tq0, P.Tq1};
        double2 z = make_double2(0.0, 0.0);
        if (arr < 4) {
#pragma unroll
          for (int ab = 0; ab < 4; ++ab) z = c2add(z, c2scale(Pq[(arr * 4 + ab) * M1 + m], Tp[ab >> 1] * Tq[ab & 1]));
        } else {
          const int grp = (arr - 4) / 5, comp = (arr - 4) - 5 * grp;
#pragma unroll
          for (int ab = 0; ab < 4; ++ab) {
            const int x = ab >> 1, y = ab & 1;
            const D4& Dpx = x == 0 ? P.Dp0 : P.Dp1;
            const D4& Dqy = y == 0 ? P.Dq0 : P.Dq1;
            const D4 w = grp == 0 ? Dpx * Dqy : (grp == 1 ? Tq[y] * Dpx : Tp[x] * Dqy);
            const double wc = comp == 0 ? w.v : w.d[comp - 1];
            z = c2add(z, c2scale(Pq[(grp * 4 + ab) * M1 + m], wc));
          }
        }
        Cq[i] = z;
      }
      __syncthreads();
      const D4& d = sR[sl].d;
      const D4& aa = sR[sl].aa;
      const D4& doa = sR[sl].doa;
      const D4& vmax = sR[sl].vmax;
      const D4& xi0 = sR[sl].xi0;
      const D4& span = sR[sl].span;
      for (int vn = tid; vn < nnodes; vn += 256) {
        D4 wq, R, cx, sx;
        if (vn < 48) {
          const D4 vv = a.core_u[vn] * vmax, wv = a.core_w[vn] * vmax;
          D4 sv = doa * dsinh(vv);
          sv = dmin(sv, 1.0);
          const D4 cv = dsqrt(dmax(1.0 - sv * sv, 0.0));
          const D4 den = cv.v >= sqrt(1e-15) ? cv : dconst(sqrt(1e-15));
          wq = (2.0 * wv) * doa * dcosh(vv) / den;
          R = dsqrt(d * d + (aa * sv) * (aa * sv));
          sx = 2.0 * (sv * cv);
          cx = 1.0 - 2.0 * (sv * sv);
        } else {
          const int ti = pr.tail_off + (vn - 48);
          const D4 xi = xi0 + a.tail_u[ti] * span;
          wq = a.tail_w[ti] * span;
          D4 sh, ch;
          dsincos(0.5 * xi, &sh, &ch);
          R = dsqrt(d * d + (aa * sh) * (aa * sh));
          sx = 2.0 * (sh * ch);
          cx = 1.0 - 2.0 * (sh * sh);
        }
        R = dmax(R, 1e-300);
        // xi of the node and its partials: dxi = cos xi dsin xi - sin xi dcos xi
        D4 dxi;
        dxi.v = 0.0;
#pragma unroll
        for (int i = 0; i < 4; ++i) dxi.d[i] = cx.v * sx.d[i] - sx.v * cx.d[i];
        // the series at xi: TT sums of Ct, Sf, Sg, Cf, Nf (values and slopes),
        // DD(Ct), DT(Mf), TD(Mg) (value and four weight partials, and the
        // slope of the value)
        const double2 z0 = make_double2(0.0, 0.0);
        double2 tc = z0, tcs = z0, ts = z0, tss = z0, tg = z0, tgs = z0, tf = z0, tfs = z0, tn = z0, tns = z0;
        double2 dd[5], dt[5], td[5], dds = z0, dts = z0, tds = z0;
#pragma unroll
        for (int j = 0; j < 5; ++j) { dd[j] = z0; dt[j] = z0; td[j] = z0; }
        double c = 1.0, s = 0.0, cp = cx.v, sp = -sx.v;
        const double tcx = 2.0 * cx.v;
        // two orders per step: the next order's 19 LDS loads are in flight
        // while this order's sums are formed (at one wavefront per SIMD nothing
        // else hides their latency)
#pragma unroll 2
        for (int m = 0; m <= a.M; ++m) {
          const double fm = (double)m;
          const double mc = fm * c, ms = fm * s, m2c = fm * mc, m2s = fm * ms, m3s = fm * m2s;
          const double2* cm = Cq + m;
          const double2 Att = cm[0], Atf = cm[M1], Aft = cm[2 * M1], Aff = cm[3 * M1];
          tc = c2add(tc, c2scale(Att, c));    tcs = c2add(tcs, c2scale(Att, -ms));
          ts = c2add(ts, c2scale(Atf, s));    tss = c2add(tss, c2scale(Atf, mc));
          tg = c2add(tg, c2scale(Aft, s));    tgs = c2add(tgs, c2scale(Aft, mc));
          tf = c2add(tf, c2scale(Aff, c));    tfs = c2add(tfs, c2scale(Aff, -ms));
          tn = c2add(tn, c2scale(Aff, m2c));  tns = c2add(tns, c2scale(Aff, -m3s));
#pragma unroll
          for (int j = 0; j < 5; ++j) {
            const double2 B = cm[(4 + j) * M1], Cc = cm[(9 + j) * M1], D = cm[(14 + j) * M1];
            dd[j] = c2add(dd[j], c2scale(B, c));
            dt[j] = c2add(dt[j], c2scale(Cc, mc));
            td[j] = c2add(td[j], c2scale(D, mc));
            if (j == 0) {
              dds = c2add(dds, c2scale(B, -ms));
              dts = c2add(dts, c2scale(Cc, -m2s));
              tds = c2add(tds, c2scale(D, -m2s));
            }
          }
          const double cn = tcx * c - cp, sn = tcx * s - sp;
          cp = c; sp = s;
          c = cn; s = sn;
        }
        auto tdual = [&](double2 v, double2 sv) -> C4 {
          return C4{dchain(dxi, v.x, sv.x), dchain(dxi, v.y, sv.y)};
        };
        auto xdual = [&](const double2 (&x)[5], double2 xs) -> C4 {
          C4 r;
          r.re.v = x[0].x;
          r.im.v = x[0].y;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            r.re.d[j] = x[j + 1].x + xs.x * dxi.d[j];
            r.im.d[j] = x[j + 1].y + xs.y * dxi.d[j];
          }
          return r;
        };
        const C4 TTc = tdual(tc, tcs), TTs = tdual(ts, tss), TTg = tdual(tg, tgs), TTf = tdual(tf, tfs);
        double skr, ckr;   // e^{-jkR} of the node, shared by g(R) and p(R)
        sincos(k * R.v, &skr, &ckr);
        // EFIE: 2 wq g(R) Phi (Phi as in efie_phi, with dual series)
        {
          const C4 g = green(R, k, skr, ckr);
          const C4 TTn = tdual(tn, tns), DD = xdual(dd, dds), DT = xdual(dt, dts), TD = xdual(td, tds);
          const D4 A1 = cx * (P.tpr * P.tqr) + P.tpz * P.tqz;
          C4 u = A1 * TTc + cx * TTf;
          const C4 w = P.tpr * TTs - P.tqr * TTg;   // - j sx w
          u.re = u.re + sx * w.im;
          u.im = u.im - sx * w.re;
          const C4 xx = DT - TD;
          C4 v = DD + TTn;
          v.re = v.re - xx.im;
          v.im = v.im + xx.re;
          const C4 phi = P.rr * u - (ik2 * P.ww) * v;
          accE = accE + (2.0 * wq) * (g * phi);
        }
        // brackets: 2 wq rr p(R) [b_tt TT(Ct) - j b_tf TT(Sf) - j b_ft TT(Sg) + b_ff TT(Cf)]
        if (a.cm != 0.0) {
          D4 b[4], Rb;
          brackets(P, cx, sx, b, &Rb);
          const C4 pk = pkern(R, k, skr, ckr);
          // -j ts = (ts.im, -ts.re)
          const C4 U = b[0] * TTc + b[1] * C4{TTs.im, -TTs.re} + b[2] * C4{TTg.im, -TTg.re} + b[3] * TTf;
          const C4 pu = pk * U;
          accK = accK + ((2.0 * wq) * P.rr) * pu.re;
        }
      }
      __syncthreads();   // every thread is done with Cq before the next pair's contraction
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i] = a.ce_re * accE.re.d[i] - a.ce_im * accE.im.d[i] + a.cm * accK.d[i];
    // mass term (1/2) cM M on the tt and ff blocks of a diagonal pair, once per pair
    // (bor_solver.py:389-397): pi cM sum_p w_p rho_p T_a T_b (Q_tt + Q_ff summed over m)
    if (pr.e == pr.f && bt.y == 0 && tid < 4 && a.cmass != 0.0) {
      const DPoint p = dpoint(Ee, a.gs[tid]);
      const D4 wr = (a.gwq[tid] * Ee.L) * p.rho;
      const double T[2] = {p.T0, p.T1};
      double2 qs = make_double2(0.0, 0.0);
      for (int ab = 0; ab < 4; ++ab) {
        double2 st = make_double2(0.0, 0.0);
        for (int m = 0; m <= a.M; ++m) {
          st = c2add(st, Pq[(0 * 4 + ab) * M1 + m]);
          st = c2add(st, Pq[(3 * 4 + ab) * M1 + m]);
        }
        qs = c2add(qs, c2scale(st, T[ab >> 1] * T[ab & 1]));
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[i] += a.cmass * qs.x * wr.d[i];
    }
  }
  block_sum4(acc, a.out + (size_t)blockIdx.x * 4);
}

// ─────────────────────────────────────────────────────────────────────────────
// right-hand side and far-field projection
// ─────────────────────────────────────────────────────────────────────────────

struct AdjExciteArgs {
  const double* nodes;
  const double* gs;
  const double* gwq;
  const SysInfo* sys;     // the problem's systems (mode m at M + m)
  QView Q;                // B part at columns [n, n + nrhs)
  const double2* X;       // solutions: s.x_off + row * s.ncols + col
  const double2* acoef;   // a' per column (far-field prefactor included)
  const double* J;        // Bessel table [M+2][P][ntheta]
  const double* theta;
  int Ne, P, M, ntheta;
  double k, cE, cM;
  double* out;            // [Ne][ntheta][2]
};

// j^|n| J'_|n|(x) d x-direction parts: value and derivative of j_n = i^|n| J_|n|
__device__ __forceinline__ double bessel_at(const double* Jp, size_t stride, int n, int M, double x) {
  if (n <= M + 1) return Jp[(size_t)n * stride];
  // J_{M+2} by one upward step (only its slope term uses it)
  if (fabs(x) < 1e-300) return 0.0;
  return (2.0 * (M + 1) / x) * Jp[(size_t)(M + 1) * stride] - Jp[(size_t)M * stride];
}
__device__ __forceinline__ double bessel_slope(const double* Jp, size_t stride, int n, int M, double x) {
  if (n == 0) return -Jp[stride];
  return 0.5 * (bessel_at(Jp, stride, n - 1, M, x) - bessel_at(Jp, stride, n + 1, M, x));
}
// i^n v for n >= 0 on a dual
__device__ __forceinline__ C4 iphase(int n, const D4& v) {
  switch (n & 3) {
    case 0: return C4{v, dconst(0.0)};
    case 1: return C4{dconst(0.0), v};
    case 2: return C4{-v, dconst(0.0)};
    default: return C4{dconst(0.0), -v};
  }
}

// One thread per (element e, aspect): the derivative along the element's two
// node radii of Re sum Lambda (.) B (its share of the right-hand sides) and of
// Re sum a' S (its share of the far-field projection at fixed currents).
__global__ __launch_bounds__(128) void k_adj_excite(AdjExciteArgs a) {
  const int e = blockIdx.y;
  const int it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= a.ntheta || e >= a.Ne) return;
  double st, ct;
  sincos(a.theta[it], &st, &ct);
  const DElem E = delem(a.nodes, e, 0);
  const size_t stride = (size_t)a.P * a.ntheta;
  const double ie = 1.0 / 376.730313668;
  double acc[2] = {0.0, 0.0};
  for (int qp = 0; qp < 4; ++qp) {
    const int pidx = 4 * e + qp;
    const DPoint p = dpoint(E, a.gs[qp]);
    const D4 w = a.gwq[qp] * E.L;
    double sp, cp;
    sincos(a.k * ct * p.z, &sp, &cp);
    const C4 Eph = rc(w * p.rho, make_double2(cp, sp));   // w rho e^{j k cos(theta) z}
    const D4 u = (a.k * st) * p.rho;                       // Bessel argument
    const double* Jp = a.J + (size_t)pidx * a.ntheta + it;
    const double T[2] = {p.T0, p.T1};
    for (int mi = 0; mi < 2 * a.M + 1; ++mi) {
      const SysInfo s = a.sys[mi];
      const int m = s.m;
      const int n0 = abs(m), nm = abs(m - 1), np = abs(m + 1);
      const D4 b0 = dchain(u, Jp[(size_t)n0 * stride], bessel_slope(Jp, stride, n0, a.M, u.v));
      const D4 bm = dchain(u, Jp[(size_t)nm * stride], bessel_slope(Jp, stride, nm, a.M, u.v));
      const D4 bp = dchain(u, Jp[(size_t)np * stride], bessel_slope(Jp, stride, np, a.M, u.v));
      const C4 jm = iphase(n0, b0), jm1 = iphase(nm, bm), jp1 = iphase(np, bp);
      // Ic = pi (j_{m-1} + j_{m+1}), Is = -j pi (j_{m-1} - j_{m+1}), I1 = 2 pi j_m
      const C4 Ic = GH_PI * (jm1 + jp1);
      const C4 dm = jm1 - jp1;
      const C4 Is = C4{GH_PI * dm.im, -GH_PI * dm.re};
      const C4 I1 = (2.0 * GH_PI) * jm;
      const C4 uE = (ct * E.tr) * Ic - (st * E.tz) * I1;
      // ── right-hand sides: rows of nodes e (T0) and e + 1 (T1) ──
      {
        C4 tE[2], tH[2], fE[2], fH[2];
        tE[0] = uE;                  tH[0] = ie * Ic;
        tE[1] = E.tr * Is;           tH[1] = (ct * ie) * Is;
        fE[0] = (-ct) * Is;          fH[0] = (-ie) * (E.tr * Is);
        fE[1] = Ic;                  fH[1] = ie * uE;
        for (int pol = 0; pol < 2; ++pol) {
          const int col = 2 * it + pol;
          if (col >= s.nrhs) continue;
          const C4 vt = (a.cE * tE[pol] + a.cM * tH[pol]) * Eph;
          const C4 vf = (a.cE * fE[pol] + a.cM * fH[pol]) * Eph;
          for (int ai = 0; ai < 2; ++ai) {
            const int rt = sys_row(s, 0, e + ai), rf = sys_row(s, 1, e + ai);
            if (rt >= 0) {
              const double2 qv = a.Q[s.a_off + (size_t)rt * s.ld + s.n + col];
#pragma unroll
              for (int i = 0; i < 2; ++i) acc[i] += T[ai] * (qv.x * vt.re.d[i] - qv.y * vt.im.d[i]);
            }
            if (rf >= 0) {
              const double2 qv = a.Q[s.a_off + (size_t)rf * s.ld + s.n + col];
#pragma unroll
              for (int i = 0; i < 2; ++i) acc[i] += T[ai] * (qv.x * vf.re.d[i] - qv.y * vf.im.d[i]);
            }
          }
        }
      }
      // ── far-field projection at fixed currents (bor_solver.py:659-696) ──
      {
        // Icos = pi (j_{m+1} + j_{m-1}) = Ic, Isin = -j pi (j_{m+1} - j_{m-1}) = -Is
        for (int pol = 0; pol < 2; ++pol) {
          const int col = 2 * it + pol;
          if (col >= s.nrhs) continue;
          double2 xt = make_double2(0, 0), xf = xt;
          for (int ai = 0; ai < 2; ++ai) {
            const int rt = sys_row(s, 0, e + ai), rf = sys_row(s, 1, e + ai);
            if (rt >= 0) xt = c2add(xt, c2scale(a.X[s.x_off + (size_t)rt * s.ncols + s.col + col], T[ai]));
            if (rf >= 0) xf = c2add(xf, c2scale(a.X[s.x_off + (size_t)rf * s.ncols + s.col + col], T[ai]));
          }
          C4 term;
          if (pol == 0) {
            // Jt (ct tr Icos - st tz I1) - Jf ct Isin = Jt uE + ct Jf Is
            term = C4{uE.re * xt.x - uE.im * xt.y, uE.re * xt.y + uE.im * xt.x};
            term = term + ct * C4{Is.re * xf.x - Is.im * xf.y, Is.re * xf.y + Is.im * xf.x};
          } else {
            // Jt tr Isin + Jf Icos = -tr Jt Is + Jf Ic
            const C4 trIs = E.tr * Is;
            term = C4{Ic.re * xf.x - Ic.im * xf.y, Ic.re * xf.y + Ic.im * xf.x} -
                   C4{trIs.re * xt.x - trIs.im * xt.y, trIs.re * xt.y + trIs.im * xt.x};
          }
          const C4 tv = term * Eph;
          const double2 ac = a.acoef[col];
#pragma unroll
          for (int i = 0; i < 2; ++i) acc[i] += ac.x * tv.re.d[i] - ac.y * tv.im.d[i];
        }
      }
    }
  }
  double* o = a.out + ((size_t)e * a.ntheta + it) * 2;
  o[0] = acc[0];
  o[1] = acc[1];
}

// ─────────────────────────────────────────────────────────────────────────────
// gather to nodes
// ─────────────────────────────────────────────────────────────────────────────

// grad[q] = the far pairs' partials of node q (test slot 0 of element q, slot
// 1 of element q - 1; source slots 2 and 3 likewise), the near batches' and
// the excitation's; zero for nodes on the axis.  One 256-thread work-group
// per node: strided partial sums, then a fixed-order tree (deterministic).
__global__ __launch_bounds__(256) void k_adj_gather(const double* far, const double* nearb, const int2* batches,
                                                    const AdjNearPair* pairs, int nbatch, const double* ex, int Ne,
                                                    int ntheta, const double* nodes, double axis_tol, double* grad) {
  const int qn = blockIdx.x;
  const int t = threadIdx.x;
  double g = 0.0;
  for (int o = t; o < Ne; o += 256) {
    if (qn < Ne) g += far[((size_t)qn * Ne + o) * 4 + 0] + far[((size_t)o * Ne + qn) * 4 + 2];
    if (qn >= 1) g += far[((size_t)(qn - 1) * Ne + o) * 4 + 1] + far[((size_t)o * Ne + qn - 1) * 4 + 3];
  }
  for (int b = t; b < nbatch; b += 256) {
    const AdjNearPair pr = pairs[batches[b].x];
    if (pr.e < 0) continue;
    const double* v = nearb + (size_t)b * 4;
    if (pr.e == qn) g += v[0];
    if (pr.e + 1 == qn) g += v[1];
    if (pr.f == qn) g += v[2];
    if (pr.f + 1 == qn) g += v[3];
  }
  for (int it = t; it < ntheta; it += 256) {
    if (qn < Ne) g += ex[((size_t)qn * ntheta + it) * 2 + 0];
    if (qn >= 1) g += ex[((size_t)(qn - 1) * ntheta + it) * 2 + 1];
  }
  __shared__ double red[256];
  red[t] = g;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) grad[qn] = nodes[2 * qn] > axis_tol ? red[0] : 0.0;
}

}  // namespace ghost
