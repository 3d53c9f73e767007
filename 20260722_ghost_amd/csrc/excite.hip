// Plane-wave excitation and monostatic far field for every mode at once.
//
// Reference algorithm: bor_solver.py:625-645 (EFIE rhs), 437-461 (MFIE rhs),
// 659-708 (far field; note the (j_{m+1}, j_{m-1}) order there), 940-950
// (CFIE mix).  The Bessel functions J_n(k rho sin th), n = 0..M+1, that the
// reference obtains per call from scipy.special.jv are produced here for all
// orders of one (point, angle) by Miller's backward recurrence normalised with
// J_0 + 2 sum J_2k = 1 (power series below x = 1).
#include "bor_common.hpp"

namespace ghost {

struct ExciteArgs {
  PointSet pts;
  int P, Nn, M, ntheta;
  double k;
  double cE, cM;          // V = cE V_E + cM V_H
  const double* theta;    // [ntheta] radians
  double* J;              // [M+2][P][ntheta] Bessel table
  double2* E;             // [P][ntheta] w_p rho_p e^{j k cos(theta) z_p}: the plane-wave phase of a
                          // point times its quadrature weight, shared by the right-hand sides and the
                          // far-field projection (both k_bessel outputs follow the same test geometry)
  const double2* zs;      // per-element surface impedance, or null: the magnetic current
                          // M = -Z_s n x J radiates too (bor_solver.py:697-707)
};

// J_n(x), n = 0..nmax written to out[n * stride].
__device__ void bessel_orders(double x, int nmax, double* out, size_t stride) {
  if (x < 1e-300) {
    out[0] = 1.0;
    for (int n = 1; n <= nmax; ++n) out[n * stride] = 0.0;
    return;
  }
  if (x < 1.0) {
    // J_n(x) = sum_k (-1)^k (x/2)^(2k+n) / (k! (n+k)!)
    const double h = 0.5 * x, h2 = h * h;
    double lead = 1.0;  // (x/2)^n / n!
    for (int n = 0; n <= nmax; ++n) {
      if (n > 0) lead *= h / n;
      double term = lead, sum = lead;
      for (int kk = 1; kk < 40; ++kk) {
        term *= -h2 / ((double)kk * (double)(n + kk));
        sum += term;
        if (fabs(term) < 1e-18 * fabs(sum)) break;
      }
      out[n * stride] = sum;
    }
    return;
  }
  // Miller: start well above max(nmax, x); rescale by 1e-250 when values grow.
  // Pass 1 finds the normalisation and the number of rescales; pass 2 repeats
  // the identical recurrence and stores each order in the final scale.
  const double top = fmax((double)nmax, x);
  int N = (int)(top + 20.0 + sqrt(40.0 * top));
  N += N & 1;
  const double BIG = 1e250, SMALL = 1e-250;
  // 2n/x as n (2/x): one division per (point, angle) instead of one per step
  // (within 1.5 ulp of the quotient, the size of each step's own rounding)
  const double tx = 2.0 / x;
  int nres = 0;
  double norm = 0.0;
  {
    double fp = 0.0, f = 1e-280;
    for (int n = N; n >= 1; --n) {
      const double fm = ((double)n * tx) * f - fp;
      fp = f;
      f = fm;
      if ((n - 1) > 0 && ((n - 1) & 1) == 0) norm += 2.0 * f;
      if (fabs(f) > BIG) { f *= SMALL; fp *= SMALL; norm *= SMALL; ++nres; }
    }
    norm += f;  // f is now f_0
  }
  {
    double fp = 0.0, f = 1e-280;
    int r = 0;
    for (int n = N; n >= 1; --n) {
      const double fm = ((double)n * tx) * f - fp;
      fp = f;
      f = fm;
      if (fabs(f) > BIG) { f *= SMALL; fp *= SMALL; ++r; }
      const int o = n - 1;
      if (o <= nmax) {
        double v = f / norm;
        for (int q = r; q < nres; ++q) v *= SMALL;
        out[o * stride] = v;
      }
    }
  }
}

__global__ void k_bessel(ExciteArgs a) {
  const size_t tot = (size_t)a.P * a.ntheta;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < tot;
       t += (size_t)gridDim.x * blockDim.x) {
    const int it = (int)(t % a.ntheta);
    const int p = (int)(t / a.ntheta);
    double st, ct;
    sincos(a.theta[it], &st, &ct);
    const double x = a.k * a.pts.rho[p] * st;
    double* Jt = a.J + (size_t)p * a.ntheta + it;
    const size_t js = (size_t)a.P * a.ntheta;
    bessel_orders(fabs(x), a.M + 1, Jt, js);
    // sin(theta) < 0 (aspects beyond 180 degrees or below 0): J_n(-x) = (-1)^n J_n(x)
    if (x < 0.0)
      for (int n = 1; n <= a.M + 1; n += 2) Jt[(size_t)n * js] = -Jt[(size_t)n * js];
    double sp, cp;
    sincos(a.k * ct * a.pts.z[p], &sp, &cp);
    const double wr = a.pts.w[p] * a.pts.rho[p];
    a.E[t] = make_double2(wr * cp, wr * sp);
  }
}

// j_n = i^n J_n(u) with n >= 0 (j_{-n} = j_n for integer n), from v = J_n
__device__ __forceinline__ double2 jn_phase(int n, double v) {
  switch (n & 3) {
    case 0: return make_double2(v, 0.0);
    case 1: return make_double2(0.0, v);
    case 2: return make_double2(-v, 0.0);
    default: return make_double2(0.0, -v);
  }
}
__device__ __forceinline__ double2 jn_val(const double* J, int n, size_t stride) {
  return jn_phase(n, J[(size_t)n * stride]);
}

// RHS columns theta-major then (VV, HH), written into the augmented system
// columns [n + col, n + col + nrhs) (bor_solver.py:750-752) when Abase is
// set, and into the FP64 copy B64 when that is set.  One thread = (node,
// aspect) over a chunk of MC consecutive systems (modes m, m+1, ...): it
// writes both of the node's rows (its J_t basis and its J_phi basis, where
// the mode's mask keeps them), and the Bessel values j_{m-1}, j_m, j_{m+1}
// of the node's 8 points slide through registers from mode to mode, so every
// Bessel table entry is read once per node and chunk.
template <typename Ra, int MC>
__global__ __launch_bounds__(128) void k_rhs(ExciteArgs a, const SysInfo* sys, int nsys, void* Abase,
                                             double2* B64) {
  typedef typename CT<Ra>::c Ca;
  const int node = blockIdx.y;
  const int it = blockIdx.x * blockDim.x + threadIdx.x;   // aspect: columns 2 it (VV), 2 it + 1 (HH)
  const int i0 = blockIdx.z * MC, i1 = min(nsys, i0 + MC);
  if (it >= a.ntheta || i0 >= i1) return;
  const double th = a.theta[it];
  double st, ct;
  sincos(th, &st, &ct);
  const size_t stride = (size_t)a.P * a.ntheta;
  const PointSet& P = a.pts;
  const int Ne = a.Nn - 1;
  const double ie = 1.0 / 376.730313668;
  // the node's points: element node - 1 (the node is its T1 end), then
  // element node (T0 end), 4 Gauss points each
  double T[8], tr[8], tz[8];
  double2 E[8];
  const double* Jp[8];
  bool ok[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int e = node - 1 + (q >> 2);
    ok[q] = e >= 0 && e < Ne;
    const int p = ok[q] ? 4 * e + (q & 3) : 0;
    T[q] = q < 4 ? P.T1[p] : P.T0[p];
    tr[q] = P.trho[p];
    tz[q] = P.tz[p];
    E[q] = a.E[(size_t)p * a.ntheta + it];   // w rho e^{j k cos(th) z}
    Jp[q] = a.J + (size_t)p * a.ntheta + it;
  }
  // J_|m-1|, J_|m|, J_|m+1| of each point for the chunk's first mode
  const int mfirst = sys[i0].m;
  double w0[8], w1[8], w2[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    w0[q] = Jp[q][(size_t)abs(mfirst - 1) * stride];
    w1[q] = Jp[q][(size_t)abs(mfirst) * stride];
    w2[q] = Jp[q][(size_t)abs(mfirst + 1) * stride];
  }
  for (int i = i0; i < i1; ++i) {
    const SysInfo s = sys[i];
    const int m = s.m;
    const int am = abs(m), am1 = abs(m - 1), ap1 = abs(m + 1);
    double2 acT[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
    double2 acF[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if (!ok[q]) continue;
      const double2 jm = jn_phase(am, w1[q]), jm1 = jn_phase(am1, w0[q]), jp1 = jn_phase(ap1, w2[q]);
      // Ic = pi (j_{m-1} + j_{m+1}),  Is = -j pi (j_{m-1} - j_{m+1}),  I1 = 2 pi j_m
      const double2 Ic = make_double2(GH_PI * (jm1.x + jp1.x), GH_PI * (jm1.y + jp1.y));
      const double2 Is = make_double2(GH_PI * (jm1.y - jp1.y), -GH_PI * (jm1.x - jp1.x));
      const double2 I1 = make_double2(2.0 * GH_PI * jm.x, 2.0 * GH_PI * jm.y);
      // u = ct tr Ic - st tz I1 (the theta-directed incident field on a t basis)
      const double2 u = make_double2(ct * tr[q] * Ic.x - st * tz[q] * I1.x, ct * tr[q] * Ic.y - st * tz[q] * I1.y);
      double2 tE[2], tH[2], fE[2], fH[2];   // [VV, HH] of the J_t and the J_phi basis
      tE[0] = u;                                          tH[0] = make_double2(Ic.x * ie, Ic.y * ie);
      tE[1] = make_double2(tr[q] * Is.x, tr[q] * Is.y);   tH[1] = make_double2(ct * Is.x * ie, ct * Is.y * ie);
      fE[0] = make_double2(-ct * Is.x, -ct * Is.y);       fH[0] = make_double2(-tr[q] * Is.x * ie, -tr[q] * Is.y * ie);
      fE[1] = Ic;                                         fH[1] = make_double2(u.x * ie, u.y * ie);
#pragma unroll
      for (int pol = 0; pol < 2; ++pol) {
        const double2 vt = make_double2(a.cE * tE[pol].x + a.cM * tH[pol].x, a.cE * tE[pol].y + a.cM * tH[pol].y);
        const double2 vf = make_double2(a.cE * fE[pol].x + a.cM * fH[pol].x, a.cE * fE[pol].y + a.cM * fH[pol].y);
        acT[pol].x += T[q] * (vt.x * E[q].x - vt.y * E[q].y);
        acT[pol].y += T[q] * (vt.x * E[q].y + vt.y * E[q].x);
        acF[pol].x += T[q] * (vf.x * E[q].x - vf.y * E[q].y);
        acF[pol].y += T[q] * (vf.x * E[q].y + vf.y * E[q].x);
      }
    }
    if (2 * it < s.nrhs) {
      // a PMCHWT pair also takes eta0 <W, H_inc> in rows n1 + r: by duality
      // -<W, E_inc>(HH) in the VV column and <W, E_inc>(VV) in the HH column
      // (bor_solver.py:647-657)
      const int n1 = s.pen ? s.pen : s.n;
      const int nf = n1 - s.nt;
      const int rt = node - s.t0, rf = node - s.f0;
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        int r;
        if (b == 0) { if (rt < 0 || rt >= s.nt) continue; r = rt; }
        else { if (rf < 0 || rf >= nf) continue; r = s.nt + rf; }
        const double2* acc = b == 0 ? acT : acF;
        const double2 h0 = make_double2(-acc[1].x, -acc[1].y), h1 = acc[0];
        if (Abase) {
          Ca* A = (Ca*)Abase + s.a_off + (size_t)r * s.ld + s.n + s.col + 2 * it;
          Ca o0, o1;
          o0.x = (Ra)acc[0].x; o0.y = (Ra)acc[0].y;
          o1.x = (Ra)acc[1].x; o1.y = (Ra)acc[1].y;
          A[0] = o0;
          A[1] = o1;
          if (s.pen) {
            Ca* AH = A + (size_t)n1 * s.ld;
            o0.x = (Ra)h0.x; o0.y = (Ra)h0.y;
            o1.x = (Ra)h1.x; o1.y = (Ra)h1.y;
            AH[0] = o0;
            AH[1] = o1;
          }
        }
        if (B64) {
          double2* bb = B64 + s.x_off + (size_t)r * s.ncols + s.col + 2 * it;
          bb[0] = acc[0];
          bb[1] = acc[1];
          if (s.pen) {
            bb[(size_t)n1 * s.ncols] = h0;
            bb[(size_t)n1 * s.ncols + 1] = h1;
          }
        }
      }
    }
    // slide the Bessel window one order up (a problem's systems are its
    // consecutive modes; any other next mode reloads the window)
    if (i + 1 < i1) {
      const int mn = sys[i + 1].m;
      if (mn == m + 1) {
        const int an = abs(m + 2);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          w0[q] = w1[q];
          w1[q] = w2[q];
          w2[q] = Jp[q][(size_t)an * stride];
        }
      } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          w0[q] = Jp[q][(size_t)abs(mn - 1) * stride];
          w1[q] = Jp[q][(size_t)abs(mn) * stride];
          w2[q] = Jp[q][(size_t)abs(mn + 1) * stride];
        }
      }
    }
  }
}

// One Gauss point's far-field term w rho e^{j k cos(theta) z} x (projection of
// the currents of its element's two nodes) for system s, aspect it, pol
// (bor_solver.py:659-708; the far-field pairs (j_{m+1}, j_{m-1}) are reversed
// relative to the excitation): node values t0v/t1v (J_t), f0v/f1v (J_phi)
// and, for a PMCHWT pair, mt0/mt1/mf0/mf1 (M / eta0).
__device__ __forceinline__ double2 ff_point_term(const ExciteArgs& a, const SysInfo& s, int p, int e, int it,
                                                 int pol, double st, double ct, size_t stride, double2 t0v,
                                                 double2 t1v, double2 f0v, double2 f1v, double2 mt0, double2 mt1,
                                                 double2 mf0, double2 mf1) {
  const PointSet& P = a.pts;
  const double T0 = P.T0[p], T1 = P.T1[p];
  const double2 Jt = make_double2(T0 * t0v.x + T1 * t1v.x, T0 * t0v.y + T1 * t1v.y);
  const double2 Jf = make_double2(T0 * f0v.x + T1 * f1v.x, T0 * f0v.y + T1 * f1v.y);
  const double* Jp = a.J + (size_t)p * a.ntheta + it;
  const double2 jm = jn_val(Jp, abs(s.m), stride);
  const double2 jm1 = jn_val(Jp, abs(s.m - 1), stride);
  const double2 jp1 = jn_val(Jp, abs(s.m + 1), stride);
  // Icos = pi (j_{m+1} + j_{m-1}),  Isin = -j pi (j_{m+1} - j_{m-1}),  I1 = 2 pi j_m
  const double2 Icos = make_double2(GH_PI * (jp1.x + jm1.x), GH_PI * (jp1.y + jm1.y));
  const double2 Isin = make_double2(GH_PI * (jp1.y - jm1.y), -GH_PI * (jp1.x - jm1.x));
  const double2 I1 = make_double2(2.0 * GH_PI * jm.x, 2.0 * GH_PI * jm.y);
  const double tr = P.trho[p], tz = P.tz[p];
  double2 term;
  if (pol == 0) {
    // Jt (ct tr Icos - st tz I1) - Jf ct Isin
    const double2 u = make_double2(ct * tr * Icos.x - st * tz * I1.x, ct * tr * Icos.y - st * tz * I1.y);
    term = make_double2(Jt.x * u.x - Jt.y * u.y - ct * (Jf.x * Isin.x - Jf.y * Isin.y),
                        Jt.x * u.y + Jt.y * u.x - ct * (Jf.x * Isin.y + Jf.y * Isin.x));
  } else {
    // Jt tr Isin + Jf Icos
    term = make_double2(tr * (Jt.x * Isin.x - Jt.y * Isin.y) + Jf.x * Icos.x - Jf.y * Icos.y,
                        tr * (Jt.x * Isin.y + Jt.y * Isin.x) + Jf.x * Icos.y + Jf.y * Icos.x);
  }
  if (a.zs) {
    // M_t = Z_s J_phi, M_phi = -Z_s J_t; with the common -j k eta0 / (4 pi):
    // F_theta += (1/eta0) proj_phi(M) = (Z_s/eta0) (Jf tr Isin - Jt Icos),
    // F_phi   -= (1/eta0) proj_theta(M) = -(Z_s/eta0) (Jf u + Jt ct Isin)
    const double2 z = a.zs[e];
    const double ie = 1.0 / 376.730313668;
    double2 mt;
    if (pol == 0) {
      mt = make_double2(tr * (Jf.x * Isin.x - Jf.y * Isin.y) - (Jt.x * Icos.x - Jt.y * Icos.y),
                        tr * (Jf.x * Isin.y + Jf.y * Isin.x) - (Jt.x * Icos.y + Jt.y * Icos.x));
    } else {
      const double2 u = make_double2(ct * tr * Icos.x - st * tz * I1.x, ct * tr * Icos.y - st * tz * I1.y);
      mt = make_double2(-(Jf.x * u.x - Jf.y * u.y) - ct * (Jt.x * Isin.x - Jt.y * Isin.y),
                        -(Jf.x * u.y + Jf.y * u.x) - ct * (Jt.x * Isin.y + Jt.y * Isin.x));
    }
    term.x += ie * (z.x * mt.x - z.y * mt.y);
    term.y += ie * (z.x * mt.y + z.y * mt.x);
  }
  if (s.pen) {
    // an independent magnetic current M = eta0 M' (bor_solver.py:697-707):
    // F_theta += proj_phi(M'), F_phi -= proj_theta(M') under the common
    // -j k eta0 / (4 pi)
    const double2 Mt = make_double2(T0 * mt0.x + T1 * mt1.x, T0 * mt0.y + T1 * mt1.y);
    const double2 Mf = make_double2(T0 * mf0.x + T1 * mf1.x, T0 * mf0.y + T1 * mf1.y);
    if (pol == 0) {
      term.x += tr * (Mt.x * Isin.x - Mt.y * Isin.y) + Mf.x * Icos.x - Mf.y * Icos.y;
      term.y += tr * (Mt.x * Isin.y + Mt.y * Isin.x) + Mf.x * Icos.y + Mf.y * Icos.x;
    } else {
      const double2 u = make_double2(ct * tr * Icos.x - st * tz * I1.x, ct * tr * Icos.y - st * tz * I1.y);
      term.x += -(Mt.x * u.x - Mt.y * u.y) + ct * (Mf.x * Isin.x - Mf.y * Isin.y);
      term.y += -(Mt.x * u.y + Mt.y * u.x) + ct * (Mf.x * Isin.y + Mf.y * Isin.x);
    }
  }
  const double2 E = a.E[(size_t)p * a.ntheta + it];   // w rho e^{j k cos(th) z}
  return make_double2(term.x * E.x - term.y * E.y, term.x * E.y + term.y * E.x);
}

// Partial far-field sums over a chunk of Gauss points for every (system, column).
// X is read from the mode's columns of the augmented system or of the n x
// ncols solution block at element offset x_off.
template <typename Rx>
__global__ __launch_bounds__(256) void k_farfield_partial(ExciteArgs a, const SysInfo* sys,
                                                          const void* Xbase, int x_in_aug, int chunk,
                                                          double2* partial, int nsys) {
  typedef typename CT<Rx>::c Cx;
  const int si = blockIdx.y;
  const SysInfo s = sys[si];
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= s.nrhs) return;
  const int pc = blockIdx.z;
  const int p0 = pc * chunk, p1 = min(a.P, p0 + chunk);
  const int it = col >> 1, pol = col & 1;
  const double th = a.theta[it];
  double st, ct;
  sincos(th, &st, &ct);
  const size_t stride = (size_t)a.P * a.ntheta;
  const Cx* X;
  size_t ldx;
  if (x_in_aug) { X = (const Cx*)Xbase + s.a_off + s.n + s.col; ldx = s.ld; }
  else { X = (const Cx*)Xbase + s.x_off + s.col; ldx = s.ncols; }
  // mode -m of a pair reads mode m's currents: x_{-m} = S [x_m(VV), -x_m(HH)]
  const double ft = (s.flip && pol == 1) ? -1.0 : 1.0, ff = (s.flip && pol == 0) ? -1.0 : 1.0;
  const int n1 = s.pen ? s.pen : s.n;
  const int nf = n1 - s.nt;
  const PointSet& P = a.pts;
  // rows: off 0 the electric current J; off n1 (a PMCHWT pair) M / eta0,
  // whose mode -m entries take the opposite signs (x_{-m} = diag(S, -S) ...)
  auto xt = [&](int node, int off, double sg) -> double2 {
    const int r = node - s.t0;
    if (r < 0 || r >= s.nt) return make_double2(0.0, 0.0);
    const Cx v = X[(size_t)(off + r) * ldx + col];
    return make_double2(sg * (double)v.x, sg * (double)v.y);
  };
  auto xf = [&](int node, int off, double sg) -> double2 {
    const int r = node - s.f0;
    if (r < 0 || r >= nf) return make_double2(0.0, 0.0);
    const Cx v = X[(size_t)(off + s.nt + r) * ldx + col];
    return make_double2(sg * (double)v.x, sg * (double)v.y);
  };
  double2 acc = make_double2(0.0, 0.0);
  int e_prev = -2;   // no element yet (element 0 must not look like a successor)
  double2 t0v = make_double2(0, 0), t1v = t0v, f0v = t0v, f1v = t0v;
  double2 mt0 = t0v, mt1 = t0v, mf0 = t0v, mf1 = t0v;   // M / eta0 of a PMCHWT pair
  for (int p = p0; p < p1; ++p) {
    const int e = p >> 2;
    if (e != e_prev) {
      // the next element's first node is this element's second: carry it over
      if (e == e_prev + 1) { t0v = t1v; f0v = f1v; mt0 = mt1; mf0 = mf1; }
      else {
        t0v = xt(e, 0, ft); f0v = xf(e, 0, ff);
        if (s.pen) { mt0 = xt(e, n1, ff); mf0 = xf(e, n1, ft); }
      }
      t1v = xt(e + 1, 0, ft);
      f1v = xf(e + 1, 0, ff);
      if (s.pen) { mt1 = xt(e + 1, n1, ff); mf1 = xf(e + 1, n1, ft); }
      e_prev = e;
    }
    const double2 v = ff_point_term(a, s, p, e, it, pol, st, ct, stride, t0v, t1v, f0v, f1v, mt0, mt1, mf0, mf1);
    acc.x += v.x;
    acc.y += v.y;
  }
  partial[((size_t)pc * nsys + si) * s.nrhs + col] = acc;
}

// Mode pairs: the far field of mode -m equals mode m's (both polarizations),
// dF [2M+1][nrhs] with mode m at row M + m.
__global__ void k_mirror_dF(double2* dF, int M, int nrhs) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M * nrhs) return;
  const int m = 1 + t / nrhs, c = t % nrhs;
  dF[(size_t)(M - m) * nrhs + c] = dF[(size_t)(M + m) * nrhs + c];
}

// Sum the chunk partials and apply -j k eta0 / (4 pi); acc: add to dF (the
// next radiating surface of a multi-region body).  A wavefront takes 16
// (system, column) entries x 4 chunk phases (lane = phase x 16 + entry): each
// lane sums the chunks of its phase in order, the four phases meet by two xor
// shuffles (lanes e, e + 16, e + 32, e + 48): a fixed order, so the sum is
// bitwise reproducible.
__global__ void k_farfield_reduce(const double2* partial, int nchunk, int nsys, int nrhs, double k,
                                  double2* dF, int acc) {
  const int lane = threadIdx.x & 63;
  const int ph = lane >> 4;
  const int t = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 16 + (lane & 15);
  const bool live = t < nsys * nrhs;
  double2 sum = make_double2(0.0, 0.0);
  if (live)
    for (int c = ph; c < nchunk; c += 4) {
      const double2 v = partial[(size_t)c * nsys * nrhs + t];
      sum.x += v.x; sum.y += v.y;
    }
  sum.x += __shfl_xor(sum.x, 16);
  sum.y += __shfl_xor(sum.y, 16);
  sum.x += __shfl_xor(sum.x, 32);
  sum.y += __shfl_xor(sum.y, 32);
  if (!live || ph != 0) return;
  const double pref = k * 376.730313668 / (4.0 * GH_PI);  // -j * pref
  double2 out = make_double2(pref * sum.y, -pref * sum.x);
  if (acc) { out.x += dF[t].x; out.y += dF[t].y; }
  dF[t] = out;
}

}  // namespace ghost
