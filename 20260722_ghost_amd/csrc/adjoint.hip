// Adjoint of the monostatic RCS with respect to the generatrix geometry.
//
// For a real objective J with dJ = Re sum_c a_c dF_c (F_c the far-field
// amplitude of column c = (theta, polarisation), summed over modes), and
// F_c = pref sum_m S_{m,c}(X_m, geometry) with X_m = A_m^{-1} B_m:
//
//   dJ = Re sum_m [ a'_c dS/dgeom . X  +  lambda_m^T (dB_m - dA_m X_m) ]
//   A_m^T lambda_m = G_m,   G_m[:, c] = a'_c dS_{m,c}/dX[:, c],   a' = a pref
//
// The kernels here form G (k_ffvec), undo the factor's row order after the
// transposed triangular solves (k_unpermute), and contract the perturbed
// systems with Q = [-Lambda X^T | Lambda] row- and column-wise
// (k_adj_rowsum / k_adj_colsum) and the per-element far-field partial sums
// with a' (k_adj_ffsum).  The far-field sum follows bor_solver.py:659-708.
#include "bor_common.hpp"

namespace ghost {

// G[r][c] = a'[c] dS_{m,c}/dX[r][c]: the far-field projection of every basis
// function, one thread per (row, column), rows as in the masked system.
__global__ __launch_bounds__(256) void k_ffvec(ExciteArgs a, const SysInfo* sys, const double2* acoef,
                                               double2* G) {
  const SysInfo s = sys[blockIdx.z];
  const int r = blockIdx.y;
  if (r >= s.n) return;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= s.nrhs) return;
  const int it = col >> 1, pol = col & 1;
  const bool phi = r >= s.nt;
  const int node = phi ? s.f0 + (r - s.nt) : s.t0 + r;
  const double th = a.theta[it];
  double st, ct;
  sincos(th, &st, &ct);
  const size_t stride = (size_t)a.P * a.ntheta;
  const PointSet& P = a.pts;
  const int Ne = a.Nn - 1;
  double2 acc = make_double2(0.0, 0.0);
  for (int side = 0; side < 2; ++side) {
    const int e = node - 1 + side;  // node is T1 of element node-1 and T0 of element node
    if (e < 0 || e >= Ne) continue;
    for (int gq = 0; gq < 4; ++gq) {
      const int p = 4 * e + gq;
      const double T = side == 0 ? P.T1[p] : P.T0[p];
      const double* Jp = a.J + (size_t)p * a.ntheta + it;
      const double2 jm = jn_val(Jp, abs(s.m), stride);
      const double2 jm1 = jn_val(Jp, abs(s.m - 1), stride);
      const double2 jp1 = jn_val(Jp, abs(s.m + 1), stride);
      // far-field orders: Icos = pi (j_{m+1} + j_{m-1}), Isin = -j pi (j_{m+1} - j_{m-1}), I1 = 2 pi j_m
      const double2 Icos = make_double2(GH_PI * (jp1.x + jm1.x), GH_PI * (jp1.y + jm1.y));
      const double2 Isin = make_double2(GH_PI * (jp1.y - jm1.y), -GH_PI * (jp1.x - jm1.x));
      const double2 I1 = make_double2(2.0 * GH_PI * jm.x, 2.0 * GH_PI * jm.y);
      const double tr = P.trho[p], tz = P.tz[p];
      double2 u;
      if (pol == 0) {  // VV: Jt (ct tr Icos - st tz I1) - Jf ct Isin
        u = phi ? make_double2(-ct * Isin.x, -ct * Isin.y)
                : make_double2(ct * tr * Icos.x - st * tz * I1.x, ct * tr * Icos.y - st * tz * I1.y);
      } else {         // HH: Jt tr Isin + Jf Icos
        u = phi ? Icos : make_double2(tr * Isin.x, tr * Isin.y);
      }
      const double2 E = a.E[(size_t)p * a.ntheta + it];   // w rho e^{j k cos(th) z}
      acc.x += T * (u.x * E.x - u.y * E.y);
      acc.y += T * (u.x * E.y + u.y * E.x);
    }
  }
  const double2 ac = acoef[col];
  G[s.x_off + (size_t)r * s.nrhs + col] = make_double2(ac.x * acc.x - ac.y * acc.y, ac.x * acc.y + ac.y * acc.x);
}

// dst[perm[r]][c] = src[r][c]: lambda = P^T z after the transposed solves.
__global__ void k_unpermute(const VecOp* ops, const int* perm, const double2* src, double2* dst) {
  const VecOp op = ops[blockIdx.z];
  const int r = blockIdx.y, c = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= op.n || c >= op.N) return;
  const int pr = perm[op.piv + r];
  dst[op.d + (size_t)pr * op.ldd + c] = src[op.s + (size_t)r * op.lds + c];
}

// a'_c of the objective J = sum_c w_c 10 log10(4 pi |F_c|^2) on one problem:
// F_c = sum over the problem's modes of dF[m][c] (fixed order), a_c = w_c
// (20 / ln 10) conj(F_c) / |F_c|^2, a'_c = a_c j pr with pr = -k eta0 / (4 pi).
__global__ void k_objective_acoef(const double2* dF, int nmodes, int nrhs, const double* w, double pr,
                                  double2* out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nrhs) return;
  double fr = 0.0, fi = 0.0;
  for (int m = 0; m < nmodes; ++m) {
    const double2 v = dF[(size_t)m * nrhs + c];
    fr += v.x;
    fi += v.y;
  }
  const double n2 = fr * fr + fi * fi;
  const double f = n2 > 0.0 ? w[c] * (20.0 / 2.302585092994045684) / n2 : 0.0;
  const double ar = f * fr, ai = -f * fi;
  out[c] = make_double2(-ai * pr, ar * pr);
}

// dst32 = (float) src64 (an n x N block)
__global__ void k_cast32(const VecOp* ops, const double2* src, float2* dst) {
  const VecOp op = ops[blockIdx.z];
  const int r = blockIdx.y, c = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= op.n || c >= op.N) return;
  const double2 v = src[op.s + (size_t)r * op.lds + c];
  dst[op.d + (size_t)r * op.ldd + c] = make_float2((float)v.x, (float)v.y);
}

// dst[perm row] (+)= (double) src[r] / scale[c]: the FP32 transposed solves'
// output in the factor's row order, unscaled and put back in system order.
__global__ void k_unpermute_widen(const VecOp* ops, const int* perm, const float2* src, const double* scale,
                                  double2* dst, int add) {
  const VecOp op = ops[blockIdx.z];
  const int r = blockIdx.y, c = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= op.n || c >= op.N) return;
  const int pr = perm[op.piv + r];
  const float2 v = src[op.s + (size_t)r * op.lds + c];
  const double f = 1.0 / scale[op.sc + c];
  double2* p = dst + op.d + (size_t)pr * op.ldd + c;
  if (add) { double2 o = *p; o.x += (double)v.x * f; o.y += (double)v.y * f; *p = o; }
  else *p = make_double2((double)v.x * f, (double)v.y * f);
}

// rs[off + r] = sum_{j < n + nrhs} Q[r][j] A[r][j] for every row r of every system.
__global__ __launch_bounds__(256) void k_adj_rowsum(const SysInfo* sys, const int* row_off, const double2* Q,
                                                    const double2* A, double2* rs) {
  const SysInfo s = sys[blockIdx.y];
  const int r = blockIdx.x;
  if (r >= s.n) return;
  const double2* q = Q + s.a_off + (size_t)r * s.ld;
  const double2* x = A + s.a_off + (size_t)r * s.ld;
  const int nc = s.n + s.nrhs;
  double re = 0.0, im = 0.0;
  for (int j = threadIdx.x; j < nc; j += 256) {
    const double2 u = q[j], v = x[j];
    re += u.x * v.x - u.y * v.y;
    im += u.x * v.y + u.y * v.x;
  }
  __shared__ double red[2][256];
  red[0][threadIdx.x] = re;
  red[1][threadIdx.x] = im;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) rs[row_off[blockIdx.y] + r] = make_double2(red[0][0], red[1][0]);
}

// cs[off + c] = sum_{r < n} Q[r][c] A[r][c] for every column c < n; 64 columns x 4 row groups.
__global__ __launch_bounds__(256) void k_adj_colsum(const SysInfo* sys, const int* row_off, const double2* Q,
                                                    const double2* A, double2* cs) {
  const SysInfo s = sys[blockIdx.z];
  const int c = blockIdx.x * 64 + threadIdx.x, rg = threadIdx.y;
  double re = 0.0, im = 0.0;
  if (c < s.n)
    for (int r = rg; r < s.n; r += 4) {
      const double2 u = Q[s.a_off + (size_t)r * s.ld + c], v = A[s.a_off + (size_t)r * s.ld + c];
      re += u.x * v.x - u.y * v.y;
      im += u.x * v.y + u.y * v.x;
    }
  __shared__ double red[2][4][64];
  red[0][rg][threadIdx.x] = re;
  red[1][rg][threadIdx.x] = im;
  __syncthreads();
  if (rg == 0 && c < s.n) {
    re = (red[0][0][threadIdx.x] + red[0][1][threadIdx.x]) + (red[0][2][threadIdx.x] + red[0][3][threadIdx.x]);
    im = (red[1][0][threadIdx.x] + red[1][1][threadIdx.x]) + (red[1][2][threadIdx.x] + red[1][3][threadIdx.x]);
    cs[row_off[blockIdx.z] + c] = make_double2(re, im);
  }
}

// ffs[e] = sum_{system, column} a'[column] part[e][system][column]: the
// far-field partial sums of one problem taken one element (4 points) per chunk.
__global__ __launch_bounds__(256) void k_adj_ffsum(const double2* part, const double2* acoef, int nsys, int nrhs,
                                                   double2* ffs) {
  const int e = blockIdx.x;
  const size_t tot = (size_t)nsys * nrhs;
  const double2* pe = part + (size_t)e * tot;
  double re = 0.0, im = 0.0;
  for (size_t t = threadIdx.x; t < tot; t += 256) {
    const double2 u = pe[t], a = acoef[t % nrhs];
    re += a.x * u.x - a.y * u.y;
    im += a.x * u.y + a.y * u.x;
  }
  __shared__ double red[2][256];
  red[0][threadIdx.x] = re;
  red[1][threadIdx.x] = im;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) ffs[e] = make_double2(red[0][0], red[1][0]);
}

}  // namespace ghost
