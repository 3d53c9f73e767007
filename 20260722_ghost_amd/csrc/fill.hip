// Modal impedance fill for the BoR-MoM hot path on gfx950.
//
// Far element pairs (|e-f| >= 2): one wavefront per element pair samples the
// azimuthal integrands on the uniform xi grid, folds +-xi (the Green's
// function and the MFIE brackets have definite parity in xi), and turns the
// folded samples into every azimuthal order at once with a cosine/sine
// transform run as a 16x16x4 MFMA contraction against a twiddle table.  The
// 16 point pairs of the element pair are contracted to the 2x2 nodal blocks
// in registers (DPP/shuffle reductions) and written as complex Z entries into
// a mode-innermost buffer: Zfar[2Nn][2Nn][M+1] for modes 0..M when the modes
// are paired (mode -m is S Z_m S, S negating the J_phi rows and columns), or
// Zfar[2Nn][2Nn][2M+1] for modes -M..M.
//
// Reference algorithm: bor_kernels.py:132-167 (G_m by FFT), 273-331 (MFIE
// brackets), bor_streaming.py:320-414 (nodal contraction, 9 EFIE order
// primitives, per-mode recombination), bor_solver.py:130-154, 293-345,
// 399-435 (mode blocks), bor_stream_kernel.c:26-106 (integrands).
//
// Near element pairs (|e-f| <= 1): graded-cell quadrature with a sinh core
// and an oscillation-scaled Gauss tail per point pair (bor_kernels.py:170-257,
// 360-431; bor_solver.py:54-110, 293-345, 399-435), in FP64.
#include "bor_common.hpp"

namespace ghost {

// ─────────────────────────────────────────────────────────────────────────────
// far fill
// ─────────────────────────────────────────────────────────────────────────────

struct FarArgs {
  PointSet pts;          // test points (rows)
  PointSet spts;         // source points (columns); the same set except in the adjoint's perturbed fills
  const double* sin2E;   // [nE/2+1]  sin^2(pi j / nE)
  const double* cosE;    // [nE]      cos(2 pi i / nE)
  const double* cxB;     // [nB/2+1]  cos(2 pi j / nB)
  const double* sxB;     // [nB/2+1]  sin(2 pi j / nB)
  const double* cosB;    // [nB]
  const double* sinB;    // [nB]
  int Ne, Nn, M, nE, nB, pe, pf, nfp, npairs;
  int Neq, Nnq;          // SPLIT: source elements and nodes (another generatrix for cross-surface operators)
  const unsigned char* skip;  // SPLIT: [Ne][Neq] 1 = near pair (integrated elsewhere); nullptr: |e - f| <= 1
  int neg;               // also write modes -n (off with pairing: they are never stored)
  int delta, imax;       // unordered pairs (not SPLIT): f = pf + 2 j with j >= i + delta, rows i < imax
  int zmodes, zoff;      // Z mode stride and the index of mode 0 (multiples of 16; zoff = 0 with pairing)
  double k, ce_re, ce_im, cm;
  double ki;             // CK: imaginary part of k (a lossy medium; k is the real part)
  const double2* zs;     // IBC: per-element surface impedance (source element's applies)
  double cz;             // IBC: bracket coefficient 2 pi c_ibc (times zs of the source element)
  void* Z;               // Zfar [2Nn][2Nnq][zmodes] complex<R>
  void* Z2;              // IBC: bracket part into its own far blocks (same layout), nullptr: summed into Z
};

// Reduce-scatter step between two lanes of a DPP row whose lane bit differs:
// a lane keeps the value of the pair selected by its own bit (b) and adds the
// partner's copy of the same one.  CTRL is the partner permutation.
template <int CTRL, typename R>
__device__ __forceinline__ R rs_step(R lo, R hi, bool b) {
  const R keep = b ? hi : lo, send = b ? lo : hi;
  return keep + dpp_mov<CTRL>(send);
}

// Lane exchanges for the (f, e) EFIE blocks of an unordered element pair (see
// k_far_fill): v_permlane16_swap exchanges the odd rows of x with the even
// rows of y (a 2 x 2 transpose of lane bit 4 against the x / y index), and
// half_swap reads the other wavefront half.
__device__ __forceinline__ void perm_swap16(float& x, float& y) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_int(x), __float_as_int(y), false, false);
  x = __int_as_float(r[0]);
  y = __int_as_float(r[1]);
}
__device__ __forceinline__ void perm_swap16(double& x, double& y) {
  const auto a = __builtin_amdgcn_permlane16_swap(__double2loint(x), __double2loint(y), false, false);
  const auto b = __builtin_amdgcn_permlane16_swap(__double2hiint(x), __double2hiint(y), false, false);
  x = __hiloint2double(b[0], a[0]);
  y = __hiloint2double(b[1], a[1]);
}
// value of lane l ^ 32 (the other wavefront half)
__device__ __forceinline__ float half_swap(float x, int lane) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_int(x), __float_as_int(x), false, false);
  return __int_as_float(lane < 32 ? r[1] : r[0]);
}
__device__ __forceinline__ double half_swap(double x, int lane) {
  const auto a = __builtin_amdgcn_permlane32_swap(__double2loint(x), __double2loint(x), false, false);
  const auto b = __builtin_amdgcn_permlane32_swap(__double2hiint(x), __double2hiint(x), false, false);
  return lane < 32 ? __hiloint2double(b[1], a[1]) : __hiloint2double(b[0], a[0]);
}

// Wavefronts (element pairs) per work-group of k_far_fill: they share one copy
// of the azimuth tables in LDS.  FP32 up to 6 order tiles (m_max <= 94): 4,
// with a register budget of three wavefronts per SIMD; otherwise 2 (the FP64
// stash is twice the size, and 7-8 tiles cannot meet that budget).  Those get
// a budget of two wavefronts per SIMD where it costs no spills: FP32, and FP64
// up to 2 tiles.
__host__ __device__ constexpr bool far_wide(int bytes, int nt) { return bytes == 4 && nt <= 6; }
__host__ __device__ constexpr int far_waves(int bytes, int nt) { return far_wide(bytes, nt) ? 4 : 2; }

// One wavefront per far element pair: an unordered pair e < f of one parity
// class (SPLIT = false), or an ordered pair (e, f) when the test and source
// point sets differ (SPLIT = true); far_waves() wavefronts per block.
// Phase 1: EFIE samples -> G_n on MFMA.  Phase 2: per-order EFIE blocks,
// reduced over the test points and stashed in LDS.  Phase 3: MFIE brackets
// contracted to nodal rows -> modes on MFMA.  Phase 4: combine, RMW Zfar.
// Unordered pairs run phases 3 and 4 twice, for (e, f) and then (f, e); the
// second phase 4 takes its EFIE blocks from the same stash by lane exchanges.
// WPB = 1 is the fallback for FP64 azimuth tables too large to sit beside a
// two-wavefront stash in LDS (n_xi = 2048 with 7-8 order tiles).  IBC: the
// bracket phase samples the rotated-PV brackets of the impedance operator
// (bor_kernels.py:449-478) in place of the MFIE ones, with the coefficient
// 2 pi c_ibc Z_s(source element) (bor_solver.py:497-518).
// CK: complex k = k + j ki (penetrable media, FP64 only): every sample's
// e^{-jkR} carries the decay e^{ki R}, the brackets' (1 + jkR) its imaginary
// part, and the EFIE blocks' 1/k^2 is complex.
template <typename R, int NT, bool SPLIT, bool IBC = false, int WPB = far_waves(sizeof(R), NT), bool CK = false>
__global__ __launch_bounds__(64 * WPB)
__attribute__((amdgpu_waves_per_eu(far_wide(sizeof(R), NT) ? 3 : (sizeof(R) == 4 || NT <= 2) ? 2 : 1))) void k_far_fill(FarArgs a) {
  typedef typename CT<R>::c C2;
  typedef typename CT<R>::v4 V4;
  __shared__ R stash[WPB][NT][8][64];
  extern __shared__ __align__(16) char dyn_smem[];
  // azimuth tables in LDS, in the working precision:
  // (cosB, sinB) pairs [nB] | bracket samples (cx, sx, w, 0) [nB/2+4] |
  // EFIE samples (sin^2(xi/2), w) [nE/2+4] | cosE [nE]
  // The sample tables run 3 entries past N/2 (the last group of 4 samples),
  // with the even-fold weight w = c_j (2 pi/N)/(4 pi), c_j = 1 at j = 0 and
  // N/2, 2 between and 0 past N/2 (where the geometry repeats entry N/2), so
  // a sample takes one table load and no selects.  The bracket transform
  // reads its cos and sin twiddle in one 2-element load.
  C2* t_csB = (C2*)dyn_smem;
  V4* t_smpB = (V4*)(t_csB + a.nB);
  C2* t_smpE = (C2*)(t_smpB + (a.nB / 2 + 4));
  R* t_cosE = (R*)(t_smpE + (a.nE / 2 + 4));
  {
    const int hE = a.nE / 2, hB = a.nB / 2;
    const R sc = (R)1 / ((R)2 * (R)a.nE), scB = (R)1 / ((R)2 * (R)a.nB);   // (2 pi / N) / (4 pi)
    for (int i = threadIdx.x; i < hE + 4; i += blockDim.x) {
      const R cw = (i == 0 || i == hE) ? (R)1 : (i > hE ? (R)0 : (R)2);
      C2 v;
      v.x = (R)a.sin2E[i < hE ? i : hE];
      v.y = cw * sc;
      t_smpE[i] = v;
    }
    for (int i = threadIdx.x; i < hB + 4; i += blockDim.x) {
      const int ic = i < hB ? i : hB;
      const R cw = (i == 0 || i == hB) ? (R)1 : (i > hB ? (R)0 : (R)2);
      t_smpB[i] = V4{(R)a.cxB[ic], (R)a.sxB[ic], cw * scB, (R)0};
    }
  }
  for (int i = threadIdx.x; i < a.nE; i += blockDim.x) t_cosE[i] = (R)a.cosE[i];
  for (int i = threadIdx.x; i < a.nB; i += blockDim.x) {
    C2 v;
    v.x = (R)a.cosB[i];
    v.y = (R)a.sinB[i];
    t_csB[i] = v;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wl = threadIdx.x >> 6;
  const int wid = blockIdx.x * WPB + wl;
  if (wid >= a.npairs) return;  // wave-uniform exit (no block barriers below)
  int e, f;
  if (SPLIT) {   // ordered pairs (e, f) of the parity class
    e = a.pe + 2 * (wid / a.nfp);
    f = a.pf + 2 * (wid % a.nfp);
    if (e >= a.Ne || f >= a.Neq) return;
    if (a.skip ? a.skip[(size_t)e * a.Neq + f] != 0 : abs(e - f) <= 1) return;
  } else {
    // unordered pairs e < f of the class, row i = (e - pe)/2 holding
    // j = (f - pf)/2 from i + delta to nfp - 1: S(i) pairs before row i
    const long long base = a.nfp - a.delta;
    auto S = [&](long long i) { return i * base - i * (i - 1) / 2; };
    int lo = 0, hi = a.imax;   // S(lo) <= wid < S(hi)
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (S(mid) <= wid) lo = mid; else hi = mid;
    }
    e = a.pe + 2 * lo;
    f = a.pf + 2 * (lo + a.delta + (int)(wid - S(lo)));
  }

  const int c16 = lane & 15, kk = lane >> 4;
  const R k = (R)a.k;
  const PointSet& P = a.pts;
  const PointSet& Q = SPLIT ? a.spts : a.pts;   // source points: the test set unless split

  // operand lane: the point pair (g, h) whose result lands in lane group
  // kk = g, accumulator element r = h (row mrow(g, h) = c16)
  const bool f64 = sizeof(R) == 8;
  const int g_in = f64 ? (c16 & 3) : (c16 >> 2), h_in = f64 ? (c16 >> 2) : (c16 & 3);
  const int ip = 4 * e + g_in, iq = 4 * f + h_in;
  const R rp = (R)P.rho[ip], zp = (R)P.z[ip], rq = (R)Q.rho[iq], zq = (R)Q.z[iq];
  const R dz = zp - zq, dr = rp - rq;
  const R d2 = dr * dr + dz * dz, rr4 = (R)4 * rp * rq;
  const int uvsub = kk >> 1, asel = kk & 1;

  // ── phase 4: Z += ce * EFIE (the stash) + cm * K (accE, accO) of the pair
  //    (test te, source se) for modes +n (and -n unpaired).  With a
  //    transposed stash the EFIE blocks come from the opposite pair: the
  //    EFIE operator is symmetric, Z(f, e) = Z(e, f)^T with tf(f, e) =
  //    -ft(e, f)^T and ft(f, e) = -tf(e, f)^T (the test and source weights of
  //    every block term factorise and G, Gc, Gs are symmetric in the point
  //    pair), so lane (uvsub, a, n) takes register b of lane (uvsub, b, n)
  //    register a for tt / ff, and of the other half (ft <-> tf), negated ──
  auto write_z = [&](int te, int se, const V4* kE, const V4* kO, bool transposed) {
  const R ce_re = (R)a.ce_re, ce_im = (R)a.ce_im, cm = (R)a.cm;
  R cmi = (R)0, cmr = cm;
  if constexpr (IBC) {
    const double2 z = a.zs[se];
    cmr = (R)(a.cz * z.x);
    cmi = (R)(a.cz * z.y);
  }
  const size_t ncol = 2 * (size_t)a.Nnq;
  C2* Z = (C2*)a.Z;
  const int rowE = (uvsub ? a.Nn : 0) + te + asel;   // tt / ff, and tf / ft rows
  const int colE = (uvsub ? a.Nnq : 0) + se;         // tt / ff column (+ b)
  const int colO = (uvsub ? 0 : a.Nnq) + se;         // tf / ft column (+ b)
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int n = 16 * t + c16;
    if (n > a.M) continue;
    // the tile's EFIE values [b][reim], even (tt / ff) and odd (tf / ft) blocks
    R ev[4], ov[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) { ev[i] = stash[wl][t][i][lane]; ov[i] = stash[wl][t][4 + i][lane]; }
    if (transposed) {
      perm_swap16(ev[0], ev[2]);
      perm_swap16(ev[1], ev[3]);
      perm_swap16(ov[0], ov[2]);
      perm_swap16(ov[1], ov[3]);
#pragma unroll
      for (int i = 0; i < 4; ++i) ov[i] = -half_swap(ov[i], lane);
    }
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) {
      const R er = ev[2 * bi], ei = ev[2 * bi + 1];
      const R orr = ov[2 * bi], oi = ov[2 * bi + 1];
      const R Ker = kE[t][2 * bi], Kei = kE[t][2 * bi + 1];
      const R Kor = kO[t][2 * bi + 1], Koi = -kO[t][2 * bi];  // -j (re + j im)
      R zer, zei, zor, zoi;
      if constexpr (IBC) {
        const R ber = cmr * Ker - cmi * Kei, bei = cmr * Kei + cmi * Ker;
        const R bor = cmr * Kor - cmi * Koi, boi = cmr * Koi + cmi * Kor;
        if (a.Z2) {
          // the bracket operator on its own (a penetrable body's coupling blocks)
          C2* Zb = (C2*)a.Z2;
          C2* qe = Zb + ((size_t)rowE * ncol + colE + bi) * a.zmodes + a.zoff;
          C2* qo = Zb + ((size_t)rowE * ncol + colO + bi) * a.zmodes + a.zoff;
          C2 v;
          v = qe[n]; v.x += ber; v.y += bei; qe[n] = v;
          v = qo[n]; v.x += bor; v.y += boi; qo[n] = v;
          if (n > 0 && a.neg) {
            v = qe[-n]; v.x += ber; v.y += bei; qe[-n] = v;
            v = qo[-n]; v.x -= bor; v.y -= boi; qo[-n] = v;
          }
          zer = ce_re * er - ce_im * ei;
          zei = ce_re * ei + ce_im * er;
          zor = ce_re * orr - ce_im * oi;
          zoi = ce_re * oi + ce_im * orr;
        } else {
          zer = ce_re * er - ce_im * ei + ber;
          zei = ce_re * ei + ce_im * er + bei;
          zor = ce_re * orr - ce_im * oi + bor;
          zoi = ce_re * oi + ce_im * orr + boi;
        }
      } else {
        zer = ce_re * er - ce_im * ei + cm * Ker;
        zei = ce_re * ei + ce_im * er + cm * Kei;
        zor = ce_re * orr - ce_im * oi + cm * Kor;
        zoi = ce_re * oi + ce_im * orr + cm * Koi;
      }
      C2* pe_ = Z + ((size_t)rowE * ncol + colE + bi) * a.zmodes + a.zoff;
      C2* po_ = Z + ((size_t)rowE * ncol + colO + bi) * a.zmodes + a.zoff;
      C2 v;
      v = pe_[n]; v.x += zer; v.y += zei; pe_[n] = v;
      v = po_[n]; v.x += zor; v.y += zoi; po_[n] = v;
      if (n > 0 && a.neg) {
        v = pe_[-n]; v.x += zer; v.y += zei; pe_[-n] = v;
        v = po_[-n]; v.x -= zor; v.y -= zoi; po_[-n] = v;
      }
    }
  }
  };

  // ── phase 1: G_n = sum_j c_j (2pi/N) g(2 pi j/N) cos(2 pi n j / N) ──
  {
    V4 accR[NT], accI[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) { accR[t] = V4{0, 0, 0, 0}; accI[t] = V4{0, 0, 0, 0}; }
    {
      const int hE = a.nE >> 1, mskE = a.nE - 1;
      for (int j0 = 0; j0 <= hE; j0 += 4) {
        const int j = j0 + kk;
        const C2 smp = t_smpE[j];   // (sin^2(xi_j/2), weight)
        R Rd = sqrt_s(d2 + rr4 * smp.x);
        Rd = Rd > (R)1e-30 ? Rd : (R)1e-30;
        R s, c;
        sincos_s(k * Rd, &s, &c);
        R amp = smp.y * rcp_s(Rd);
        if constexpr (CK) amp *= exp((R)a.ki * Rd);
        const R ar = amp * c, ai = -amp * s;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const R bv = t_cosE[((16 * t + c16) * j) & mskE];
          accR[t] = mfma4(ar, bv, accR[t]);
          accI[t] = mfma4(ai, bv, accI[t]);
        }
      }
    }

    // ── phase 2: EFIE blocks per order of the pair (test te, source se); lane
    //    holds G[g = kk][h = r][n = 16 t + c16] ──
    auto efie_blocks = [&](int te, int se) {
    const int ig = 4 * te + kk;
    const R trp = (R)P.trho[ig], tzp = (R)P.tz[ig];
    const R wrp = (R)(P.w[ig] * P.rho[ig]);
    const R Twa[2] = {(R)P.T0[ig] * wrp, (R)P.T1[ig] * wrp};
    const R Swa[2] = {(R)(P.T0[ig] * P.w[ig]), (R)(P.T1[ig] * P.w[ig])};
    const R Dwa[2] = {(R)(P.D0[ig] * P.w[ig]), (R)(P.D1[ig] * P.w[ig])};
    const R trq = (R)Q.trho[4 * se], tzq = (R)Q.tz[4 * se];
    // 1/k^2: a real factor of the divergence terms, or (CK) complex and
    // applied to the source sums sDG, sSG below
    R ik2 = (R)1 / (k * k), i2r = ik2, i2i = (R)0;
    if constexpr (CK) {
      const R kr_ = k, ki_ = (R)a.ki, den = (kr_ * kr_ + ki_ * ki_) * (kr_ * kr_ + ki_ * ki_);
      i2r = (kr_ * kr_ - ki_ * ki_) / den;
      i2i = -2 * kr_ * ki_ / den;
      ik2 = (R)1;
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const R am = (R)(16 * t + c16);
      // The weights are separable (test side: a and the lane's g; source side:
      // b and h), so the source points are contracted first: per source node b
      // ten sums over h of the order data (Tw G, Tw Gc, Tw Gs, Dw G, Sw G; re
      // and im), then the 2 x 2 x 4 blocks from those sums and the test weights:
      //   tt = Tw_a (trp trq Gc + tzp tzq G)_b - Dw_a/k^2 (Dw G)_b
      //   tf = Tw_a trp (Tw Gs)_b - j am Dw_a/k^2 (Sw G)_b
      //   ft = -Tw_a trq (Tw Gs)_b + j am Sw_a/k^2 (Dw G)_b
      //   ff = Tw_a (Tw Gc)_b - am^2 Sw_a/k^2 (Sw G)_b
      R sTG[2][2], sTC[2][2], sTS[2][2], sDG[2][2], sSG[2][2];   // [b][reim]
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int ri = 0; ri < 2; ++ri) { sTG[bi][ri] = 0; sTC[bi][ri] = 0; sTS[bi][ri] = 0; sDG[bi][ri] = 0; sSG[bi][ri] = 0; }
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const int iq2 = 4 * se + h;
        const R wr = (R)(Q.w[iq2] * Q.rho[iq2]);
        const R Twb[2] = {(R)Q.T0[iq2] * wr, (R)Q.T1[iq2] * wr};
        const R Swb[2] = {(R)(Q.T0[iq2] * Q.w[iq2]), (R)(Q.T1[iq2] * Q.w[iq2])};
        const R Dwb[2] = {(R)(Q.D0[iq2] * Q.w[iq2]), (R)(Q.D1[iq2] * Q.w[iq2])};
        const R Gr = accR[t][h], Gi = accI[t][h];
        // neighbour orders by row rotation: lane c receives lane c-1 (ror 1) / c+1 (ror 15)
        R Pr = row_ror<15>(accR[t][h]), Pi = row_ror<15>(accI[t][h]);  // G_{n+1}
        R Mr = row_ror<1>(accR[t][h]), Mi = row_ror<1>(accI[t][h]);    // G_{n-1}
        if (t + 1 < NT) {
          const R xr = row_ror<15>(accR[(t + 1) % NT][h]), xi = row_ror<15>(accI[(t + 1) % NT][h]);
          if (c16 == 15) { Pr = xr; Pi = xi; }
        }
        if (t > 0) {
          const R xr = row_ror<1>(accR[(t + NT - 1) % NT][h]), xi = row_ror<1>(accI[(t + NT - 1) % NT][h]);
          if (c16 == 0) { Mr = xr; Mi = xi; }
        } else if (c16 == 0) {
          Mr = Pr; Mi = Pi;  // G_{|-1|} = G_1
        }
        const R Gcr = (R)0.5 * (Mr + Pr), Gci = (R)0.5 * (Mi + Pi);
        const R Gsr = (R)0.5 * (Mi - Pi), Gsi = (R)-0.5 * (Mr - Pr);  // (M - P)/(2j)
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) {
          sTG[bi][0] += Twb[bi] * Gr;  sTG[bi][1] += Twb[bi] * Gi;
          sTC[bi][0] += Twb[bi] * Gcr; sTC[bi][1] += Twb[bi] * Gci;
          sTS[bi][0] += Twb[bi] * Gsr; sTS[bi][1] += Twb[bi] * Gsi;
          sDG[bi][0] += Dwb[bi] * Gr;  sDG[bi][1] += Dwb[bi] * Gi;
          sSG[bi][0] += Swb[bi] * Gr;  sSG[bi][1] += Swb[bi] * Gi;
        }
      }
      if constexpr (CK) {
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) {
          R x = sDG[bi][0], y = sDG[bi][1];
          sDG[bi][0] = i2r * x - i2i * y;
          sDG[bi][1] = i2r * y + i2i * x;
          x = sSG[bi][0]; y = sSG[bi][1];
          sSG[bi][0] = i2r * x - i2i * y;
          sSG[bi][1] = i2r * y + i2i * x;
        }
      }
      const R tt_c = trp * trq, tt_z = tzp * tzq;
      R z[32];  // [blk(4)][a(2)][b(2)][reim]: blk 0 tt, 1 tf, 2 ft, 3 ff
#pragma unroll
      for (int ai = 0; ai < 2; ++ai) {
        const R kT = Twa[ai], kD = Dwa[ai] * ik2, kS = Swa[ai] * ik2;
        const R kDm = kD * am, kSm = kS * am, kSmm = kSm * am;
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) {
          const int ab = 2 * ai + bi;
          const R Kr = tt_c * sTC[bi][0] + tt_z * sTG[bi][0], Ki = tt_c * sTC[bi][1] + tt_z * sTG[bi][1];
          z[0 * 8 + 2 * ab] = kT * Kr - kD * sDG[bi][0];                        // tt
          z[0 * 8 + 2 * ab + 1] = kT * Ki - kD * sDG[bi][1];
          z[1 * 8 + 2 * ab] = kT * trp * sTS[bi][0] + kDm * sSG[bi][1];          // tf
          z[1 * 8 + 2 * ab + 1] = kT * trp * sTS[bi][1] - kDm * sSG[bi][0];
          z[2 * 8 + 2 * ab] = -kT * trq * sTS[bi][0] - kSm * sDG[bi][1];         // ft
          z[2 * 8 + 2 * ab + 1] = -kT * trq * sTS[bi][1] + kSm * sDG[bi][0];
          z[3 * 8 + 2 * ab] = kT * sTC[bi][0] - kSmm * sSG[bi][0];               // ff
          z[3 * 8 + 2 * ab + 1] = kT * sTC[bi][1] - kSmm * sSG[bi][1];
        }
      }
      // reduce over g (lanes c16 + 16 kk): stage 1 (wavefront halves, uvsub =
      // lane bit 5) keeps {tt, tf} (uvsub 0) or {ff, ft}
      R y16[16];
#pragma unroll
      for (int i = 0; i < 8; ++i) y16[i] = xadd32(z[0 * 8 + i], z[3 * 8 + i]);       // tt <-> ff
#pragma unroll
      for (int i = 0; i < 8; ++i) y16[8 + i] = xadd32(z[1 * 8 + i], z[2 * 8 + i]);   // tf <-> ft
      // stage 2 (row pairs, asel = lane bit 4) keeps a = asel; stash (even, odd) x (b, reim)
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int i = 0; i < 4; ++i) stash[wl][t][4 * q + i][lane] = xadd16(y16[8 * q + i], y16[8 * q + 4 + i]);
    }
    };
    efie_blocks(e, f);
  }

  // ── phase 3: MFIE brackets of the pair (test te, source se), contracted to
  //    nodal (a, b) rows before the transform ──
  V4 accE[NT], accO[NT];
  auto mfie_brackets = [&](int te, int se) {
#pragma unroll
  for (int t = 0; t < NT; ++t) { accE[t] = V4{0, 0, 0, 0}; accO[t] = V4{0, 0, 0, 0}; }
  {
    const int ip = 4 * te + g_in, iq = 4 * se + h_in;
    const R rp = (R)P.rho[ip], zp = (R)P.z[ip], rq = (R)Q.rho[iq], zq = (R)Q.z[iq];
    const R dz = zp - zq;
    const R trp = (R)P.trho[ip], tzp = (R)P.tz[ip], trq = (R)Q.trho[iq], tzq = (R)Q.tz[iq];
    const R wp = (R)(P.w[ip] * P.rho[ip]), wq = (R)(Q.w[iq] * Q.rho[iq]);
    // The nodal weight of the point pair factorises, W_ab = Wa_a(g) Wb_b(h), so
    // the sum over the 16 point pairs of a DPP row runs as two 2-bit
    // reduce-scatters: over h with Wb (scattering b and reim), then over g
    // with Wa (scattering a and uvsub).  The lane bits of (a, uvsub, b, reim)
    // follow the accumulator row of the transform, mrow(a + 2 uvsub, reim + 2 b):
    // FP32 uvsub 3, a 2, b 1, reim 0 (g = bits 3:2, h = bits 1:0); FP64 b 3,
    // reim 2, uvsub 1, a 0 (h = bits 3:2, g = bits 1:0).  Partners: bit 0
    // quad_perm [1,0,3,2], bit 1 quad_perm [2,3,0,1], bit 2 row_ror 4 (run
    // while bit 3 is still unscattered: the partner i +- 4 may differ there),
    // bit 3 row_ror 8.
    const bool f32 = sizeof(R) == 4;
    const int bb = f32 ? (lane >> 1) & 1 : (lane >> 3) & 1;    // b
    const int ab = f32 ? (lane >> 2) & 1 : lane & 1;           // a
    const bool rb = f32 ? (lane & 1) : ((lane >> 2) & 1);       // reim
    const bool ub = f32 ? ((lane >> 3) & 1) : ((lane >> 1) & 1);  // uvsub
    const R Wa[2] = {(R)P.T0[ip] * wp, (R)P.T1[ip] * wp};
    const R Wb[2] = {(R)Q.T0[iq] * wq, (R)Q.T1[iq] * wq};
    const R Wak = Wa[ab], Was = Wa[ab ^ 1], Wbk = Wb[bb], Wbs = Wb[bb ^ 1];
    // the partner's weight for this lane's a (b): its own "other" weight, so
    // a stage is Wk x + Wp dpp(x) (the move folds into the multiply)
    const R Wap = f32 ? dpp_mov<0x124>(Was) : dpp_mov<0xB1>(Was);
    const R Wbp = dpp_mov<0x4E>(Wbs);   // FP32 b stage (bit 1)
    const int hB = a.nB >> 1, mskB = a.nB - 1;
    for (int j0 = 0; j0 <= hB; j0 += 4) {
      const int j = j0 + kk;
      const V4 smp = t_smpB[j];   // (cos xi_j, sin xi_j, weight, 0)
      const R cx = smp[0], sx = smp[1];
      const R Rx = rp - rq * cx, Ry = rq * sx;
      R Rd = sqrt_s(Rx * Rx + Ry * Ry + dz * dz);
      Rd = Rd > (R)1e-30 ? Rd : (R)1e-30;
      R s, c;
      sincos_s(k * Rd, &s, &c);
      const R inv = rcp_s(Rd);
      R pre = smp[2] * inv * inv * inv;
      const R kr = k * Rd;
      R pr, pim;
      if constexpr (CK) {
        // (1 + jkR) e^{-jkR} with k = k + j ki: e^{ki R} ((1 - ki R) + j kR) (c - j s)
        pre *= exp((R)a.ki * Rd);
        const R u = (R)1 - (R)a.ki * Rd;
        pr = pre * (u * c + kr * s);
        pim = pre * (kr * c - u * s);
      } else {
        pr = pre * (c + kr * s);
        pim = pre * (kr * c - s);
      }
      const R Wt_t = trp * trq * cx + tzp * tzq, Wt_f = trp * sx, Wf_t = -trq * sx;
      // brackets x4 = tt, ff (even in xi), tf, ft (odd), times (re, im)
      R bx[4];
      if constexpr (IBC) {
        // p(R) [(W_u . n_q)(R . f_v) - (W_u . f_v)(R . n_q)], source normal n_q
        const R D = rp * cx - rq;
        const R Wt_n = -trp * tzq * cx + tzp * trq, Wf_n = tzq * sx;
        const R R_t = trq * D + tzq * dz, R_f = rp * sx, R_n = -tzq * D + trq * dz;
        bx[0] = Wt_n * R_t - Wt_t * R_n;
        bx[1] = Wf_n * R_f - cx * R_n;
        bx[2] = Wt_n * R_f - Wt_f * R_n;
        bx[3] = Wf_n * R_t - Wf_t * R_n;
      } else {
        // -p(R) [(W_u . R)(n_p . f_v) - (W_u . f_v)(n_p . R)], test normal n_p
        const R WtR = trp * Rx + tzp * dz, nR = -tzp * Rx + trp * dz;
        const R n_t = -tzp * trq * cx + trp * tzq, n_f = -tzp * sx;
        bx[0] = -(WtR * n_t - Wt_t * nR);
        bx[1] = -(Ry * n_f - cx * nR);
        bx[2] = -(WtR * n_f - Wt_f * nR);
        bx[3] = -(Ry * n_t - Wf_t * nR);
      }
      R X[4][2];
#pragma unroll
      for (int x = 0; x < 4; ++x) { X[x][0] = bx[x] * pr; X[x][1] = bx[x] * pim; }
      R Y[4];   // sum over h of Wb_b(h) X[x][reim] for the lane's (b, reim)
      if constexpr (sizeof(R) == 4) {
        R U[4][2];
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
          for (int ri = 0; ri < 2; ++ri) U[x][ri] = Wbk * X[x][ri] + Wbp * dpp_mov<0x4E>(X[x][ri]);  // b: bit 1
#pragma unroll
        for (int x = 0; x < 4; ++x) Y[x] = rs_step<0xB1>(U[x][0], U[x][1], rb);                       // reim: bit 0
      } else {
        R U[2][4];
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
          for (int x = 0; x < 4; ++x) U[bi][x] = rs_step<0x124>(Wb[bi] * X[x][0], Wb[bi] * X[x][1], rb);  // reim: bit 2
#pragma unroll
        for (int x = 0; x < 4; ++x) Y[x] = rs_step<0x128>(U[0][x], U[1][x], bb);                          // b: bit 3
      }
      R V[4];   // then over g with Wa_a(g) for the lane's a
      R ve, vo;
      if constexpr (sizeof(R) == 4) {
#pragma unroll
        for (int x = 0; x < 4; ++x) V[x] = Wak * Y[x] + Wap * dpp_mov<0x124>(Y[x]);  // a: bit 2
        ve = rs_step<0x128>(V[0], V[1], ub);                                          // uvsub: bit 3
        vo = rs_step<0x128>(V[2], V[3], ub);
      } else {
#pragma unroll
        for (int x = 0; x < 4; ++x) V[x] = Wak * Y[x] + Wap * dpp_mov<0xB1>(Y[x]);   // a: bit 0
        ve = rs_step<0x4E>(V[0], V[1], ub);                                           // uvsub: bit 1
        vo = rs_step<0x4E>(V[2], V[3], ub);
      }
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int idx = ((16 * t + c16) * j) & mskB;
        const C2 cs = t_csB[idx];
        accE[t] = mfma4(ve, cs.x, accE[t]);
        accO[t] = mfma4(vo, cs.y, accO[t]);
      }
    }
  }
  };

  mfie_brackets(e, f);
  write_z(e, f, accE, accO, false);
  if (!SPLIT) {   // (f, e): EFIE from the transposed stash, MFIE (no such symmetry) on its own
    mfie_brackets(f, e);
    write_z(f, e, accE, accO, true);
  }
}

// ─────────────────────────────────────────────────────────────────────────────
// near element pairs, stage 1: modal values per cell point pair (FP64)
// ─────────────────────────────────────────────────────────────────────────────

struct NearPairInfo {
  int e, f, kind;   // cell table: same surface 0 diag, 1 corner10 (f = e+1), 2 corner01 (f = e-1);
                    // two surfaces 0 tensor grid, 1 + 2a + b corner<a><b> (shared ends a, b)
  int pp_off;       // first point pair in the NP buffer
  int n_tail;       // Gauss order of the [xi0, pi] tail
  int tail_off;     // offset of that rule in the tail tables
};

struct NearArgs {
  const double* nodes;    // [Nn][2]  test element geometry
  const double* L;        // [Ne]
  const double* trho;     // [Ne]
  const double* tz;       // [Ne]
  const double* snodes;   // source element geometry (the same arrays except in the
  const double* sL;       // adjoint's perturbed fills)
  const double* strho;
  const double* stz;
  const double* cell_s;   // concatenated cell points of the kinds
  const double* cell_sp;
  const double* cell_w;
  int cell_off[8], cell_cnt[8];
  const double* core_u;   // [48] Gauss nodes on [0,1]
  const double* core_w;
  const double* tail_u;   // concatenated tail rules on [0,1]
  const double* tail_w;
  const NearPairInfo* pairs;
  const int2* batches;    // (pair slot, first cell point) per work-group
  int Ne, M, nord;        // nord = M + 2 orders stored per point pair
  int ibc;                // brackets: 0 MFIE, 1 rotated-PV (IBC) (bor_kernels.py:515-560)
  double k;
  double ki;              // imaginary part of k (lossy medium), 0 in free space
  double2* NP;            // [pp][nord][5] : G, Ktt, Ktf, Kft, Kff
};

__device__ __forceinline__ void cell_geom(const double* nodes, int e, double s, double& rho, double& z) {
  const double r0 = nodes[2 * e], z0 = nodes[2 * e + 1];
  const double r1 = nodes[2 * e + 2], z1 = nodes[2 * e + 3];
  rho = r0 + s * (r1 - r0);
  z = z0 + s * (z1 - z0);
}

// exp of the near band's node values in either precision (the decay of a
// lossy medium); sqrt, sincos and reciprocals use the far fill's sample forms
__device__ __forceinline__ double nexp(double x) { return exp(x); }
__device__ __forceinline__ float nexp(float x) { return expf(x); }

// Accumulate one thread's C orders over one batch of 8 nodes: nvs = the
// point pair's node values (12 per node), st0 / st1 = its chunk starts
// e^{j c C xi} (re / im, [node * 9 + chunk]).  EFIE and even brackets
// project on cos(n xi), odd brackets on sin(n xi); cos / sin of the chunk's
// orders follow the three-term recurrence f_{n+1} = 2 cos(xi) f_n - f_{n-1}
// from the two exact starts e^{j n0 xi} and e^{j (n0 + 1) xi} (9 steps: the
// error stays a few ulp of R).
template <int C, typename R>
__device__ __forceinline__ void near_acc8(const double* nvs, const double* st0, const double* st1, int cs,
                                          R (&gr)[C], R (&gi)[C], R (&kr)[4][C], R (&ki)[4][C]) {
  for (int u = 0; u < 8; ++u) {
    const double* xd = nvs + 12 * u;
    R x[10];
#pragma unroll
    for (int t = 0; t < 10; ++t) x[t] = (R)xd[t];
    const R c1 = (R)xd[10], s1 = (R)xd[11], tc = (R)2 * c1;
    R cp = (R)st0[u * 9 + cs], sp = (R)st1[u * 9 + cs];
    R cc = cp * c1 - sp * s1, sc = cp * s1 + sp * c1;
#pragma unroll
    for (int i = 0; i < C; ++i) {
      const R cr = i == 0 ? cp : cc, ci = i == 0 ? sp : sc;
      gr[i] += x[0] * cr;    gi[i] += x[1] * cr;
      kr[0][i] += x[2] * cr; ki[0][i] += x[3] * cr;
      kr[1][i] += x[4] * ci; ki[1][i] += x[5] * ci;
      kr[2][i] += x[6] * ci; ki[2][i] += x[7] * ci;
      kr[3][i] += x[8] * cr; ki[3][i] += x[9] * cr;
      if (i > 0) {
        const R cn = tc * cc - cp, sn = tc * sc - sp;
        cp = cc; sp = sc;
        cc = cn; sc = sn;
      }
    }
  }
}

// 256 threads = 32 point pairs x 8 order chunks of C orders.  Node values are
// evaluated once per (point pair, node) and shared through LDS in batches of 8.
// The sums over the 48 nodes of the singular sinh core run in FP64; those over
// the smooth [xi0, pi] tail in RT (FP64, or FP32 in mixed precision, where the
// far blocks are FP32 already).
template <int C, typename RT = double>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(C <= 9 ? 2 : 1))) void k_near_modal(NearArgs a) {
  const int2 bt = a.batches[blockIdx.x];
  const NearPairInfo pr = a.pairs[bt.x];
  const int slot = threadIdx.x >> 3, cs = threadIdx.x & 7;
  const int q = bt.y + slot;
  const int cnt = a.cell_cnt[pr.kind];
  const bool valid = q < cnt;
  const int qi = a.cell_off[pr.kind] + (valid ? q : 0);
  const double s = a.cell_s[qi], sp = a.cell_sp[qi];
  double rp, zp, rq, zq;
  cell_geom(a.nodes, pr.e, s, rp, zp);
  cell_geom(a.snodes, pr.f, sp, rq, zq);
  const double trp = a.trho[pr.e], tzp = a.tz[pr.e], trq = a.strho[pr.f], tzq = a.stz[pr.f];
  const double k = a.k;
  const double d2 = (rp - rq) * (rp - rq) + (zp - zq) * (zp - zq);
  const double rr4 = 4.0 * rp * rq;
  const bool axis0 = rr4 <= 1e-30;
  const double d = sqrt(d2 > 1e-300 ? d2 : 1e-300);
  const double aa = sqrt(rr4 > 1e-300 ? rr4 : 1e-300);
  const double s0 = axis0 ? 1.0 : fmin(0.25, 20.0 * d / aa);
  const int nnodes = 48 + pr.n_tail;
  const int n0 = cs * C;
  // The point pair's geometry goes to LDS (one copy per slot) and the node
  // values below read it from there: held in registers across the node loop
  // it would sit beside the C x 10 accumulators and push them into scratch.
  __shared__ double pg[32][15];
  if (cs == 0) {
    double* g = pg[slot];
    g[0] = asinh((aa / d) * s0);          // vmax
    g[1] = d;
    g[2] = aa;
    g[3] = d / aa;
    g[4] = 2.0 * asin(fmin(s0, 1.0));      // xi0
    g[5] = GH_PI - g[4];                   // span
    g[6] = rp;
    g[7] = rq;
    g[8] = zp - zq;                        // Rz
    g[9] = trp;
    g[10] = tzp;
    g[11] = trq;
    g[12] = tzq;
    g[13] = axis0 ? 1.0 : 0.0;
    g[14] = d2;
  }
  __syncthreads();

  // per point-pair slot: 8 nodes x 12 values (slot stride 97 doubles, odd, so
  // the 8 slots of a wavefront fall on different banks), and 8 nodes x 8 chunk
  // starts e^{j n0 xi} in separate re / im planes at [slot * 72 + node * 9 +
  // chunk]: reads (chunk = lane's chunk, one node) and writes (node = lane's
  // node, one chunk) both cover distinct banks
  __shared__ double nv[32][97];
  __shared__ double st[2][32 * 72];

  double gR[C], gI[C], kR[4][C], kI[4][C];
#pragma unroll
  for (int i = 0; i < C; ++i) {
    gR[i] = 0.0; gI[i] = 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u) { kR[u][i] = 0.0; kI[u][i] = 0.0; }
  }
  // Node values from (w, R, sin xi, cos xi): the kernel values in T, FP64
  // over the sinh core and RT over the smooth tail (in FP32 the distance, the
  // phase kR and the reciprocal on the hardware sqrt, sin / cos and rcp, as
  // in the far fill's samples, at the level of the FP32 tail sums); xi,
  // cos xi, sin xi and the chunk starts always in FP64.
  auto emit = [&](auto wq, auto Rd, double sxd, double cxd, double* out) {
    using T = decltype(wq);
    const T sx = (T)sxd, cx = (T)cxd;
    const double* g = pg[slot];
    const T rp = (T)g[6], rq = (T)g[7], Rz = (T)g[8], trp = (T)g[9], tzp = (T)g[10], trq = (T)g[11], tzq = (T)g[12];
    const bool axis = g[13] != 0.0;
    const T tiny = sizeof(T) == 8 ? (T)1e-300 : (T)1e-30f;
    Rd = Rd > tiny ? Rd : tiny;
    T sk, ck;
    sincos_s((T)k * Rd, &sk, &ck);
    const T inv = rcp_s(Rd);
    // complex k = k + j ki: e^{-jkR} carries the decay e^{ki R}, and
    // (1 + jkR) = (1 - ki R) + j k R
    const T dec = a.ki != 0.0 ? nexp((T)a.ki * Rd) : (T)1, u = (T)1 - (T)a.ki * Rd;
    // EFIE: 2 g w  (axis pairs use the closed form instead)
    const T fpi = (T)(4.0 * GH_PI);
    const T ge = axis ? (T)0 : (T)2 * wq * inv / fpi * dec;
    out[0] = ge * ck;
    out[1] = -ge * sk;
    // brackets: 2 w p(R) b_uv(xi); a zero weight (the empty tail of an axis
    // pair) stays zero
    const T pre = wq == (T)0 ? (T)0 : (T)2 * wq * inv * inv * inv / fpi * dec;
    const T kr = (T)k * Rd;
    const T pr_ = pre * (u * ck + kr * sk), pi_ = pre * (kr * ck - u * sk);
    const T Rx = rp - rq * cx, Ry = rq * sx;
    const T Wt_t = trp * trq * cx + tzp * tzq, Wt_f = trp * sx, Wf_t = -trq * sx;
    T btt, btf, bft, bff;
    if (a.ibc) {
      const T D = rp * cx - rq;
      const T Wt_n = -trp * tzq * cx + tzp * trq, Wf_n = tzq * sx;
      const T R_t = trq * D + tzq * Rz, R_f = rp * sx, R_n = -tzq * D + trq * Rz;
      btt = Wt_n * R_t - Wt_t * R_n;
      btf = Wt_n * R_f - Wt_f * R_n;
      bft = Wf_n * R_t - Wf_t * R_n;
      bff = Wf_n * R_f - cx * R_n;
    } else {
      const T WtR = trp * Rx + tzp * Rz, nR = -tzp * Rx + trp * Rz;
      const T n_t = -tzp * trq * cx + trp * tzq, n_f = -tzp * sx;
      btt = -(WtR * n_t - Wt_t * nR);
      btf = -(WtR * n_f - Wt_f * nR);
      bft = -(Ry * n_t - Wf_t * nR);
      bff = -(Ry * n_f - cx * nR);
    }
    out[2] = btt * pr_; out[3] = btt * pi_;
    out[4] = btf * pr_; out[5] = btf * pi_;
    out[6] = bft * pr_; out[7] = bft * pi_;
    out[8] = bff * pr_; out[9] = bff * pi_;
    out[10] = cxd; out[11] = sxd;
    // chunk starts e^{j c C xi}, c = 0..7, from e^{j C xi} = (e^{j xi})^C, in
    // FP64 (a modulus error of e^{j xi} grows with the power)
    double cC = 1.0, sC = 0.0;
    {
      double br = cxd, bi = sxd;
#pragma unroll
      for (int bit = 1; bit <= C; bit <<= 1) {
        if (C & bit) {
          const double t = cC * br - sC * bi;
          sC = cC * bi + sC * br;
          cC = t;
        }
        const double t = br * br - bi * bi;
        bi = 2.0 * br * bi;
        br = t;
      }
    }
    double zr = 1.0, zi = 0.0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      st[0][slot * 72 + cs * 9 + c] = zr;
      st[1][slot * 72 + cs * 9 + c] = zi;
      const double t = zr * cC - zi * sC;
      zi = zr * sC + zi * cC;
      zr = t;
    }
  };
  auto node_values = [&](int v0) {
    // ── node value for node v0 + cs of this point pair ──
    const int vn = v0 + cs;
    double* out = &nv[slot][12 * cs];
    if (valid && vn < nnodes) {
      const double* g = pg[slot];
      const double vmax = g[0], d = g[1], aa = g[2], doa = g[3], xi0 = g[4], span = g[5];
      // xi enters through s = sin(xi/2) only: cos xi = 1 - 2 s^2 and
      // sin xi = 2 s cos(xi/2), with cos(xi/2) = sqrt(1 - s^2) on the core
      // (xi = 2 asin s, s <= 1), so neither asin nor a sincos of xi is needed
      if (vn < 48) {
        const double vv = a.core_u[vn] * vmax, wv = a.core_w[vn] * vmax;
        const double em = expm1(vv), ev = em + 1.0;
        const double shv = 0.5 * (em + em / ev), chv = 0.5 * (ev + 1.0 / ev);   // sinh, cosh
        double sv = doa * shv;
        sv = sv < 1.0 ? sv : 1.0;
        const double cv = sqrt(fmax(1.0 - sv * sv, 0.0));
        const double wq = wv * 2.0 * doa * chv / fmax(cv, sqrt(1e-15));
        const double Rd = sqrt(d * d + (aa * sv) * (aa * sv));
        emit(wq, Rd, 2.0 * sv * cv, 1.0 - 2.0 * sv * sv, out);
      } else {
        const int ti = pr.tail_off + (vn - 48);
        const double xi = xi0 + a.tail_u[ti] * span;
        const RT wq = (RT)(a.tail_w[ti] * span);
        // sin / cos of xi/2 in FP64 in either case: an angle rounded to
        // FP32 would put an error of n x 6e-8 on the phase of order n
        double sh, ch;
        sincos(0.5 * xi, &sh, &ch);
        const RT dd = (RT)d, ad = (RT)aa, shr = (RT)sh;
        const RT Rd = sqrt_s(dd * dd + (ad * shr) * (ad * shr));
        emit(wq, Rd, 2.0 * sh * ch, 1.0 - 2.0 * sh * sh, out);
      }
    } else {
#pragma unroll
      for (int i = 0; i < 12; ++i) out[i] = 0.0;
      out[10] = 1.0;
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        st[0][slot * 72 + cs * 9 + c] = 0.0;
        st[1][slot * 72 + cs * 9 + c] = 0.0;
      }
    }
  };
  // singular core (nodes 0..47) in FP64
  for (int v0 = 0; v0 < 48; v0 += 8) {
    node_values(v0);
    __syncthreads();
    if (valid) near_acc8<C, double>(nv[slot], &st[0][slot * 72], &st[1][slot * 72], cs, gR, gI, kR, kI);
    __syncthreads();
  }
  // smooth tail in RT
  {
    RT tR[C], tI[C], uR[4][C], uI[4][C];
#pragma unroll
    for (int i = 0; i < C; ++i) {
      tR[i] = (RT)gR[i]; tI[i] = (RT)gI[i];
#pragma unroll
      for (int u = 0; u < 4; ++u) { uR[u][i] = (RT)kR[u][i]; uI[u][i] = (RT)kI[u][i]; }
    }
    for (int v0 = 48; v0 < nnodes; v0 += 8) {
      node_values(v0);
      __syncthreads();
      if (valid) near_acc8<C, RT>(nv[slot], &st[0][slot * 72], &st[1][slot * 72], cs, tR, tI, uR, uI);
      __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < C; ++i) {
      gR[i] = (double)tR[i]; gI[i] = (double)tI[i];
#pragma unroll
      for (int u = 0; u < 4; ++u) { kR[u][i] = (double)uR[u][i]; kI[u][i] = (double)uI[u][i]; }
    }
  }
  if (!valid) return;
  const bool axis = pg[slot][13] != 0.0;
  const double d2g = pg[slot][14];
  double2* out = a.NP + ((size_t)(pr.pp_off + q) * a.nord) * 5;
#pragma unroll
  for (int i = 0; i < C; ++i) {
    const int n = n0 + i;
    if (n >= a.nord) break;
    double2 G = make_double2(gR[i], gI[i]);
    if (axis) {
      // closed form: only order 0, 2 pi g(d)
      if (n == 0) {
        const double R0 = sqrt(d2g);
        const double Rc = R0 > 1e-300 ? R0 : 1e-300;
        double sk, ck;
        sincos(k * R0, &sk, &ck);
        const double dec = a.ki != 0.0 ? exp(a.ki * R0) : 1.0;
        G = make_double2(2.0 * GH_PI * ck / (4.0 * GH_PI * Rc) * dec, -2.0 * GH_PI * sk / (4.0 * GH_PI * Rc) * dec);
      } else {
        G = make_double2(0.0, 0.0);
      }
    }
    double2* o = out + (size_t)n * 5;
    o[0] = G;
    o[1] = make_double2(kR[0][i], kI[0][i]);
    o[2] = make_double2(kI[1][i], -kR[1][i]);  // -j * sum
    o[3] = make_double2(kI[2][i], -kR[2][i]);
    o[4] = make_double2(kR[3][i], kI[3][i]);
  }
}

// ─────────────────────────────────────────────────────────────────────────────
// near element pairs, stage 2: 2x2 nodal blocks per mode (FP64)
// ─────────────────────────────────────────────────────────────────────────────

struct NearBlockArgs {
  NearArgs na;
  const double* gw;      // Gauss weights x L per point [P] (mass term)
  const double* grho;    // [P]
  const double* gT0;
  const double* gT1;
  double ce_re, ce_im;   // EFIE coefficient
  double cM;             // MFIE coefficient (brackets of NearArgs::ibc == 0)
  const double2* zs;     // IBC (NearArgs::ibc == 1): per-element surface impedance
  double cI;             // IBC coefficient of the brackets x Z_s(source element)
  double cMass;          // IBC coefficient of (1/2) Z_s mass (cI for the IBC-EFIE, 0 for a bare rotated-PV operator)
  double2* NB;           // [3 Ne][16][2M+1]
  double2* NB2;          // bracket and mass part into its own blocks (same layout), nullptr: summed into NB
};

// one work-group per near pair slot; thread = (|m| lane, cell group): with
// mlanes >= M + 1 lanes per group, blockDim / mlanes groups share the cells
// (small M would otherwise leave most threads idle) and are summed in LDS
__global__ __launch_bounds__(256) void k_near_blocks(NearBlockArgs b, int mlanes) {
  const NearArgs& a = b.na;
  const int slot = blockIdx.x;
  const NearPairInfo pr = a.pairs[slot];
  if (pr.e < 0) return;
  const int G = (int)blockDim.x / mlanes;
  const int am = (int)threadIdx.x % mlanes, grp = (int)threadIdx.x / mlanes;
  const bool act = am <= a.M && grp < G;
  const int e = pr.e, f = pr.f;
  const double trp = a.trho[e], tzp = a.tz[e], trq = a.strho[f], tzq = a.stz[f];
  const double Le = a.L[e], Lf = a.sL[f];
  const double dre = a.nodes[2 * e + 2] - a.nodes[2 * e];
  const double drf = a.snodes[2 * f + 2] - a.snodes[2 * f];
  const double k = a.k;
  // 1/k^2: real, or (lossy medium) complex and applied to ksc below
  double ik2 = 1.0 / (k * k), i2r = ik2, i2i = 0.0;
  if (a.ki != 0.0) {
    const double n2 = k * k + a.ki * a.ki;
    i2r = (k * k - a.ki * a.ki) / (n2 * n2);
    i2i = -2.0 * k * a.ki / (n2 * n2);
    ik2 = 1.0;
  }
  const int cnt = a.cell_cnt[pr.kind], off = a.cell_off[pr.kind];
  double2 E[16], K[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) { E[i] = make_double2(0, 0); K[i] = make_double2(0, 0); }
  const int om1 = am > 0 ? am - 1 : 1;
  for (int q = act ? grp : cnt; q < cnt; q += G) {
    const double s = a.cell_s[off + q], sp = a.cell_sp[off + q];
    const double W = a.cell_w[off + q] * Le * Lf;
    double rp, zp, rq, zq;
    cell_geom(a.nodes, e, s, rp, zp);
    cell_geom(a.snodes, f, sp, rq, zq);
    const double Tp[2] = {1.0 - s, s}, Tq[2] = {1.0 - sp, sp};
    const double Dp[2] = {(dre * (1.0 - s) - rp) / Le, (dre * s + rp) / Le};
    const double Dq[2] = {(drf * (1.0 - sp) - rq) / Lf, (drf * sp + rq) / Lf};
    const double rr = rp * rq * W;
    const double2* v = a.NP + (size_t)(pr.pp_off + q) * a.nord * 5;
    const double2 G = v[am * 5], Gm = v[om1 * 5], Gp = v[(am + 1) * 5];
    const double2 Gc = make_double2(0.5 * (Gm.x + Gp.x), 0.5 * (Gm.y + Gp.y));
    const double2 Gs = make_double2(0.5 * (Gm.y - Gp.y), -0.5 * (Gm.x - Gp.x));
    const double2 ktt = make_double2(rr * (trp * trq * Gc.x + tzp * tzq * G.x),
                                     rr * (trp * trq * Gc.y + tzp * tzq * G.y));
    const double2 ktf = make_double2(rr * trp * Gs.x, rr * trp * Gs.y);
    const double2 kft = make_double2(-rr * trq * Gs.x, -rr * trq * Gs.y);
    const double2 kff = make_double2(rr * Gc.x, rr * Gc.y);
    double2 ksc = make_double2(W * G.x, W * G.y);
    if (a.ki != 0.0) ksc = make_double2(i2r * ksc.x - i2i * ksc.y, i2r * ksc.y + i2i * ksc.x);
    const double m = (double)am;
#pragma unroll
    for (int ai = 0; ai < 2; ++ai)
#pragma unroll
      for (int bi = 0; bi < 2; ++bi) {
        const int ab = 2 * ai + bi;
        const double TT = Tp[ai] * Tq[bi];
        const double DD = Dp[ai] * Dq[bi] * ik2, DT = Dp[ai] * Tq[bi] * m * ik2;
        const double TD = Tp[ai] * Dq[bi] * m * ik2, SS = TT * m * m * ik2;
        E[0 + ab].x += TT * ktt.x - DD * ksc.x;   E[0 + ab].y += TT * ktt.y - DD * ksc.y;
        E[4 + ab].x += TT * ktf.x + DT * ksc.y;   E[4 + ab].y += TT * ktf.y - DT * ksc.x;
        E[8 + ab].x += TT * kft.x - TD * ksc.y;   E[8 + ab].y += TT * kft.y + TD * ksc.x;
        E[12 + ab].x += TT * kff.x - SS * ksc.x;  E[12 + ab].y += TT * kff.y - SS * ksc.y;
#pragma unroll
        for (int uv = 0; uv < 4; ++uv) {
          const double2 Kv = v[am * 5 + 1 + uv];
          const double wgt = 2.0 * GH_PI * TT * rr;
          K[4 * uv + ab].x += wgt * Kv.x;
          K[4 * uv + ab].y += wgt * Kv.y;
        }
      }
  }
  // sum the cell groups (group 0 keeps the totals)
  __shared__ double2 red[256][17];
#define GROUP_SUM(V)                                          \
  {                                                           \
    _Pragma("unroll") for (int i = 0; i < 16; ++i) red[threadIdx.x][i] = V[i]; \
    __syncthreads();                                          \
    if (grp == 0 && act)                                      \
      for (int g = 1; g < G; ++g)                             \
        _Pragma("unroll") for (int i = 0; i < 16; ++i) {      \
          const double2 w = red[g * mlanes + am][i];          \
          V[i].x += w.x;                                      \
          V[i].y += w.y;                                      \
        }                                                     \
    __syncthreads();                                          \
  }
  GROUP_SUM(E)
  GROUP_SUM(K)
#undef GROUP_SUM
  if (grp != 0 || !act) return;
  // mass term 2 pi int rho T_a T_b dt on the element itself (diagonal pairs),
  // a single integral taken over the test geometry
  double Mab[4] = {0, 0, 0, 0};
  if (e == f) {
    for (int gq = 0; gq < 4; ++gq) {
      const int ip = 4 * e + gq;
      const double wr = b.gw[ip] * b.grho[ip];
      const double T[2] = {b.gT0[ip], b.gT1[ip]};
      for (int ai = 0; ai < 2; ++ai)
        for (int bi = 0; bi < 2; ++bi) Mab[2 * ai + bi] += 2.0 * GH_PI * wr * T[ai] * T[bi];
    }
  }
  const int nmodes = 2 * a.M + 1;
  double2* out = b.NB + (size_t)slot * 16 * nmodes;
  // bracket coefficient kc and mass coefficient mc: MFIE Z_M = (1/2) M - K
  // (bor_solver.py:428-435) with c_mfie; IBC (1/2) M_Zs + K'_Zs with c_ibc,
  // Z_s of the source element on the brackets and of the element itself on
  // the mass (bor_solver.py:519-552)
  double2 kc = make_double2(-b.cM, 0.0), mc = make_double2(0.5 * b.cM, 0.0);
  if (a.ibc) {
    kc = make_double2(b.cI * b.zs[f].x, b.cI * b.zs[f].y);
    mc = make_double2(0.5 * b.cMass * b.zs[e].x, 0.5 * b.cMass * b.zs[e].y);
  }
  // NB2 set: the EFIE part to NB, the bracket and mass part to NB2
  double2* out2 = b.NB2 ? b.NB2 + (size_t)slot * 16 * nmodes : nullptr;
  for (int blk = 0; blk < 4; ++blk)
    for (int ab = 0; ab < 4; ++ab) {
      const double2 Ev = E[4 * blk + ab], Kv = K[4 * blk + ab];
      const double xr = b.ce_re * Ev.x - b.ce_im * Ev.y, xi = b.ce_re * Ev.y + b.ce_im * Ev.x;
      const double kr = kc.x * Kv.x - kc.y * Kv.y, ki = kc.x * Kv.y + kc.y * Kv.x;
      const bool odd = (blk == 1 || blk == 2);
      const double mr = odd ? 0.0 : mc.x * Mab[ab], mi = odd ? 0.0 : mc.y * Mab[ab];
      const double sg = odd ? -1.0 : 1.0;   // mode -m: odd blocks change sign, the mass does not
      double2* o = out + (size_t)(4 * blk + ab) * nmodes + a.M;
      if (out2) {
        double2* o2 = out2 + (size_t)(4 * blk + ab) * nmodes + a.M;
        o[am] = make_double2(xr, xi);
        o2[am] = make_double2(kr + mr, ki + mi);
        if (am > 0) {
          o[-am] = make_double2(sg * xr, sg * xi);
          o2[-am] = make_double2(sg * kr + mr, sg * ki + mi);
        }
      } else {
        o[am] = make_double2(xr + kr + mr, xi + ki + mi);
        if (am > 0) o[-am] = make_double2(sg * (xr + kr) + mr, sg * (xi + ki) + mi);
      }
    }
}

// ─────────────────────────────────────────────────────────────────────────────
// assembly: Zfar + near band -> masked row-major mode systems
// ─────────────────────────────────────────────────────────────────────────────

struct AsmArgs {
  const void* Zfar;          // [2Nn][2Nn][zmodes] complex<Rf>, mode m at m + zoff
  const double2* NB;         // [3 Ne][16][nmodes]
  const SysInfo* sys;        // systems of this problem, index = m + M
  int Nn, Ne, M;
  int mlo;                   // first mode index assembled (M with pairing: modes -m are not stored)
  int zmodes, zoff;
  void* A;                   // augmented systems (complex<Ra>), base pointer
  double2* A64;              // optional FP64 copy of the n x n block (nullptr if unused)
};

template <typename Rf, typename Ra>
__global__ __launch_bounds__(256) void k_assemble(AsmArgs a) {
  typedef typename CT<Rf>::c Cf;
  typedef typename CT<Ra>::c Ca;
  __shared__ double2 tile[32][33];
  const int i_row = blockIdx.x;            // node-level row in [0, 2Nn)
  const int j0 = blockIdx.y * 32;          // node-level column chunk
  const int m0 = a.mlo + blockIdx.z * 32;  // mode chunk (index into [0, nmodes))
  const int nmodes = 2 * a.M + 1;
  const size_t ncol = 2 * (size_t)a.Nn;
  const Cf* Z = (const Cf*)a.Zfar;
  // load [32 cols][32 modes], modes fastest
  for (int t = threadIdx.x; t < 32 * 32; t += 256) {
    const int jj = t >> 5, mm = t & 31;
    const int j = j0 + jj, mi = m0 + mm;
    double2 v = make_double2(0, 0);
    if (j < (int)ncol && mi < nmodes) {
      const Cf z = Z[((size_t)i_row * ncol + j) * a.zmodes + (mi - a.M + a.zoff)];
      v = make_double2((double)z.x, (double)z.y);
    }
    tile[jj][mm] = v;
  }
  __syncthreads();
  const int bi = i_row >= a.Nn ? 1 : 0;
  const int ii = i_row - bi * a.Nn;
  for (int t = threadIdx.x; t < 32 * 32; t += 256) {
    const int mm = t >> 5, jj = t & 31;
    const int j = j0 + jj, mi = m0 + mm;
    if (j >= (int)ncol || mi >= nmodes) continue;
    const SysInfo s = a.sys[mi];
    const int nf = s.n - s.nt;
    int r;
    if (bi == 0) { r = ii - s.t0; if (r < 0 || r >= s.nt) continue; }
    else { r = ii - s.f0; if (r < 0 || r >= nf) continue; r += s.nt; }
    const int bj = j >= a.Nn ? 1 : 0;
    const int jj2 = j - bj * a.Nn;
    int c;
    if (bj == 0) { c = jj2 - s.t0; if (c < 0 || c >= s.nt) continue; }
    else { c = jj2 - s.f0; if (c < 0 || c >= nf) continue; c += s.nt; }
    double2 v = tile[jj][mm];
    if (abs(ii - jj2) <= 2) {
      const int blk = 2 * bi + bj;
#pragma unroll
      for (int ai = 0; ai < 2; ++ai)
#pragma unroll
        for (int bb = 0; bb < 2; ++bb) {
          const int e = ii - ai, f = jj2 - bb;
          if (e < 0 || f < 0 || e >= a.Ne || f >= a.Ne || abs(e - f) > 1) continue;
          const int slot = 3 * e + (f - e + 1);
          const double2 w = a.NB[((size_t)slot * 16 + 4 * blk + 2 * ai + bb) * nmodes + mi];
          v.x += w.x; v.y += w.y;
        }
    }
    Ca* A = (Ca*)a.A + s.a_off;
    Ca o; o.x = (Ra)v.x; o.y = (Ra)v.y;
    A[(size_t)r * s.ld + c] = o;
    if (a.A64) a.A64[s.a64_off + (size_t)r * s.n + c] = v;
  }
}

// Near-band part of one node-level entry (row node ii, column node jj, block
// blk = 2 (row half) + column half) of mode index mi: the element pairs
// (e, f) with e in {ii - 1, ii}, f in {jj - 1, jj} and |e - f| <= 1.
__device__ inline double2 near_entry(const double2* NB, int ii, int jj, int blk, int mi, int Ne, int nmodes) {
  double2 v = make_double2(0.0, 0.0);
  if (abs(ii - jj) > 2) return v;
#pragma unroll
  for (int ai = 0; ai < 2; ++ai)
#pragma unroll
    for (int bb = 0; bb < 2; ++bb) {
      const int e = ii - ai, f = jj - bb;
      if (e < 0 || f < 0 || e >= Ne || f >= Ne || abs(e - f) > 1) continue;
      const int slot = 3 * e + (f - e + 1);
      const double2 w = NB[((size_t)slot * 16 + 4 * blk + 2 * ai + bb) * nmodes + mi];
      v.x += w.x;
      v.y += w.y;
    }
  return v;
}

// PMCHWT assembly of a homogeneous penetrable body (bor_solver.py:1047-1056):
// from the exterior and interior EFIE operators T and rotated-PV bracket
// operators B (far blocks + near band, node based), the masked system
//   [[Te + Ti, -eta0 (Pe + Pi)], [eta0 (Pe + Pi), Te + (eta0/eta_i)^2 Ti]]
// for the unknowns (J, M / eta0), where P takes B's columns remapped,
// P = [[-Btf, Btt], [-Bff, Bft]] (bor_solver.py:554-575): column j < Nn of P
// is -(column j + Nn of B), column j >= Nn is column j - Nn.  FP64 only.
struct AsmPenArgs {
  const double2* Te;         // far blocks [2Nn][2Nn][zmodes] of the four operators
  const double2* Be;
  const double2* Ti;
  const double2* Bi;
  const double2* NTe;        // near blocks [3 Ne][16][nmodes]
  const double2* NBe;
  const double2* NTi;
  const double2* NBi;
  const SysInfo* sys;        // systems of this problem, index = m + M (pen = single-surface size)
  int Nn, Ne, M, mlo, zmodes, zoff;
  double eta0;
  double r2_re, r2_im;       // (eta0 / eta_i)^2
  double2* A;                // augmented systems
  double2* A64;              // optional FP64 copy of the n x n block
};

__global__ __launch_bounds__(256) void k_assemble_pen(AsmPenArgs a) {
  __shared__ double2 tile[4][32][33];   // Te, Ti at column j; Be, Bi at the partner column
  const int i_row = blockIdx.x;
  const int j0 = blockIdx.y * 32;
  const int m0 = a.mlo + blockIdx.z * 32;
  const int nmodes = 2 * a.M + 1;
  const int Nn = a.Nn;
  const size_t ncol = 2 * (size_t)Nn;
  for (int t = threadIdx.x; t < 4 * 32 * 32; t += 256) {
    const int op = t >> 10, jj = (t >> 5) & 31, mm = t & 31;
    const int j = j0 + jj, mi = m0 + mm;
    double2 v = make_double2(0, 0);
    if (j < (int)ncol && mi < nmodes) {
      const int jc = op < 2 ? j : (j < Nn ? j + Nn : j - Nn);
      const double2* Z = op == 0 ? a.Te : op == 1 ? a.Ti : op == 2 ? a.Be : a.Bi;
      v = Z[((size_t)i_row * ncol + jc) * a.zmodes + (mi - a.M + a.zoff)];
    }
    tile[op][jj][mm] = v;
  }
  __syncthreads();
  const int bi = i_row >= Nn ? 1 : 0;
  const int ii = i_row - bi * Nn;
  for (int t = threadIdx.x; t < 32 * 32; t += 256) {
    const int mm = t >> 5, jj = t & 31;
    const int j = j0 + jj, mi = m0 + mm;
    if (j >= (int)ncol || mi >= nmodes) continue;
    const SysInfo s = a.sys[mi];
    const int n1 = s.pen, nf = n1 - s.nt;
    int r;
    if (bi == 0) { r = ii - s.t0; if (r < 0 || r >= s.nt) continue; }
    else { r = ii - s.f0; if (r < 0 || r >= nf) continue; r += s.nt; }
    const int bj = j >= Nn ? 1 : 0;
    const int jn = j - bj * Nn;
    int c;
    if (bj == 0) { c = jn - s.t0; if (c < 0 || c >= s.nt) continue; }
    else { c = jn - s.f0; if (c < 0 || c >= nf) continue; c += s.nt; }
    const int blk = 2 * bi + bj, blkp = 2 * bi + (1 - bj);
    double2 te = tile[0][jj][mm], ti = tile[1][jj][mm], be = tile[2][jj][mm], bq = tile[3][jj][mm];
    const double2 nte = near_entry(a.NTe, ii, jn, blk, mi, a.Ne, nmodes);
    const double2 nti = near_entry(a.NTi, ii, jn, blk, mi, a.Ne, nmodes);
    const double2 nbe = near_entry(a.NBe, ii, jn, blkp, mi, a.Ne, nmodes);
    const double2 nbi = near_entry(a.NBi, ii, jn, blkp, mi, a.Ne, nmodes);
    te.x += nte.x; te.y += nte.y;
    ti.x += nti.x; ti.y += nti.y;
    // eta0 (Pe + Pi) at column j
    const double ps = (bj == 0 ? -a.eta0 : a.eta0);
    const double2 pv = make_double2(ps * (be.x + nbe.x + bq.x + nbi.x), ps * (be.y + nbe.y + bq.y + nbi.y));
    const double2 v00 = make_double2(te.x + ti.x, te.y + ti.y);
    const double2 v11 = make_double2(te.x + a.r2_re * ti.x - a.r2_im * ti.y, te.y + a.r2_re * ti.y + a.r2_im * ti.x);
    double2* A = a.A + s.a_off;
    A[(size_t)r * s.ld + c] = v00;
    A[(size_t)r * s.ld + n1 + c] = make_double2(-pv.x, -pv.y);
    A[(size_t)(n1 + r) * s.ld + c] = pv;
    A[(size_t)(n1 + r) * s.ld + n1 + c] = v11;
    if (a.A64) {
      double2* B = a.A64 + s.a64_off;
      B[(size_t)r * s.n + c] = v00;
      B[(size_t)r * s.n + n1 + c] = make_double2(-pv.x, -pv.y);
      B[(size_t)(n1 + r) * s.n + c] = pv;
      B[(size_t)(n1 + r) * s.n + n1 + c] = v11;
    }
  }
}

// ─────────────────────────────────────────────────────────────────────────────
// multi-region systems (cross-surface operators, block sums, junction map)
// ─────────────────────────────────────────────────────────────────────────────

// Near pairs of a cross-surface operator added into its far blocks: one
// work-group per pair of one parity class (e mod 2, f mod 2), whose 2 x 2
// node footprints are disjoint, so the adds need no atomics and run in a
// fixed order.  NB / NB2 [pair][16][nmodes] (EFIE / brackets); Z / Z2
// [2Np][2Nq][zmodes], modes from mlo.
__global__ __launch_bounds__(256) void k_cross_near_add(const double2* NB, const double2* NB2,
                                                        const NearPairInfo* pairs, const int* plist, int M,
                                                        int Np, int Nq, int zmodes, int zoff, int mlo,
                                                        double2* Z, double2* Z2) {
  const int pi = plist[blockIdx.x];
  const NearPairInfo pr = pairs[pi];
  const int nmodes = 2 * M + 1, nm = nmodes - mlo;
  const size_t ncol = 2 * (size_t)Nq;
  for (int t = threadIdx.x; t < 16 * nm; t += blockDim.x) {
    const int q = t / nm, mi = mlo + t % nm;
    const int blk = q >> 2, ai = (q >> 1) & 1, bi = q & 1;
    const int row = ((blk >> 1) ? Np : 0) + pr.e + ai, col = ((blk & 1) ? Nq : 0) + pr.f + bi;
    const size_t zi = ((size_t)row * ncol + col) * zmodes + zoff + (mi - M);
    const size_t ni = ((size_t)pi * 16 + q) * nmodes + mi;
    double2 v = Z[zi], w = NB[ni];
    v.x += w.x; v.y += w.y;
    Z[zi] = v;
    if (Z2) {
      v = Z2[zi]; w = NB2[ni];
      v.x += w.x; v.y += w.y;
      Z2[zi] = v;
    }
  }
}

// One operator block added into the full (node-based, unreduced) system of
// every mode: Afull[m][row_off + i][col_off + j] += coef x op(i, j, m), op
// = T, or P = B with columns remapped (column j < Nq is -(column j + Nq),
// column j >= Nq is column j - Nq; bor_solver.py:554-575).  The operator's
// far blocks hold modes 0..M only when the handle pairs modes: mode -m is
// then S op_m S (S = +1 on J_t rows / columns, -1 on J_phi), as its near
// blocks (stored for every mode) already are.  Same-surface operators add
// their near band from NB (|e - f| <= 1 slots); cross-surface ones have it
// in their far blocks.  Modes sit on the LDS tile's fast axis for the far
// block reads, the columns for the writes.
struct RegionBlockArgs {
  const double2* Z;          // far blocks [2Np][2Nq][zmodes]
  const double2* NB;         // near band [3 Ne][16][nmodes] of a same-surface operator, or nullptr
  int Np, Nq, Ne, M, zmodes, zoff, paired, remap;
  double cre, cim;
  int row_off, col_off, n_full;
  double2* A;                // [nmodes][n_full][n_full]
};

__global__ __launch_bounds__(256) void k_region_block(RegionBlockArgs a) {
  __shared__ double2 tile[32][33];
  const int i = blockIdx.x;
  const int j0 = blockIdx.y * 32, m0 = blockIdx.z * 32;
  const int nmodes = 2 * a.M + 1, ncol = 2 * a.Nq;
  const double si = i < a.Np ? 1.0 : -1.0;
  const int bi = i < a.Np ? 0 : 1, ii = i - bi * a.Np;
  for (int t = threadIdx.x; t < 32 * 32; t += 256) {
    const int jj = t >> 5, mm = t & 31;
    const int j = j0 + jj, mi = m0 + mm;
    double2 v = make_double2(0.0, 0.0);
    if (j < ncol && mi < nmodes) {
      const int m = mi - a.M;
      const int jc = a.remap ? (j < a.Nq ? j + a.Nq : j - a.Nq) : j;
      const int zi = a.paired ? abs(m) : a.zoff + m;
      v = a.Z[((size_t)i * ncol + jc) * a.zmodes + zi];
      if (a.paired && m < 0) {
        const double sg = si * (jc < a.Nq ? 1.0 : -1.0);
        v.x *= sg; v.y *= sg;
      }
      if (a.NB) {
        const int bj = jc < a.Nq ? 0 : 1;
        const double2 w = near_entry(a.NB, ii, jc - bj * a.Nq, 2 * bi + bj, mi, a.Ne, nmodes);
        v.x += w.x; v.y += w.y;
      }
      if (a.remap && j < a.Nq) { v.x = -v.x; v.y = -v.y; }
    }
    tile[jj][mm] = v;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < 32 * 32; t += 256) {
    const int mm = t >> 5, jj = t & 31;
    const int j = j0 + jj, mi = m0 + mm;
    if (j >= ncol || mi >= nmodes) continue;
    const double2 v = tile[jj][mm];
    double2* o = a.A + ((size_t)mi * a.n_full + a.row_off + i) * a.n_full + a.col_off + j;
    double2 r = *o;
    r.x += a.cre * v.x - a.cim * v.y;
    r.y += a.cre * v.y + a.cim * v.x;
    *o = r;
  }
}

// Junction map of one mode, as lists: reduced index r gathers the full
// indices mem[ptr[r] .. ptr[r+1]) with coefficients cf (x_full = Q x_red,
// each full index in at most one list).
struct RegionMap {
  const int* ptr;            // [nmodes][n_full + 1] (only n_red + 1 used per mode)
  const int* mem;            // [nmodes][n_full]
  const double* cf;          // [nmodes][n_full]
  int n_full;
};

// A_red = Q^T A_full Q into the mode's augmented system (and its FP64 copy).
__global__ __launch_bounds__(256) void k_region_reduce(RegionMap q, const double2* Afull, const SysInfo* sys,
                                                       double2* A, double2* A64) {
  const int mi = blockIdx.z;
  const SysInfo s = sys[mi];
  const int r = blockIdx.x;
  const int c = blockIdx.y * blockDim.x + threadIdx.x;
  if (r >= s.n || c >= s.n) return;
  const int* ptr = q.ptr + (size_t)mi * (q.n_full + 1);
  const int* mem = q.mem + (size_t)mi * q.n_full;
  const double* cf = q.cf + (size_t)mi * q.n_full;
  const double2* F = Afull + (size_t)mi * q.n_full * q.n_full;
  double2 v = make_double2(0.0, 0.0);
  for (int u = ptr[r]; u < ptr[r + 1]; ++u)
    for (int w = ptr[c]; w < ptr[c + 1]; ++w) {
      const double2 x = F[(size_t)mem[u] * q.n_full + mem[w]];
      const double g = cf[u] * cf[w];
      v.x += g * x.x;
      v.y += g * x.y;
    }
  A[s.a_off + (size_t)r * s.ld + c] = v;
  if (A64) A64[s.a64_off + (size_t)r * s.n + c] = v;
}

// B_red = Q^T B_full into the augmented columns (and the FP64 copy).
__global__ __launch_bounds__(256) void k_region_rhs(RegionMap q, const double2* Bfull, const SysInfo* sys,
                                                    double2* A, double2* B64) {
  const int mi = blockIdx.z;
  const SysInfo s = sys[mi];
  const int r = blockIdx.x;
  const int c = blockIdx.y * blockDim.x + threadIdx.x;
  if (r >= s.n || c >= s.nrhs) return;
  const int* ptr = q.ptr + (size_t)mi * (q.n_full + 1);
  const int* mem = q.mem + (size_t)mi * q.n_full;
  const double* cf = q.cf + (size_t)mi * q.n_full;
  const double2* F = Bfull + (size_t)mi * q.n_full * s.nrhs;
  double2 v = make_double2(0.0, 0.0);
  for (int u = ptr[r]; u < ptr[r + 1]; ++u) {
    const double2 x = F[(size_t)mem[u] * s.nrhs + c];
    v.x += cf[u] * x.x;
    v.y += cf[u] * x.y;
  }
  A[s.a_off + (size_t)r * s.ld + s.n + s.col + c] = v;
  if (B64) B64[s.x_off + (size_t)r * s.ncols + s.col + c] = v;
}

// x_full = Q x_red from the augmented columns, every full index (0 where
// the map drops it), for the far field.
__global__ __launch_bounds__(256) void k_region_expand(RegionMap q, const double2* A, const SysInfo* sys,
                                                       double2* Xfull) {
  const int mi = blockIdx.z;
  const SysInfo s = sys[mi];
  const int r = blockIdx.x;
  const int c = blockIdx.y * blockDim.x + threadIdx.x;
  if (r >= s.n || c >= s.nrhs) return;
  const int* ptr = q.ptr + (size_t)mi * (q.n_full + 1);
  const int* mem = q.mem + (size_t)mi * q.n_full;
  const double* cf = q.cf + (size_t)mi * q.n_full;
  const double2 x = A[s.a_off + (size_t)r * s.ld + s.n + s.col + c];
  for (int u = ptr[r]; u < ptr[r + 1]; ++u)
    Xfull[((size_t)mi * q.n_full + mem[u]) * s.nrhs + c] = make_double2(cf[u] * x.x, cf[u] * x.y);
}

// ─────────────────────────────────────────────────────────────────────────────
// compat samplers (bor_stream_kernel.c signatures; host entry points wrap these)
// ─────────────────────────────────────────────────────────────────────────────

__global__ void k_sample_g(int nr, int np_, int nxi, const double* rho_p, const double* z_p,
                           const double* rho_q, const double* z_q, double k,
                           const double* sin2, double2* out) {
  const size_t tot = (size_t)nr * np_ * nxi;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < tot;
       t += (size_t)gridDim.x * blockDim.x) {
    const int l = (int)(t % nxi);
    const size_t ij = t / nxi;
    const int j = (int)(ij % np_), i = (int)(ij / np_);
    const double dr = rho_p[i] - rho_q[j], dz = z_p[i] - z_q[j];
    double R = sqrt(dr * dr + dz * dz + 4.0 * rho_p[i] * rho_q[j] * sin2[l]);
    if (R < 1e-30) R = 1e-30;
    const double amp = 1.0 / (4.0 * GH_PI * R);
    double s, c;
    sincos(k * R, &s, &c);
    out[t] = make_double2(amp * c, -amp * s);
  }
}

template <bool IBC>
__global__ void k_sample_brackets(int nr, int np_, int nxi, const double* rho_p, const double* z_p,
                                  const double* tr_p, const double* tz_p, const double* rho_q,
                                  const double* z_q, const double* tr_q, const double* tz_q, double k,
                                  const double* cx_t, const double* sx_t, double2* o_tt, double2* o_tf,
                                  double2* o_ft, double2* o_ff) {
  const size_t tot = (size_t)nr * np_ * nxi;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < tot;
       t += (size_t)gridDim.x * blockDim.x) {
    const int l = (int)(t % nxi);
    const size_t ij = t / nxi;
    const int j = (int)(ij % np_), i = (int)(ij / np_);
    const double cx = cx_t[l], sx = sx_t[l];
    const double Rx = rho_p[i] - rho_q[j] * cx, Ry = rho_q[j] * sx, Rz = z_p[i] - z_q[j];
    double R = sqrt(Rx * Rx + Ry * Ry + Rz * Rz);
    if (R < 1e-30) R = 1e-30;
    const double pre = 1.0 / (4.0 * GH_PI * R * R * R);
    double s, c;
    sincos(k * R, &s, &c);
    const double kr = k * R;
    const double pr = pre * (c + kr * s), pim = pre * (kr * c - s);
    const double trp = tr_p[i], tzp = tz_p[i], trq = tr_q[j], tzq = tz_q[j];
    double ftt, ftf, fft, fff;
    if (!IBC) {
      const double WtR = trp * Rx + tzp * Rz, nR = -tzp * Rx + trp * Rz;
      const double n_t = -tzp * trq * cx + trp * tzq, n_f = -tzp * sx;
      const double Wt_t = trp * trq * cx + tzp * tzq, Wt_f = trp * sx, Wf_t = -trq * sx;
      ftt = -(WtR * n_t - Wt_t * nR);
      ftf = -(WtR * n_f - Wt_f * nR);
      fft = -(Ry * n_t - Wf_t * nR);
      fff = -(Ry * n_f - cx * nR);
    } else {
      const double Wt_n = -trp * tzq * cx + tzp * trq, Wf_n = tzq * sx;
      const double D = rho_p[i] * cx - rho_q[j];
      const double R_t = trq * D + tzq * Rz, R_f = rho_p[i] * sx, R_n = -tzq * D + trq * Rz;
      const double Wt_t = trp * trq * cx + tzp * tzq, Wt_f = trp * sx, Wf_t = -trq * sx;
      ftt = Wt_n * R_t - Wt_t * R_n;
      ftf = Wt_n * R_f - Wt_f * R_n;
      fft = Wf_n * R_t - Wf_t * R_n;
      fff = Wf_n * R_f - cx * R_n;
    }
    o_tt[t] = make_double2(ftt * pr, ftt * pim);
    o_tf[t] = make_double2(ftf * pr, ftf * pim);
    o_ft[t] = make_double2(fft * pr, fft * pim);
    o_ff[t] = make_double2(fff * pr, fff * pim);
  }
}

__global__ void k_zero(double2* p, size_t n) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
    p[t] = make_double2(0.0, 0.0);
}

}  // namespace ghost
