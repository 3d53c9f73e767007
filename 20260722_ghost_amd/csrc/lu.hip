// Batched blocked complex LU with partial pivoting, triangular solves and the
// complex MFMA GEMM that carries the trailing updates, for a batch of
// row-major systems of uneven size (one per azimuthal mode and frequency).
//
// Replaces numpy.linalg.solve (LAPACK zgesv) at bor_solver.py:753: pivot
// choice by max |Re| + |Im| as LAPACK's izamax, row interchanges applied to
// whole rows, unit-lower L and upper U stored in place.  Right-hand sides are
// appended as extra columns so the trailing updates also perform the forward
// substitution; back substitution runs blocked (diagonal TRSM + GEMM).
#include "bor_common.hpp"

namespace ghost {

struct GemmOp {      // C[M x N] -= A[M x K] B[K x N], complex, row-major
  long long a, b, c; // element offsets from the respective base pointers
  int lda, ldb, ldc;
  int M, N, K;
  int t0[3];         // first block of this op in its launch, per tile shape
                     // (64x64, 64x128, 128x128): launches are flattened over
                     // the tiles of all their ops, so systems of uneven size
                     // cost no idle blocks
};

// Op of flattened block `bid`: the last op whose first block is <= bid (ops
// with no tiles share the next op's start and are skipped).
__device__ __forceinline__ int find_op(const GemmOp* ops, int nops, int shape, int bid) {
  int lo = 0, hi = nops - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ops[mid].t0[shape] <= bid) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// Flattened tile of workgroup b of nb: XCD x = b mod 8 takes the contiguous
// range [x q + min(x, r), ...) (q = nb / 8, r = nb mod 8; the first r XCDs
// take q + 1), so neighbouring tiles, which share A row panels and B column
// panels of one system, run on one L2.  A bijection of [0, nb).
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int q = nb >> 3, r = nb & 7, x = b & 7, i = b >> 3;
  return x * q + min(x, r) + i;
}

struct PanelOp {     // factor rows [k, n) x cols [k, k + w)
  long long a;       // system base (element offset)
  int n, ld, k, w;
  int piv;           // offset of this system's pivot array
  int sys;           // system index (zero-pivot count)
};

struct SwapOp {      // apply pivots k..k+w-1 to cols [0, c_lo) and [c_hi, ncols)
  long long a;
  int ld, k, w, c_lo, c_hi, ncols, piv;
};

struct TrsmOp {      // diagonal block T (w x w) of a factor and the slot of its W = I - T^{-1}
  long long t;       // element offset of T's (0, 0) in the factor buffer
  long long wo;      // element offset of the 64 x 64 W slot
  int ldt, w;
};

// ─────────────────────────────────────────────────────────────────────────────
// complex GEMM update on 16x16x4 MFMA: 64 x 64 tile per 256-thread block
// ─────────────────────────────────────────────────────────────────────────────

// TA / TB: the operand is stored transposed (A as K x M with lda, B as N x K with ldb).
template <typename T, bool TA = false, bool TB = false>
__global__ __launch_bounds__(256) void k_gemm_update(const GemmOp* ops, int nops, const void* Abase,
                                                     const void* Bbase, void* Cbase) {
  typedef typename CT<T>::c C;
  typedef typename CT<T>::v4 V4;
  const GemmOp op = ops[find_op(ops, nops, 0, (int)blockIdx.x)];
  if (op.M <= 0 || op.N <= 0 || op.K <= 0) return;
  const int tn = (op.N + 63) >> 6;
  const int tiles = ((op.M + 63) >> 6) * tn;
  const int tile = (int)blockIdx.x - op.t0[0];
  if (tile < 0 || tile >= tiles) return;
  const int m0 = (tile / tn) * 64, n0 = (tile % tn) * 64;
  const C* A = (const C*)Abase + op.a;
  const C* B = (const C*)Bbase + op.b;
  C* Cm = (C*)Cbase + op.c;

  __shared__ T Ar[16][64 + 1], Ai[16][64 + 1];   // [k][row]
  __shared__ T Br[16][64 + 1], Bi[16][64 + 1];   // [k][col]

  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wy = wv >> 1, wx = wv & 1;
  const int c16 = lane & 15, kk = lane >> 4;
  V4 accR[2][2], accI[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) { accR[i][j] = V4{0, 0, 0, 0}; accI[i][j] = V4{0, 0, 0, 0}; }

  for (int k0 = 0; k0 < op.K; k0 += 16) {
    // A tile: 64 rows x 16 k, thread -> (row, 4 consecutive k)
    {
      const int row = tid >> 2, kq = (tid & 3) * 4;
      const int gr = m0 + row;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int gk = k0 + kq + q;
        C v;
        v.x = 0; v.y = 0;
        if (gr < op.M && gk < op.K) v = TA ? A[(size_t)gk * op.lda + gr] : A[(size_t)gr * op.lda + gk];
        Ar[kq + q][row] = v.x;
        Ai[kq + q][row] = v.y;
      }
    }
    // B tile: 16 k x 64 cols, thread -> (k, 4 consecutive cols)
    {
      const int kr = tid >> 4, cq = (tid & 15) * 4;
      const int gk = k0 + kr;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int gc = n0 + cq + q;
        C v;
        v.x = 0; v.y = 0;
        if (gk < op.K && gc < op.N) v = TB ? B[(size_t)gc * op.ldb + gk] : B[(size_t)gk * op.ldb + gc];
        Br[kr][cq + q] = v.x;
        Bi[kr][cq + q] = v.y;
      }
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const int kx = 4 * ks + kk;
      T ar[2], ai[2], nai[2], br[2], bi[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        ar[i] = Ar[kx][wy * 32 + i * 16 + c16];
        ai[i] = Ai[kx][wy * 32 + i * 16 + c16];
        nai[i] = -ai[i];
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        br[j] = Br[kx][wx * 32 + j * 16 + c16];
        bi[j] = Bi[kx][wx * 32 + j * 16 + c16];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          accR[i][j] = mfma4(ar[i], br[j], accR[i][j]);
          accR[i][j] = mfma4(nai[i], bi[j], accR[i][j]);
          accI[i][j] = mfma4(ar[i], bi[j], accI[i][j]);
          accI[i][j] = mfma4(ai[i], br[j], accI[i][j]);
        }
    }
    __syncthreads();
  }
  // C -= acc
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gr = m0 + wy * 32 + i * 16 + mrow<T>(kk, r);
        const int gc = n0 + wx * 32 + j * 16 + c16;
        if (gr < op.M && gc < op.N) {
          C* p = Cm + (size_t)gr * op.ldc + gc;
          C v = *p;
          v.x -= accR[i][j][r];
          v.y -= accI[i][j][r];
          *p = v;
        }
      }
}

// C[M x 2] -= A[M x K] B[K x 2], FP64 complex: the error probe's two-column
// products.  One wavefront per row of A (coalesced along K), B read through
// the cache, the two dot products reduced across the wave.  Memory-bound: a
// 64 x 64 MFMA tile would spend its time multiplying B's zero padding.
__global__ __launch_bounds__(256) void k_gemm_n2(const GemmOp* ops, const double2* Abase, const double2* Bbase,
                                                 double2* Cbase) {
  const GemmOp op = ops[blockIdx.y];
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= op.M || op.N != 2 || op.K <= 0) return;
  const double2* a = Abase + op.a + (size_t)row * op.lda;
  const double2* b = Bbase + op.b;
  double r0 = 0, i0 = 0, r1 = 0, i1 = 0;
  for (int k = lane; k < op.K; k += 64) {
    const double2 x = a[k], u = b[(size_t)k * op.ldb], v = b[(size_t)k * op.ldb + 1];
    r0 += x.x * u.x - x.y * u.y;
    i0 += x.x * u.y + x.y * u.x;
    r1 += x.x * v.x - x.y * v.y;
    i1 += x.x * v.y + x.y * v.x;
  }
  for (int o = 32; o > 0; o >>= 1) {
    r0 += __shfl_xor(r0, o);
    i0 += __shfl_xor(i0, o);
    r1 += __shfl_xor(r1, o);
    i1 += __shfl_xor(i1, o);
  }
  if (lane == 0) {
    double2* c = Cbase + op.c + (size_t)row * op.ldc;
    c[0].x -= r0; c[0].y -= i0;
    c[1].x -= r1; c[1].y -= i1;
  }
}

// Block tile (32 WM) x (32 WN) complex, 2 x 2 wavefronts of WM x WN MFMA
// tiles each; K in chunks of 16 staged through two LDS buffers with the next
// chunk's global loads in flight while the current one feeds the MFMAs.
template <typename T, int WM, int WN, bool TA = false, bool TB = false, bool PRE = false>
__global__ __launch_bounds__(256) void k_gemm2(const GemmOp* ops, int nops, const void* Abase,
                                               const void* Bbase, void* Cbase) {
  typedef typename CT<T>::c C;
  typedef typename CT<T>::v4 V4;
  constexpr int BM = 32 * WM, BN = 32 * WN, KC = 16;
  constexpr int AP = BM * KC / 256, BP = BN * KC / 256;   // complex elements per thread
  constexpr int SH = (BM == 64 && BN == 64) ? 0 : (BM == 64 ? 1 : 2);
  const int bid = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  const GemmOp op = ops[find_op(ops, nops, SH, bid)];
  if (op.M <= 0 || op.N <= 0 || op.K <= 0) return;
  const int tn = (op.N + BN - 1) / BN;
  const int tiles = ((op.M + BM - 1) / BM) * tn;
  const int tile = bid - op.t0[SH];
  if (tile < 0 || tile >= tiles) return;
  // tiles in groups of 8 tile rows, column-major inside a group: the blocks
  // resident on one XCD at a time share a few A row panels and B column panels
  const int tm = (op.M + BM - 1) / BM, per_group = 8 * tn;
  const int g0 = (tile / per_group) * 8, gm = min(tm - g0, 8), t_in = tile % per_group;
  const int m0 = (g0 + t_in % gm) * BM, n0 = (t_in / gm) * BN;
  const C* A = (const C*)Abase + op.a;
  const C* B = (const C*)Bbase + op.b;
  C* Cm = (C*)Cbase + op.c;

  __shared__ T As[2][2][KC][BM + 1];   // [buffer][re, im][k][row]
  __shared__ T Bs[2][2][KC][BN + 1];   // [buffer][re, im][k][col]

  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wy = wv >> 1, wx = wv & 1;
  const int c16 = lane & 15, kk = lane >> 4;

  // global -> register assignment
  //   A (row-major): thread = (row, AP consecutive k); TA: (k, AP consecutive rows)
  //   B (row-major): thread = (k, BP consecutive cols); TB: (col, BP consecutive k)
  const int a_r = TA ? (tid % (BM / AP)) * AP : tid / (KC / AP);
  const int a_k = TA ? tid / (BM / AP) : (tid % (KC / AP)) * AP;
  const int b_k = TB ? (tid % (KC / BP)) * BP : tid / (BN / BP);
  const int b_c = TB ? tid / (KC / BP) : (tid % (BN / BP)) * BP;
  C ra[AP], rb[BP];
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < AP; ++q) {
      const int gr = m0 + (TA ? a_r + q : a_r), gk = k0 + (TA ? a_k : a_k + q);
      C v;
      v.x = 0; v.y = 0;
      if (gr < op.M && gk < op.K) v = TA ? A[(size_t)gk * op.lda + gr] : A[(size_t)gr * op.lda + gk];
      ra[q] = v;
    }
#pragma unroll
    for (int q = 0; q < BP; ++q) {
      const int gk = k0 + (TB ? b_k + q : b_k), gc = n0 + (TB ? b_c : b_c + q);
      C v;
      v.x = 0; v.y = 0;
      if (gk < op.K && gc < op.N) v = TB ? B[(size_t)gc * op.ldb + gk] : B[(size_t)gk * op.ldb + gc];
      rb[q] = v;
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int q = 0; q < AP; ++q) {
      const int r = TA ? a_r + q : a_r, k = TA ? a_k : a_k + q;
      As[buf][0][k][r] = ra[q].x;
      As[buf][1][k][r] = ra[q].y;
    }
#pragma unroll
    for (int q = 0; q < BP; ++q) {
      const int k = TB ? b_k + q : b_k, c = TB ? b_c : b_c + q;
      Bs[buf][0][k][c] = rb[q].x;
      Bs[buf][1][k][c] = rb[q].y;
    }
  };

  V4 accR[WM][WN], accI[WM][WN];
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j) { accR[i][j] = V4{0, 0, 0, 0}; accI[i][j] = V4{0, 0, 0, 0}; }
  const int nk = (op.K + KC - 1) / KC;
  gload(0);
  // the C tile is loaded up front into registers of its own: its latency
  // overlaps the first chunk, and with the short K of the LU updates (64) the
  // epilogue is a subtraction and a store rather than a dependent load.  A B
  // still accumulates from zero, so the rounding of the update stays at its
  // own scale (accumulating onto C would round every MFMA step at C's scale)
  // (PRE: the register cost lowers occupancy, so it is a tile variant)
  C cpre[WM][WN][4];
  if (PRE) {
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
      for (int j = 0; j < WN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gr = m0 + wy * 16 * WM + i * 16 + mrow<T>(kk, r);
          const int gc = n0 + wx * 16 * WN + j * 16 + c16;
          C v;
          v.x = 0; v.y = 0;
          if (gr < op.M && gc < op.N) v = Cm[(size_t)gr * op.ldc + gc];
          cpre[i][j][r] = v;
        }
  }
  lstore(0);
  __syncthreads();
  for (int kc = 0; kc < nk; ++kc) {
    const int buf = kc & 1;
    if (kc + 1 < nk) gload((kc + 1) * KC);
#pragma unroll
    for (int ks = 0; ks < KC / 4; ++ks) {
      const int kx = 4 * ks + kk;
      T ar[WM], ai[WM], nai[WM], br[WN], bi[WN];
#pragma unroll
      for (int i = 0; i < WM; ++i) {
        ar[i] = As[buf][0][kx][wy * 16 * WM + i * 16 + c16];
        ai[i] = As[buf][1][kx][wy * 16 * WM + i * 16 + c16];
        nai[i] = -ai[i];
      }
#pragma unroll
      for (int j = 0; j < WN; ++j) {
        br[j] = Bs[buf][0][kx][wx * 16 * WN + j * 16 + c16];
        bi[j] = Bs[buf][1][kx][wx * 16 * WN + j * 16 + c16];
      }
#pragma unroll
      for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j) {
          accR[i][j] = mfma4(ar[i], br[j], accR[i][j]);
          accR[i][j] = mfma4(nai[i], bi[j], accR[i][j]);
          accI[i][j] = mfma4(ar[i], bi[j], accI[i][j]);
          accI[i][j] = mfma4(ai[i], br[j], accI[i][j]);
        }
    }
    if (kc + 1 < nk) lstore(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gr = m0 + wy * 16 * WM + i * 16 + mrow<T>(kk, r);
        const int gc = n0 + wx * 16 * WN + j * 16 + c16;
        if (gr < op.M && gc < op.N) {
          C v = PRE ? cpre[i][j][r] : Cm[(size_t)gr * op.ldc + gc];
          v.x -= accR[i][j][r];
          v.y -= accI[i][j][r];
          Cm[(size_t)gr * op.ldc + gc] = v;
        }
      }
}

// ─────────────────────────────────────────────────────────────────────────────
// FP32 complex GEMM on the bf16 MFMA: each FP32 operand is split exactly into
// three bf16 parts (x = x0 + x1 + x2, 8 significant bits each), and the six
// products x_i y_j with i + j <= 2 are accumulated in FP32 (the MFMA's own
// accumulation rounds to nearest, checked on gfx950); the three dropped ones
// stay below 2 x 2^-24 of the product, so the result carries FP32 accuracy.
// A 32x32x16 block of real products (32768 flops) takes six
// v_mfma_f32_32x32x16_bf16 of 32 cycles = 192 cycles, against 16
// v_mfma_f32_16x16x4_f32 of 32 cycles = 512 cycles on the FP32 MFMA.
// ─────────────────────────────────────────────────────────────────────────────

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

// x = p0 + p1 + p2 exactly (x finite, normal range), two values (k, k + 1) at
// a time, packed into one dword per part (k in the low half) as the MFMA
// operand expects: p0 is x rounded to bf16, p1 the rounded remainder (at most
// 16 significant bits before rounding), p2 what is left (at most 8, so exact).
// Rounding to nearest lets p1 and p2 take either sign, so the three dropped
// products (p1 q2, p2 q1, p2 q2) are unbiased; with truncated parts they all
// carry the sign of the product, and their coherent sum moved the monostatic
// RCS by 4.4e-3 dB (against 1.7e-3 dB rounded and 1.0e-3 dB for the FP32
// MFMA; 1200-element ogive, |m| <= 70, measured).
typedef float f2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void split3x2(float x0, float x1, unsigned& p0, unsigned& p1, unsigned& p2) {
  const f2v x = {x0, x1};
  const bf2v h0 = __builtin_convertvector(x, bf2v);
  const f2v r = x - __builtin_convertvector(h0, f2v);
  const bf2v h1 = __builtin_convertvector(r, bf2v);
  const f2v t = r - __builtin_convertvector(h1, f2v);
  const bf2v h2 = __builtin_convertvector(t, bf2v);
  p0 = __builtin_bit_cast(unsigned, h0);
  p1 = __builtin_bit_cast(unsigned, h1);
  p2 = __builtin_bit_cast(unsigned, h2);
}

// One 32x32x16 block.  Lane l supplies A[l%32][8(l/32)+j] and B[8(l/32)+j][l%32]
// (j = 0..7); accumulator element r holds D[8(r/4) + 4(l/32) + r%4][l%32]
// (checked on gfx950 against a host product).
__device__ __forceinline__ f16v mfma32(bf16x8 a, bf16x8 b, f16v c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// C -= A B, complex FP32, 64 x 64 tile per 256-thread block (2 x 2 wavefronts
// of one 32 x 32 block each), K in chunks of 16 staged through one LDS buffer
// (37 KB: four work-groups per CU) as the three bf16 parts of the real and
// imaginary parts, with the next chunk's global loads in flight during the
// current chunk's MFMAs.  (Two LDS buffers, 74 KB and two work-groups per CU,
// measured slower: 36.6 against 35.1 ms per step at n ~ 2400; so were
// 64 x 128 and 128 x 128 tiles, held to one work-group per CU.)  Same op list
// and tile order as k_gemm2<float, 2, 2>.  Every thread stages four
// consecutive k of one row of A and one column of B (one 8-byte LDS write per
// part); when the operand's k is the strided index in memory the rows /
// columns run over the lanes, so the global loads stay coalesced.  Chunks
// inside the matrix take their loads from pointers advanced per chunk with
// no bounds tests.  Three accumulators: Re A Re B, Im A Im B and the
// imaginary part; Re C = first - second.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm_bf3(const GemmOp* ops, int nops, const void* Abase,
                                                  const void* Bbase, void* Cbase) {
  typedef float2 C;
  constexpr int BM = 64, BN = 64, KC = 16, KP = KC + 8;   // KP: plane row pitch (48 bytes)
  const int bid = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  const GemmOp op = ops[find_op(ops, nops, 0, bid)];
  if (op.M <= 0 || op.N <= 0 || op.K <= 0) return;
  const int tn = (op.N + BN - 1) / BN;
  const int tiles = ((op.M + BM - 1) / BM) * tn;
  const int tile = bid - op.t0[0];
  if (tile < 0 || tile >= tiles) return;
  const int tm = (op.M + BM - 1) / BM, per_group = 8 * tn;
  const int g0 = (tile / per_group) * 8, gm = min(tm - g0, 8), t_in = tile % per_group;
  const int m0 = (g0 + t_in % gm) * BM, n0 = (t_in / gm) * BN;
  const C* A = (const C*)Abase + op.a;
  const C* B = (const C*)Bbase + op.b;
  C* Cm = (C*)Cbase + op.c;

  // [part][re, im][row or col][k]: each lane's eight k of one row / column
  // are one 16-byte read
  __shared__ __attribute__((aligned(16))) __bf16 As[3][2][BM][KP];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[3][2][BN][KP];

  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wy = wv >> 1, wx = wv & 1;
  const int l32 = lane & 31, kh = lane >> 5;

  // staging: row / column rc, k = k0 + 4 kq + q (q = 0..3)
  const int a_rc = TA ? (tid & 63) : (tid >> 2), a_kq = TA ? (tid >> 6) : (tid & 3);
  const int b_rc = TB ? (tid >> 2) : (tid & 63), b_kq = TB ? (tid & 3) : (tid >> 6);
  const int gr = m0 + a_rc, gc = n0 + b_rc;
  const bool rowin = gr < op.M, colin = gc < op.N;
  // element (k0 + 4 kq + q) of this thread's row / column at pa[q * sa] (+ k0 * ka)
  const size_t sa = TA ? (size_t)op.lda : 1, ka = TA ? (size_t)op.lda : 1;
  const size_t sb = TB ? 1 : (size_t)op.ldb, kb = TB ? 1 : (size_t)op.ldb;
  const C* pa = A + (TA ? (size_t)(4 * a_kq) * op.lda + (rowin ? gr : 0) : (size_t)(rowin ? gr : 0) * op.lda + 4 * a_kq);
  const C* pb = B + (TB ? (size_t)(colin ? gc : 0) * op.ldb + 4 * b_kq : (size_t)(4 * b_kq) * op.ldb + (colin ? gc : 0));
  C ra[4], rb[4];
  auto gload = [&](int k0) {
    if (k0 + KC <= op.K) {
      const C* qa = pa + (size_t)k0 * ka;
      const C* qb = pb + (size_t)k0 * kb;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        ra[q] = qa[q * sa];
        rb[q] = qb[q * sb];
      }
      if (!rowin) {
#pragma unroll
        for (int q = 0; q < 4; ++q) { ra[q].x = 0; ra[q].y = 0; }
      }
      if (!colin) {
#pragma unroll
        for (int q = 0; q < 4; ++q) { rb[q].x = 0; rb[q].y = 0; }
      }
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        C v;
        v.x = 0; v.y = 0;
        if (rowin && k0 + 4 * a_kq + q < op.K) v = pa[(size_t)k0 * ka + q * sa];
        ra[q] = v;
        C w;
        w.x = 0; w.y = 0;
        if (colin && k0 + 4 * b_kq + q < op.K) w = pb[(size_t)k0 * kb + q * sb];
        rb[q] = w;
      }
    }
  };
  auto put = [&](__bf16* plane0, size_t pstride, size_t reim, const C* v) {
    unsigned re[3][2], im[3][2];
    split3x2(v[0].x, v[1].x, re[0][0], re[1][0], re[2][0]);
    split3x2(v[2].x, v[3].x, re[0][1], re[1][1], re[2][1]);
    split3x2(v[0].y, v[1].y, im[0][0], im[1][0], im[2][0]);
    split3x2(v[2].y, v[3].y, im[0][1], im[1][1], im[2][1]);
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      *(uint2*)(plane0 + p * pstride) = make_uint2(re[p][0], re[p][1]);
      *(uint2*)(plane0 + p * pstride + reim) = make_uint2(im[p][0], im[p][1]);
    }
  };
  auto lstore = [&]() {
    put(&As[0][0][a_rc][4 * a_kq], (size_t)2 * BM * KP, (size_t)BM * KP, ra);
    put(&Bs[0][0][b_rc][4 * b_kq], (size_t)2 * BN * KP, (size_t)BN * KP, rb);
  };

  f16v pRR, pII, pIm;
#pragma unroll
  for (int r = 0; r < 16; ++r) { pRR[r] = 0.f; pII[r] = 0.f; pIm[r] = 0.f; }
  const int nk = (op.K + KC - 1) / KC;
  const int arow = wy * 32 + l32, bcol = wx * 32 + l32, koff = 8 * kh;
  gload(0);
  for (int kc = 0; kc < nk; ++kc) {
    lstore();
    __syncthreads();
    if (kc + 1 < nk) gload((kc + 1) * KC);
    bf16x8 ar[3], ai[3], br[3], bi[3];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      ar[p] = *(const bf16x8*)&As[p][0][arow][koff];
      ai[p] = *(const bf16x8*)&As[p][1][arow][koff];
      br[p] = *(const bf16x8*)&Bs[p][0][bcol][koff];
      bi[p] = *(const bf16x8*)&Bs[p][1][bcol][koff];
    }
    // the smallest terms first: (2,0) (1,1) (0,2), (1,0) (0,1), (0,0)
#pragma unroll
    for (int t = 5; t >= 0; --t) {
      const int u = t == 5 ? 2 : (t == 4 ? 1 : (t == 2 ? 1 : 0));
      const int v = t == 3 ? 2 : (t == 4 ? 1 : (t == 1 ? 1 : 0));
      pRR = mfma32(ar[u], br[v], pRR);
      pII = mfma32(ai[u], bi[v], pII);
      pIm = mfma32(ar[u], bi[v], pIm);
      pIm = mfma32(ai[u], br[v], pIm);
    }
    __syncthreads();   // every wave has read the chunk before the next is staged
  }
  const int gcol = n0 + bcol;
  if (gcol < op.N) {
    C* pc = Cm + (size_t)(m0 + wy * 32 + 4 * kh) * op.ldc + gcol;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int dr = 8 * (r >> 2) + (r & 3);
      if (m0 + wy * 32 + 4 * kh + dr < op.M) {
        C* p = pc + (size_t)dr * op.ldc;
        C v = *p;
        v.x -= pRR[r] - pII[r];
        v.y -= pIm[r];
        *p = v;
      }
    }
  }
}

// Two-column triangular solves of the error probe with the FP32 factors, one
// 1024-thread work-group per system (in place on the n x 2 block at x):
// forward (UPPER = false) through the unit lower factor, y_b = T^-1 y_b with
// W_l = I - T^-1, then the rows below take y_b's contribution; backward
// (UPPER = true) the same with U and the rows above.  The same algebra as the
// blocked GEMM schedule of the wide solves, without its ~150 launches of a
// few microseconds each: every row of the factor is streamed once per sweep.
struct TrsvOp {
  long long a;       // element offset of the system's factor
  long long w;       // element offset of the system's first W slot (block b at w + 2 * 4096 b)
  long long x;       // element offset of the n x 2 block
  int ld, n;
};

template <bool UPPER>
__global__ __launch_bounds__(1024) void k_trsv2(const TrsvOp* ops, const float2* F, const float2* Wbuf,
                                                float2* X) {
  const TrsvOp op = ops[blockIdx.x];
  const int n = op.n, tid = threadIdx.x;
  if (n <= 0) return;
  const int nblk = (n + 63) / 64;
  __shared__ float2 Wt[64][65];
  __shared__ float2 yb[64][2];
  float2* x = X + op.x;
  for (int s = 0; s < nblk; ++s) {
    const int b = UPPER ? nblk - 1 - s : s;
    const int k = 64 * b, w = min(64, n - k);
    const float2* W = Wbuf + op.w + (size_t)b * 2 * 4096 + (UPPER ? 4096 : 0);
    for (int i = tid; i < 4096; i += 1024) Wt[i >> 6][i & 63] = W[i];
    if (tid < 128) {
      const int r = tid >> 1, c = tid & 1;
      yb[r][c] = r < w ? x[(size_t)(k + r) * 2 + c] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    // y_b = y_b - W y_b (W row-major 64 x 64; its padding beyond w is not read)
    float2 v = make_float2(0.f, 0.f);
    const int r = tid >> 1, c = tid & 1;
    if (tid < 128 && r < w) {
      v = yb[r][c];
      for (int q = 0; q < w; ++q) {
        const float2 a = Wt[r][q], y = yb[q][c];
        v.x -= a.x * y.x - a.y * y.y;
        v.y -= a.x * y.y + a.y * y.x;
      }
    }
    __syncthreads();
    if (tid < 128 && r < w) {
      yb[r][c] = v;
      x[(size_t)(k + r) * 2 + c] = v;
    }
    __syncthreads();
    // the other rows: below the block (forward) or above it (backward).  A
    // 16-lane group per row, 4 columns per lane (two 16-byte loads), so a
    // wavefront reads 4 rows x 512 contiguous bytes; the group's partial dot
    // products are summed by row rotations.
    const int lo = UPPER ? 0 : k + w, hi = UPPER ? k : n;
    const int lane = tid & 63, wv = tid >> 6, sub = lane >> 4, ql = (lane & 15) * 4;
    for (int row0 = lo + 4 * wv; row0 < hi; row0 += 64) {
      const int row = row0 + sub;
      float p0r = 0.f, p0i = 0.f, p1r = 0.f, p1i = 0.f;
      if (row < hi) {
        const float2* fr = F + op.a + (size_t)row * op.ld + k + ql;
        float2 l[4];
        if (ql + 4 <= w) {
          const float4 u0 = *(const float4*)fr, u1 = *(const float4*)(fr + 2);
          l[0] = make_float2(u0.x, u0.y); l[1] = make_float2(u0.z, u0.w);
          l[2] = make_float2(u1.x, u1.y); l[3] = make_float2(u1.z, u1.w);
        } else {
#pragma unroll
          for (int j = 0; j < 4; ++j) l[j] = ql + j < w ? fr[j] : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int q = min(ql + j, 63);   // l[j] = 0 past w
          const float2 y0 = yb[q][0], y1 = yb[q][1];
          p0r += l[j].x * y0.x - l[j].y * y0.y;
          p0i += l[j].x * y0.y + l[j].y * y0.x;
          p1r += l[j].x * y1.x - l[j].y * y1.y;
          p1i += l[j].x * y1.y + l[j].y * y1.x;
        }
      }
      p0r += row_ror<8>(p0r); p0i += row_ror<8>(p0i); p1r += row_ror<8>(p1r); p1i += row_ror<8>(p1i);
      p0r += row_ror<4>(p0r); p0i += row_ror<4>(p0i); p1r += row_ror<4>(p1r); p1i += row_ror<4>(p1i);
      p0r += row_ror<2>(p0r); p0i += row_ror<2>(p0i); p1r += row_ror<2>(p1r); p1i += row_ror<2>(p1i);
      p0r += row_ror<1>(p0r); p0i += row_ror<1>(p0i); p1r += row_ror<1>(p1r); p1i += row_ror<1>(p1i);
      if ((lane & 15) == 0 && row < hi) {
        float2* xr = x + (size_t)row * 2;
        const float2 a0 = xr[0], a1 = xr[1];
        xr[0] = make_float2(a0.x - p0r, a0.y - p0i);
        xr[1] = make_float2(a1.x - p1r, a1.y - p1i);
      }
    }
    __syncthreads();
  }
}

// ─────────────────────────────────────────────────────────────────────────────
// panel factorisation: one 1024-thread block per system, SW-column sub-panels
// held in registers (rows r0 + tid + 1024 i), pivots by max |Re| + |Im|
// ─────────────────────────────────────────────────────────────────────────────

template <typename T> __device__ __forceinline__ typename CT<T>::c cdiv1(typename CT<T>::c z) {
  // 1 / z with scaling (Smith's method)
  typename CT<T>::c r;
  if (fabs(z.x) >= fabs(z.y)) {
    const T q = z.y / z.x, d = z.x + z.y * q;
    r.x = (T)1 / d; r.y = -q / d;
  } else {
    const T q = z.x / z.y, d = z.y + z.x * q;
    r.x = q / d; r.y = (T)-1 / d;
  }
  return r;
}

// Pivot arg-max on DPP (no LDS round trips): (v, i) takes (w, j) when w is
// larger, or equal with the smaller row index (the first maximum, as izamax).
template <typename T> __device__ __forceinline__ void amax_take(T& v, int& i, T w, int j) {
  if (w > v || (w == v && j < i)) { v = w; i = j; }
}
// DPP move of x into the rows selected by RM; other rows keep `old`
template <int CTRL, int RM> __device__ __forceinline__ int dpp_upd(int old, int x) {
  return __builtin_amdgcn_update_dpp(old, x, CTRL, RM, 0xF, false);
}
template <int CTRL, int RM> __device__ __forceinline__ float dpp_upd(float old, float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(x), CTRL, RM, 0xF, false));
}
template <int CTRL, int RM> __device__ __forceinline__ double dpp_upd(double old, double x) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(x), CTRL, RM, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(x), CTRL, RM, 0xF, false);
  return __hiloint2double(hi, lo);
}
template <int CTRL, int RM, typename T> __device__ __forceinline__ void amax_dpp(T& v, int& i) {
  amax_take(v, i, dpp_upd<CTRL, RM>(v, v), dpp_upd<CTRL, RM>(i, i));
}
// every lane of a 16-lane DPP row ends with the row's arg-max: quad_perm
// [1,0,3,2] and [2,3,0,1] combine a quad, row_ror 4 and 8 the four quads
// (the operation is idempotent, so rotations serve as butterflies)
template <typename T> __device__ __forceinline__ void row_argmax(T& v, int& i) {
  amax_dpp<0xB1, 0xF>(v, i);
  amax_dpp<0x4E, 0xF>(v, i);
  amax_dpp<0x124, 0xF>(v, i);
  amax_dpp<0x128, 0xF>(v, i);
}
// the wavefront's arg-max, complete in lane 63: row_bcast:15 brings row r's
// result into row r + 1 (rows 1 and 3 written), row_bcast:31 brings row 1's
// into rows 2 and 3
template <typename T> __device__ __forceinline__ void wave_argmax_to63(T& v, int& i) {
  row_argmax(v, i);
  amax_dpp<0x142, 0xA>(v, i);
  amax_dpp<0x143, 0xC>(v, i);
}
__device__ __forceinline__ float readlane_t(float x, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ double readlane_t(double x, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), l),
                          __builtin_amdgcn_readlane(__double2loint(x), l));
}

// SW contiguous elements of one sub-panel row in 16-byte accesses.  The
// segment is 16-byte aligned: matrix offsets and leading dimensions are
// multiples of 16 elements and sub-panels start at multiples of SW (even).
template <int SW> __device__ __forceinline__ void load_seg(const float2* p, float2 (&v)[SW]) {
#pragma unroll
  for (int c = 0; c < SW; c += 2) {
    const float4 q = *(const float4*)(p + c);
    v[c] = make_float2(q.x, q.y);
    v[c + 1] = make_float2(q.z, q.w);
  }
}
template <int SW> __device__ __forceinline__ void load_seg(const double2* p, double2 (&v)[SW]) {
#pragma unroll
  for (int c = 0; c < SW; ++c) v[c] = p[c];
}
template <int SW> __device__ __forceinline__ void store_seg(float2* p, const float2 (&v)[SW]) {
#pragma unroll
  for (int c = 0; c < SW; c += 2) *(float4*)(p + c) = make_float4(v[c].x, v[c].y, v[c + 1].x, v[c + 1].y);
}
template <int SW> __device__ __forceinline__ void store_seg(double2* p, const double2 (&v)[SW]) {
#pragma unroll
  for (int c = 0; c < SW; ++c) p[c] = v[c];
}

// Block step of the two-level recursive panel: columns [k0, k0 + KB) are
// factored (rows [k0, n) hold L, row-interchanged like every panel column),
// columns [k0 + KB, k0 + KB + TW) are not yet touched by them.  U = L11^-1 A12
// for rows [k0, k0 + KB) of the target columns (unit lower L11, one thread
// per column, in LDS), then the rank-KB update A22 -= L21 U of rows
// [k0 + KB, n) on the 16x16x4 MFMA: 16 wavefronts walk the 16 x 16 tiles, each
// complex product four real MFMAs (Re A Re B, Im A Im B, and both cross terms
// into one accumulator), K = KB in steps of 4.  One barrier-separated step of
// the 1024-thread panel work-group; it replaces the VALU rank-SW updates of
// the target columns by the factored block's sub-panels.
template <typename T, int KB, int TW>
__device__ void panel_block_update(typename CT<T>::c* A, int n, int ld, int k0, int tid) {
  typedef typename CT<T>::c C;
  typedef typename CT<T>::v4 V4;
  __shared__ C Lb[KB][KB + 1];
  __shared__ C Ub[KB][TW + 1];
  for (int i = tid; i < KB * KB; i += 1024) {
    const int r = i / KB, c = i % KB;
    Lb[r][c] = A[(size_t)(k0 + r) * ld + k0 + c];
  }
  for (int i = tid; i < KB * TW; i += 1024) {
    const int r = i / TW, c = i % TW;
    Ub[r][c] = A[(size_t)(k0 + r) * ld + k0 + KB + c];
  }
  __syncthreads();
  if (tid < TW) {
    // forward substitution down column tid of U (L11 unit lower)
    for (int r = 1; r < KB; ++r) {
      C z = Ub[r][tid];
      for (int q = 0; q < r; ++q) z = csub(z, cmul(Lb[r][q], Ub[q][tid]));
      Ub[r][tid] = z;
    }
    for (int r = 0; r < KB; ++r) A[(size_t)(k0 + r) * ld + k0 + KB + tid] = Ub[r][tid];
  }
  __syncthreads();
  const int lane = tid & 63, wv = tid >> 6;
  const int rows = n - (k0 + KB);
  constexpr int CT_ = TW / 16;
  const int ntile = ((rows + 15) >> 4) * CT_;
  const int lr = lane & 15, lk = lane >> 4;
  for (int t = wv; t < ntile; t += 16) {
    const int r0 = k0 + KB + 16 * (t / CT_), cb = 16 * (t % CT_);
    V4 arr, aii, aim;
#pragma unroll
    for (int e = 0; e < 4; ++e) { arr[e] = (T)0; aii[e] = (T)0; aim[e] = (T)0; }
    const int row = r0 + lr;
    const C* lrow = A + (size_t)(row < n ? row : k0) * ld + k0;
#pragma unroll
    for (int kk = 0; kk < KB / 4; ++kk) {
      const int q = 4 * kk + lk;
      C a = lrow[q];
      if (row >= n) { a.x = (T)0; a.y = (T)0; }
      const C b = Ub[q][cb + lr];
      arr = mfma4(a.x, b.x, arr);
      aii = mfma4(a.y, b.y, aii);
      aim = mfma4(a.x, b.y, aim);
      aim = mfma4(a.y, b.x, aim);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int rr = r0 + mrow<T>(lk, e);
      if (rr < n) {
        C* pc = A + (size_t)rr * ld + k0 + KB + cb + lr;
        C z = *pc;
        z.x -= arr[e] - aii[e];
        z.y -= aim[e];
        *pc = z;
      }
    }
  }
  __syncthreads();
}

template <typename T, int RPT, int SW>
__global__ __launch_bounds__(1024) void k_lu_panel(const PanelOp* ops, void* base, int* pivots, int* info) {
  typedef typename CT<T>::c C;
  const PanelOp op = ops[blockIdx.x];
  if (op.w <= 0) return;
  C* A = (C*)base + op.a;
  int* ipiv = pivots + op.piv;
  const int n = op.n, ld = op.ld, k = op.k, w = op.w;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;

  __shared__ T red_v[2][16];
  __shared__ int red_i[2][16];
  __shared__ C prow[SW];
  __shared__ C srow[SW];
  __shared__ C L11[SW][SW];
  __shared__ C U12[SW][64];
  __shared__ int piv_s[SW];

  for (int j0 = 0; j0 < w; j0 += SW) {
    const int ws = min(SW, w - j0);
    const int c0 = k + j0;
    C v[RPT][SW];
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int row = c0 + tid + 1024 * i;
      if (ws == SW && row < n) {
        load_seg<SW>(A + (size_t)row * ld + c0, v[i]);
      } else {
#pragma unroll
        for (int c = 0; c < SW; ++c) {
          C z;
          z.x = 0; z.y = 0;
          if (row < n && c < ws) z = A[(size_t)row * ld + c0 + c];
          v[i][c] = z;
        }
      }
    }
    // columns unrolled: jj is a constant in each copy, so the column's values
    // are plain registers and the rank-1 update covers only the columns right of it
#pragma unroll
    for (int jj = 0; jj < SW; ++jj) {
      if (jj >= ws) break;
      const int col = c0 + jj;
      // ── pivot search ──
      T best = (T)-1;
      int bidx = 0x7fffffff;
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        const int row = c0 + tid + 1024 * i;
        if (row >= col && row < n) {
          C z;
#pragma unroll
          for (int c = 0; c < SW; ++c) if (c == jj) z = v[i][c];
          const T mag = fabs(z.x) + fabs(z.y);
          if (mag > best || (mag == best && row < bidx)) { best = mag; bidx = row; }
        }
      }
      wave_argmax_to63(best, bidx);
      // the 16 wave maxima, double-buffered by column parity: a wave writes the
      // next column's entry only after every wave has passed this column's
      // second barrier, i.e. after every read of this buffer
      if (lane == 63) { red_v[jj & 1][wv] = best; red_i[jj & 1][wv] = bidx; }
      __syncthreads();
      // every wave reduces the 16 entries itself (row 0 of the wave, then a
      // broadcast from lane 0)
      best = lane < 16 ? red_v[jj & 1][lane] : (T)-1;
      bidx = lane < 16 ? red_i[jj & 1][lane] : 0x7fffffff;
      row_argmax(best, bidx);
      best = readlane_t(best, 0);
      bidx = __builtin_amdgcn_readlane(bidx, 0);
      const bool singular = !(best > (T)0);
      const int pr = singular ? col : bidx;
      if (tid == 0) {
        if (singular && info) atomicAdd(info + op.sys, 1);
        ipiv[col] = pr;
        piv_s[jj] = pr;
      }
      // ── interchange rows col and pr inside the sub-panel ──
      const int oc = col - c0, op_ = pr - c0;
      if ((oc & 1023) == tid) {
#pragma unroll
        for (int i = 0; i < RPT; ++i)
          if (i == (oc >> 10))
#pragma unroll
            for (int c = 0; c < SW; ++c) srow[c] = v[i][c];
      }
      if ((op_ & 1023) == tid) {
#pragma unroll
        for (int i = 0; i < RPT; ++i)
          if (i == (op_ >> 10))
#pragma unroll
            for (int c = 0; c < SW; ++c) prow[c] = v[i][c];
      }
      __syncthreads();
      if (pr != col) {
        if ((oc & 1023) == tid) {
#pragma unroll
          for (int i = 0; i < RPT; ++i)
            if (i == (oc >> 10))
#pragma unroll
              for (int c = 0; c < SW; ++c) v[i][c] = prow[c];
        }
        if ((op_ & 1023) == tid) {
#pragma unroll
          for (int i = 0; i < RPT; ++i)
            if (i == (op_ >> 10))
#pragma unroll
              for (int c = 0; c < SW; ++c) v[i][c] = srow[c];
        }
      }
      // ── scale the column, rank-1 update of the sub-panel ──
      C pv;
#pragma unroll
      for (int c = 0; c < SW; ++c) if (c == jj) pv = prow[c];
      const bool ok = (pv.x != (T)0 || pv.y != (T)0);
      const C inv = ok ? cdiv1<T>(pv) : pv;
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        const int row = c0 + tid + 1024 * i;
        if (row > col && row < n && ok) {
          C l;
#pragma unroll
          for (int c = 0; c < SW; ++c) if (c == jj) l = v[i][c];
          const C lv = cmul(l, inv);
#pragma unroll
          for (int c = 0; c < SW; ++c) {
            if (c == jj) v[i][c] = lv;
            else if (c > jj && c < ws) v[i][c] = csub(v[i][c], cmul(lv, prow[c]));
          }
        }
      }
      // no barrier here: prow / srow are rewritten only after the next column's
      // first barrier, which every wave reaches after finishing this update
    }
    __syncthreads();
    // ── write the sub-panel back; keep L11 ──
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int row = c0 + tid + 1024 * i;
      if (row < n) {
        if (ws == SW) {
          store_seg<SW>(A + (size_t)row * ld + c0, v[i]);
        } else {
#pragma unroll
          for (int c = 0; c < SW; ++c)
            if (c < ws) A[(size_t)row * ld + c0 + c] = v[i][c];
        }
      }
    }
    if (tid < ws) {
#pragma unroll
      for (int c = 0; c < SW; ++c) L11[tid][c] = v[0][c];
    }
    __syncthreads();
    // ── interchanges on the other panel columns [k, c0) and [c0 + ws, k + w) ──
    const int nleft = c0 - k, nswap = k + w - (c0 + ws);
    // full panels run as a two-level recursion over 16-column quarters: a
    // sub-panel updates only its own quarter's columns; the next quarter (or
    // the right half, after the left half) takes the factored block's rank-16
    // (rank-32) update on the MFMA (panel_block_update, below)
    const bool split = w == 64;
    const int nright = split ? k + 16 * (j0 / 16 + 1) - (c0 + ws) : nswap;
    if (tid < nleft + nswap) {
      const int cc = tid < nleft ? k + tid : c0 + ws + (tid - nleft);
      for (int jj = 0; jj < ws; ++jj) {
        const int r1 = c0 + jj, r2 = piv_s[jj];
        if (r1 != r2) {
          C* p1 = A + (size_t)r1 * ld + cc;
          C* p2 = A + (size_t)r2 * ld + cc;
          const C t1 = *p1;
          *p1 = *p2;
          *p2 = t1;
        }
      }
    }
    __syncthreads();
    if (nright > 0) {
      // U12 = L11^{-1} A12 (unit lower), one thread per right column
      if (tid < nright) {
        const int cc = c0 + ws + tid;
#pragma unroll
        for (int r = 0; r < SW; ++r) {
          if (r < ws) {
            C z = A[(size_t)(c0 + r) * ld + cc];
#pragma unroll
            for (int q = 0; q < SW; ++q)
              if (q < r) z = csub(z, cmul(L11[r][q], U12[q][tid]));
            A[(size_t)(c0 + r) * ld + cc] = z;
            U12[r][tid] = z;
          }
        }
      }
      __syncthreads();
      // rank-ws update of rows [c0 + ws, n) on the right columns (nright <= 64 - ws):
      // 8 lanes per row over the columns with stride 8, so every access of a wave
      // covers 8 rows x 8 contiguous elements; the multipliers of the row are
      // read back from the sub-panel just written (same block, after the barrier)
      {
        const int rsub = tid >> 3, cq = tid & 7;
        for (int row = c0 + ws + rsub; row < n; row += 128) {
          const C* lrow = A + (size_t)row * ld + c0;
          C l[SW];
          if (ws == SW) {
            load_seg<SW>(lrow, l);
          } else {
#pragma unroll
            for (int q = 0; q < SW; ++q)
              if (q < ws) l[q] = lrow[q];
          }
          C* p = A + (size_t)row * ld + c0 + ws;
          for (int c = cq; c < nright; c += 8) {
            C z = p[c];
#pragma unroll
            for (int q = 0; q < SW; ++q)
              if (q < ws) z = csub(z, cmul(l[q], U12[q][c]));
            p[c] = z;
          }
        }
      }
      __syncthreads();
    }
    if (split && j0 + ws == 16) panel_block_update<T, 16, 16>(A, n, ld, k, tid);        // quarter 0 -> 1
    if (split && j0 + ws == 32) panel_block_update<T, 32, 32>(A, n, ld, k, tid);        // left half -> right half
    if (split && j0 + ws == 48) panel_block_update<T, 16, 16>(A, n, ld, k + 32, tid);   // quarter 2 -> 3
  }
}

// row interchanges of one panel on the columns outside it
template <typename T>
__global__ __launch_bounds__(256) void k_lu_swap(const SwapOp* ops, void* base, const int* pivots) {
  typedef typename CT<T>::c C;
  const SwapOp op = ops[blockIdx.y];
  if (op.w <= 0) return;
  const int nout = op.c_lo + (op.ncols - op.c_hi);
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nout) return;
  const int cc = t < op.c_lo ? t : op.c_hi + (t - op.c_lo);
  C* A = (C*)base + op.a;
  const int* ipiv = pivots + op.piv;
  for (int jj = 0; jj < op.w; ++jj) {
    const int r1 = op.k + jj, r2 = ipiv[op.k + jj];
    if (r1 != r2) {
      C* p1 = A + (size_t)r1 * op.ld + cc;
      C* p2 = A + (size_t)r2 * op.ld + cc;
      const C a1 = *p1;
      *p1 = *p2;
      *p2 = a1;
    }
  }
}

// W = I - T^{-1} for the w x w diagonal block T (w <= 64) of the factor:
// unit lower (UPPER = false) or upper (UPPER = true).  A following
// k_gemm_update with A = W and B = C = X then leaves X = T^{-1} X: the whole
// B column slice is read before any C element of that slice is written.
template <typename T, bool UPPER>
__global__ __launch_bounds__(256) void k_tri_inv(const TrsmOp* ops, const void* tbase, void* wbuf) {
  typedef typename CT<T>::c C;
  const TrsmOp op = ops[blockIdx.x];
  C* W = (C*)wbuf + op.wo;
  if (op.w <= 0) return;
  // T zero-padded to 64 x 64; with the padding rows and columns >= w come
  // out as the identity in T^{-1}, i.e. zero in W
  __shared__ C Tm[64][65];
  __shared__ C dinv[64];   // UPPER: 1 / U[r][r] (1 on the padding)
  const C* Tp = (const C*)tbase + op.t;
  const int w = op.w;
  for (int i = threadIdx.x; i < 64 * 64; i += 256) {
    const int r = i >> 6, c = i & 63;
    C z;
    z.x = 0; z.y = 0;
    if (r < w && c < w) z = Tp[(size_t)r * op.ldt + c];
    Tm[r][c] = z;
  }
  __syncthreads();
  if (UPPER && threadIdx.x < 64) {
    const int r = threadIdx.x;
    C one;
    one.x = (T)1; one.y = (T)0;
    dinv[r] = r < w ? cdiv1<T>(Tm[r][r]) : one;
  }
  if (UPPER) __syncthreads();
  // column j of the inverse, rows r = 4 i + s held by lane s of a quad in x[i]:
  // x_r = delta_rj - sum_{q < r} T[r][q] x_q (lower, unit diagonal), or
  // x_r = (delta_rj - sum_{q > r} U[r][q] x_q) / U[r][r] (upper); each lane
  // sums its own q = s (mod 4) and the quad reduces by DPP
  const int j = threadIdx.x >> 2, sl = threadIdx.x & 3;
  C x[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) { x[i].x = (T)0; x[i].y = (T)0; }
#pragma clang loop unroll(full)
  for (int step = 0; step < 64; ++step) {
    const int r = UPPER ? 63 - step : step;
    T pr = 0, pi = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int q = 4 * i + sl;
      const bool use = UPPER ? (4 * i + 3 > r && q > r) : (4 * i < r && q < r);
      if (UPPER ? (4 * i + 3 > r) : (4 * i < r)) {   // static: some lane of the quad may use it
        if (use) {
          const C t = Tm[r][q];
          pr += t.x * x[i].x - t.y * x[i].y;
          pi += t.x * x[i].y + t.y * x[i].x;
        }
      }
    }
    pr += dpp_mov<0xB1>(pr);
    pi += dpp_mov<0xB1>(pi);
    pr += dpp_mov<0x4E>(pr);
    pi += dpp_mov<0x4E>(pi);
    C z;
    z.x = (r == j ? (T)1 : (T)0) - pr;
    z.y = -pi;
    if (UPPER) z = cmul(z, dinv[r]);
    if (sl == (r & 3)) x[r >> 2] = z;
  }
  // W = I - T^{-1}, row-major 64 x 64
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int r = 4 * i + sl;
    C z;
    z.x = (r == j ? (T)1 : (T)0) - x[i].x;
    z.y = -x[i].y;
    W[r * 64 + j] = z;
  }
}

// ─────────────────────────────────────────────────────────────────────────────
// helpers for the mixed-precision refinement
// ─────────────────────────────────────────────────────────────────────────────

struct VecOp {       // an n x N block: src at s (lds), dst at d (ldd)
  long long s, d;
  int lds, ldd, n, N, piv;
  int sc;            // offset of this block's N column scales
};

// Power-of-two column scales of an n x N FP64 block (src at s, lds):
// scale[c] = 2^-e with max_r |v[r][c]| in [2^(e-1), 2^e) (1 for a zero
// column), so that the scaled column has max modulus in [1/2, 1) and survives
// the cast to FP32 whatever its magnitude (right-hand sides of high azimuthal
// orders fall to 1e-60 and below).  64 columns x 4 row groups per block.
__global__ __launch_bounds__(256) void k_col_scale(const VecOp* ops, const double2* src, double* scale) {
  const VecOp op = ops[blockIdx.z];
  const int c = blockIdx.x * 64 + threadIdx.x, rg = threadIdx.y;
  __shared__ double red[4][64];
  double m = 0.0;
  if (c < op.N)
    for (int r = rg; r < op.n; r += 4) {
      const double2 v = src[op.s + (size_t)r * op.lds + c];
      m = fmax(m, fmax(fabs(v.x), fabs(v.y)));
    }
  red[rg][threadIdx.x] = m;
  __syncthreads();
  if (rg == 0 && c < op.N) {
    m = fmax(fmax(red[0][threadIdx.x], red[1][threadIdx.x]), fmax(red[2][threadIdx.x], red[3][threadIdx.x]));
    // exponent clamped so that the scale stays a finite double (a column of
    // denormals, such as an excitation that vanishes to rounding, keeps
    // 2^1000 and stays tiny rather than overflowing)
    int e = 0;
    if (m > 0.0) frexp(m, &e);
    e = max(-1000, min(1000, e));
    scale[op.sc + c] = ldexp(1.0, -e);
  }
}

// Error probe ratios per system: eta = ||E / scale|| / ||XS|| over the n x 2
// probe block (E: FP32 corrections at op.d, XS: FP64 at op.s, both ld 2;
// column scales at op.sc), 0 for a zero block and +inf for E != 0 with XS = 0.
// One work-group per system; the sums run in a fixed order.
__global__ __launch_bounds__(256) void k_probe_eta(const VecOp* ops, const float2* E, const double2* XS,
                                                   const double* scale, double* eta) {
  const VecOp op = ops[blockIdx.x];
  __shared__ double re[256], rx[256];
  double e2 = 0.0, x2 = 0.0;
  const double f0 = 1.0 / scale[op.sc], f1 = 1.0 / scale[op.sc + 1];
  for (int i = threadIdx.x; i < 2 * op.n; i += 256) {
    const float2 e = E[op.d + i];
    const double2 x = XS[op.s + i];
    const double f = (i & 1) ? f1 : f0;
    e2 += ((double)e.x * e.x + (double)e.y * e.y) * f * f;
    x2 += x.x * x.x + x.y * x.y;
  }
  re[threadIdx.x] = e2;
  rx[threadIdx.x] = x2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) { re[threadIdx.x] += re[threadIdx.x + o]; rx[threadIdx.x] += rx[threadIdx.x + o]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double a = re[0], b = rx[0];
    eta[blockIdx.x] = b > 0 ? sqrt(a / b) : (a > 0 ? INFINITY : 0.0);
  }
}

// dst32[r][c] = (float) (src64[perm ? perm[r] : r][c] * scale[c])
__global__ void k_scale_cast(const VecOp* ops, const int* perm, const double2* src, const double* scale,
                             float2* dst) {
  const VecOp op = ops[blockIdx.z];
  const int r = blockIdx.y, c = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= op.n || c >= op.N) return;
  const int pr = perm ? perm[op.piv + r] : r;
  const double2 v = src[op.s + (size_t)pr * op.lds + c];
  const double f = scale[op.sc + c];
  dst[op.d + (size_t)r * op.ldd + c] = make_float2((float)(v.x * f), (float)(v.y * f));
}

// dst64 (+)= (double) src32 / scale[c]
__global__ void k_widen_scaled(const VecOp* ops, const float2* src, const double* scale, double2* dst, int add) {
  const VecOp op = ops[blockIdx.z];
  const int r = blockIdx.y, c = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= op.n || c >= op.N) return;
  const float2 v = src[op.s + (size_t)r * op.lds + c];
  const double f = 1.0 / scale[op.sc + c];
  double2* p = dst + op.d + (size_t)r * op.ldd + c;
  if (add) { double2 o = *p; o.x += (double)v.x * f; o.y += (double)v.y * f; *p = o; }
  else *p = make_double2((double)v.x * f, (double)v.y * f);
}

__global__ void k_zero_int(int* p, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0;
}

// dst64 = src64 (copy of an n x N block)
__global__ void k_copy64(const VecOp* ops, const double2* src, double2* dst) {
  const VecOp op = ops[blockIdx.z];
  const int r = blockIdx.y, c = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= op.n || c >= op.N) return;
  dst[op.d + (size_t)r * op.ldd + c] = src[op.s + (size_t)r * op.lds + c];
}

// permutation of the factor rows from LAPACK-style interchanges (one thread per system)
// One work-group per system (n <= 4096, the panel limit): the interchanges are
// replayed in order by one lane on an LDS copy (their chain is sequential;
// LDS keeps each step's latency short), then written out by the group.
__global__ __launch_bounds__(256) void k_perm_from_pivots(const VecOp* ops, int nsys, const int* pivots,
                                                          int* perm) {
  const int s = blockIdx.x;
  if (s >= nsys) return;
  const VecOp op = ops[s];
  if (op.n > 8192) return;   // the build rejects such systems
  __shared__ int p[8192];
  for (int i = threadIdx.x; i < op.n; i += blockDim.x) p[i] = i;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int* ip = pivots + op.piv;
    for (int i = 0; i < op.n; ++i) {
      const int j = ip[i];
      if (j != i) { const int t = p[i]; p[i] = p[j]; p[j] = t; }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < op.n; i += blockDim.x) perm[op.piv + i] = p[i];
}

}  // namespace ghost
