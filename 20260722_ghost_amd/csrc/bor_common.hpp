// Shared device helpers for the BoR-MoM HIP library (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#define GH_PI 3.14159265358979323846

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

template <typename R> struct CT;
template <> struct CT<double> { typedef double2 c; typedef d4v v4; };
template <> struct CT<float> { typedef float2 c; typedef f4v v4; };

// One 16x16x4 MFMA step.  Lane l supplies A[l%16][l/16] and B[l/16][l%16];
// the accumulator holds D[4*(l/16)+r][l%16] in element r.
__device__ __forceinline__ d4v mfma4(double a, double b, d4v c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4v mfma4(float a, float b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Row of the 16x16 result held by lane group kk = l/16 in accumulator element r
// (measured on gfx950: FP32 4*kk + r, FP64 kk + 4*r); the column is l%16.
template <typename R> __device__ __forceinline__ constexpr int mrow(int kk, int r);
template <> __device__ __forceinline__ constexpr int mrow<float>(int kk, int r) { return 4 * kk + r; }
template <> __device__ __forceinline__ constexpr int mrow<double>(int kk, int r) { return kk + 4 * r; }

// Far-fill sample arithmetic.  FP64: the library's sqrt, sincos and
// division.  FP32 (mixed precision): the hardware v_sqrt_f32, v_rcp_f32 and
// v_sin_f32 / v_cos_f32 (argument in revolutions, first reduced to
// [-1/2, 1/2]).  In FP32 the phase k R of a far pair (up to ~70 rad at 10
// wavelengths) already carries ~4e-6 rad of rounding, which a full-range
// library reduction cannot recover, so the cheaper forms lose nothing that
// matters and cut the sample cost several-fold.
__device__ __forceinline__ double sqrt_s(double x) { return sqrt(x); }
__device__ __forceinline__ float sqrt_s(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double rcp_s(double x) { return 1.0 / x; }
__device__ __forceinline__ float rcp_s(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ void sincos_s(double x, double* s, double* c) { sincos(x, s, c); }
__device__ __forceinline__ void sincos_s(float x, float* s, float* c) {
  const float r = x * 0.159154943091895336f;   // 1 / (2 pi)
  const float f = r - rintf(r);
  *s = __builtin_amdgcn_sinf(f);
  *c = __builtin_amdgcn_cosf(f);
}

// Row-rotate a 64-bit value within each 16-lane DPP row: lane c receives
// the value of lane (c - N) mod 16 of the same row.
template <int N> __device__ __forceinline__ double row_ror(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_mov_dpp(lo, 0x120 + N, 0xF, 0xF, true);
  hi = __builtin_amdgcn_mov_dpp(hi, 0x120 + N, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}
template <int N> __device__ __forceinline__ float row_ror(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x120 + N, 0xF, 0xF, true));
}

// Value of the DPP partner lane (ctrl: row_mirror 0x140, row_half_mirror
// 0x141, row_ror 0x121-0x12F, quad_perm codes 0x00-0xFF), 64-bit and 32-bit.
// Only permutations, where every lane has a source lane: bound_ctrl then
// changes nothing, and setting it lets the compiler fold the move into the
// instruction that uses the value (VOP2 with a DPP operand).
template <int CTRL> __device__ __forceinline__ double dpp_mov(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, true);
  hi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}
template <int CTRL> __device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}

// Reduce-scatter across wavefront halves (xadd32) or row pairs (xadd16) on
// gfx950's permlane swaps: lanes 0-31 return lo_l + lo_(l^32) and lanes 32-63
// hi_l + hi_(l^32); xadd16 the same with rows of 16 (even rows keep lo).
// v_permlane32_swap(lo, hi) leaves exactly the two addends in its two outputs
// (measured lane layout: output 0 holds lanes 0-31 of lo in the lower half and
// of hi in the upper half, output 1 lanes 32-63 of lo and of hi).
__device__ __forceinline__ float xadd32(float lo, float hi) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_int(lo), __float_as_int(hi), false, false);
  return __int_as_float(r[0]) + __int_as_float(r[1]);
}
__device__ __forceinline__ float xadd16(float lo, float hi) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_int(lo), __float_as_int(hi), false, false);
  return __int_as_float(r[0]) + __int_as_float(r[1]);
}
__device__ __forceinline__ double xadd32(double lo, double hi) {
  const auto a = __builtin_amdgcn_permlane32_swap(__double2loint(lo), __double2loint(hi), false, false);
  const auto b = __builtin_amdgcn_permlane32_swap(__double2hiint(lo), __double2hiint(hi), false, false);
  return __hiloint2double(b[0], a[0]) + __hiloint2double(b[1], a[1]);
}
__device__ __forceinline__ double xadd16(double lo, double hi) {
  const auto a = __builtin_amdgcn_permlane16_swap(__double2loint(lo), __double2loint(hi), false, false);
  const auto b = __builtin_amdgcn_permlane16_swap(__double2hiint(lo), __double2hiint(hi), false, false);
  return __hiloint2double(b[0], a[0]) + __hiloint2double(b[1], a[1]);
}

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// Per-Gauss-point generatrix data, element-major (4 points per element).
struct PointSet {
  const double* rho;
  const double* z;
  const double* trho;
  const double* tz;
  const double* w;    // Gauss weight x element length
  const double* T0;
  const double* T1;
  const double* D0;   // d(rho T0)/dt
  const double* D1;   // d(rho T1)/dt
};

// One mode system inside a batch.  Paired modes: Z_{-m} = S Z_m S with S =
// diag(I on the J_t rows, -I on the J_phi rows), and the plane-wave right-hand
// sides obey S b_{-m} = [b_m(VV), -b_m(HH)] column by column, so the currents
// of mode -m are x_{-m} = S [x_m(VV), -x_m(HH)]: only m >= 0 is solved, and
// the entry of mode -m (flip = 1) reads mode m's storage with those signs.
struct SysInfo {
  int prob;      // owning problem (body x frequency)
  int m;         // azimuthal mode
  int n;         // masked system size
  int nt;        // active J_t bases (rows [0, nt))
  int t0;        // first active J_t node
  int f0;        // first active J_phi node (rows [nt, n))
  int ld;        // leading dimension of the augmented row-major system
  int nrhs;      // this mode's right-hand sides (2 x n_theta)
  int ncols;     // right-hand-side columns of the solved system
  int col;       // first of this mode's columns among them
  int flip;      // 1: mode -m of a pair (J_phi rows and HH columns change sign)
  int owner;     // index of the solved system (in solve order)
  int pen;       // PMCHWT pair (penetrable body): the single-surface size n1,
                 // rows [0, n1) J and [n1, 2 n1) M / eta0, each laid out as
                 // above (J_t then J_phi); 0 for a single-surface system
  long long a_off;   // element offset of the augmented system [n x ld]
  long long x_off;   // element offset of the n x ncols solution / RHS copy
  long long a64_off; // element offset of the FP64 n x n copy (mixed precision)
};

// Sign of row r under the pairing's diagonal: S = diag(I, -I) on (J_t, J_phi)
// of a single-surface system, diag(S, -S) on (J, M / eta0) of a PMCHWT pair.
__host__ __device__ inline double pair_row_sign(const SysInfo& s, int r) {
  const bool second = s.pen && r >= s.pen;
  const int rr = second ? r - s.pen : r;
  const double b = rr >= s.nt ? -1.0 : 1.0;
  return second ? -b : b;
}
