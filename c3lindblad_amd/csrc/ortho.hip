// One-reduction Gram-Schmidt step with delayed reorthogonalisation (see ortho.h).
//
// Per step three launches on the caller's stream:
//   reduce_kernel         one pass over basis columns 0..j (column j = q) against q and w,
//                         one wave per column per chunk of rows, q and w held in registers:
//                         partials of s = V^H q (s_j = |q|^2), z = V^H w (z_j = q^H w),
//                         |w|^2 and sink . q;
//   finish_kernel         one workgroup per reduced number;
//   update_column_kernel  every workgroup derives the step scalars (r, c, g) from the
//                         reduced numbers; the first ones write the final column j-1,
//                         sink . v_j, H s and the tentative column j (one wave per
//                         Hessenberg row), the others make one pass over columns 0..j-1:
//                         v_j = (q - V s)/r and q_{j+1} = w/g - (c/(r g)) q - V (z - (c/r) s)/g.
// Algebra (columns 0..j-1 orthonormal, q the once-projected v_j, nu its tentative
// sub-diagonal, A v_{j-1} = V h + nu q):
//   r = sqrt(|q|^2 - |s|^2), v_j = (q - V s)/r, H[:j, j-1] = h + nu s, H[j, j-1] = nu r;
//   c = (q^H w - s^H z)/r = v_j^H w;
//   A v_j = (w - A V s)/r = V_{j+1} ([z; c] - H s)/r + u/r, u = w - V z - c v_j,
//   |u|^2 = |w|^2 - |z|^2 - |c|^2, q_{j+1} = u/|u|, tentative nu_j = |u|/r.

#include "ortho.h"

#include <algorithm>
#include <cmath>
#include <string>
#include <type_traits>

#include "../../include/c3l_hip.h"

extern "C" int c3l_internal_fail(int code, const char* message);

namespace c3l_ortho {
namespace {

// Sum over the 64 lanes of a wave with DPP moves (no LDS traffic), valid in lane 63:
// an inclusive scan inside each row of 16 lanes (row_shr 1, 2, 4, 8, zero-filled at
// the row start), then the row totals carried forward with row_bcast 15 (rows 1, 3)
// and row_bcast 31 (rows 2, 3).
template <int CTRL, int ROW_MASK, bool ZERO_FILL>
__device__ __forceinline__ double dpp_move(double v) {
  const int2 b = __builtin_bit_cast(int2, v);
  const int lo = __builtin_amdgcn_update_dpp(0, b.x, CTRL, ROW_MASK, 0xf, ZERO_FILL);
  const int hi = __builtin_amdgcn_update_dpp(0, b.y, CTRL, ROW_MASK, 0xf, ZERO_FILL);
  return __builtin_bit_cast(double, make_int2(lo, hi));
}

__device__ __forceinline__ double wave_total(double v) {
  v += dpp_move<0x111, 0xf, true>(v);  // row_shr:1
  v += dpp_move<0x112, 0xf, true>(v);  // row_shr:2
  v += dpp_move<0x114, 0xf, true>(v);  // row_shr:4
  v += dpp_move<0x118, 0xf, true>(v);  // row_shr:8
  v += dpp_move<0x142, 0xa, false>(v);  // row_bcast:15 into rows 1 and 3 (others add 0)
  v += dpp_move<0x143, 0xc, false>(v);  // row_bcast:31 into rows 2 and 3
  return v;
}

template <typename T>
struct Num;

template <>
struct Num<double> {
  static __device__ __forceinline__ double zero() { return 0.0; }
  static __device__ __forceinline__ double make(double re) { return re; }
  static __device__ __forceinline__ double re(double a) { return a; }
  static __device__ __forceinline__ double add(double a, double b) { return a + b; }
  static __device__ __forceinline__ double sub(double a, double b) { return a - b; }
  static __device__ __forceinline__ double mul(double a, double b) { return a * b; }
  static __device__ __forceinline__ double scale(double a, double s) { return a * s; }
  static __device__ __forceinline__ double abs2(double a) { return a * a; }
  static __device__ __forceinline__ double fma(double a, double b, double acc) { return acc + a * b; }   // acc + a b
  static __device__ __forceinline__ double cfma(double a, double b, double acc) { return acc + a * b; }  // acc + conj(a) b
  static __device__ __forceinline__ double wave_total(double v) { return c3l_ortho::wave_total(v); }
};

template <>
struct Num<double2> {
  static __device__ __forceinline__ double2 zero() { return make_double2(0.0, 0.0); }
  static __device__ __forceinline__ double2 make(double re) { return make_double2(re, 0.0); }
  static __device__ __forceinline__ double re(double2 a) { return a.x; }
  static __device__ __forceinline__ double2 add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
  static __device__ __forceinline__ double2 sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
  static __device__ __forceinline__ double2 mul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
  }
  static __device__ __forceinline__ double2 scale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
  static __device__ __forceinline__ double abs2(double2 a) { return a.x * a.x + a.y * a.y; }
  static __device__ __forceinline__ double2 fma(double2 a, double2 b, double2 acc) {
    return make_double2(acc.x + a.x * b.x - a.y * b.y, acc.y + a.x * b.y + a.y * b.x);
  }
  static __device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 acc) {
    return make_double2(acc.x + a.x * b.x + a.y * b.y, acc.y + a.x * b.y - a.y * b.x);
  }
  static __device__ __forceinline__ double2 wave_total(double2 v) {
    return make_double2(c3l_ortho::wave_total(v.x), c3l_ortho::wave_total(v.y));
  }
};

// Basis storage: FP64 (real or complex) or FP32 (real) columns; arithmetic in T.
__device__ __forceinline__ double widen(float v) { return (double)v; }
__device__ __forceinline__ double widen(double v) { return v; }
__device__ __forceinline__ double2 widen(double2 v) { return v; }
template <typename S, typename T>
__device__ __forceinline__ S narrow_to(T v) {
  if constexpr (std::is_same<S, T>::value) return v;
  else return static_cast<S>(v);
}

// Rows per reduction chunk (q and w stay in registers), basis columns loaded per
// iteration of a reduction wave, waves per update workgroup and basis rows per update
// lane.  Complex bases are short (tens of thousands of rows): 512-row chunks and 8-wave
// update groups of one row per lane keep the GPU full.  Real bases (the CCSD Newton
// correction, stored in FP32) are long: 256-row chunks (four rows per lane) give the
// reduction more waves (at 1.14 M rows 8-20 % less time than 512-row chunks), and four
// rows per update lane cut the update grid fourfold, which at a few columns is bound by
// the number of workgroups rather than by the bytes.
template <typename T>
struct Shape;
template <>
struct Shape<double> {
  static constexpr int kRows = 256;
  static constexpr int kColsPerIter = 2;
  static constexpr int kWaves = 4;
  static constexpr int kRowsPerLane = 4;
};
template <>
struct Shape<double2> {
  static constexpr int kRows = 512;
  static constexpr int kColsPerIter = 2;
  static constexpr int kWaves = 8;
  static constexpr int kRowsPerLane = 1;
};

// Grid (chunks, column groups), 4 waves per group; wave `wave` of group g takes
// columns g*4*cpw + wave + 4c (c < cpw).  Column nv is the extra one: its q-dot
// is sink . q (unconjugated, 0 without a sink) and its w-dot |w|^2.
// Partials: s-side output k at partial[k*chunks + chunk], w-side at
// partial[(nv+1+k)*chunks + chunk].
// Kernel bodies are device functions with restrict parameters, inlined into thin
// kernels that pick the instance's pointers (blockIdx.z).
template <typename T, typename S, int ROWS, int CPI>
__device__ __forceinline__ void reduce_body(const S* __restrict__ V, int64_t ld, int nv, int cpw,
                                            const T* __restrict__ w, const T* __restrict__ sink, int64_t dim,
                                            T* __restrict__ partial, int64_t chunks) {
  using N = Num<T>;
  constexpr int kPer = ROWS / 64;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int kbase = blockIdx.y * 4 * cpw + wave;
  if (kbase > nv) return;  // whole wave
  const int64_t r0 = blockIdx.x * (int64_t)ROWS + lane;
  const S* __restrict__ q = V + (int64_t)(nv - 1) * ld;
  // rows past the end read the last row (always valid) against zero q and w entries; q and
  // the basis columns stay in their storage type in registers and widen at the multiply
  S qv[kPer];
  T wv[kPer];
#pragma unroll
  for (int t = 0; t < kPer; ++t) {
    const int64_t i = r0 + t * 64;
    qv[t] = i < dim ? q[i] : S{};
    wv[t] = (w != nullptr && i < dim) ? w[i] : N::zero();
  }
  // basis columns, CPI per iteration so that CPI kPer loads are in flight
  for (int c = 0; c < cpw; c += CPI) {
    const int k0 = kbase + 4 * c;
    if (k0 >= nv) break;
    S v[CPI][kPer];
#pragma unroll
    for (int u = 0; u < CPI; ++u) {
      const int k = k0 + 4 * u;
      const S* __restrict__ col = V + (int64_t)((c + u < cpw && k < nv) ? k : k0) * ld;
#pragma unroll
      for (int t = 0; t < kPer; ++t) v[u][t] = col[min(r0 + t * 64, dim - 1)];
    }
    T as[CPI], az[CPI];
#pragma unroll
    for (int u = 0; u < CPI; ++u) {
      as[u] = N::zero();
      az[u] = N::zero();
#pragma unroll
      for (int t = 0; t < kPer; ++t) {
        const T a = widen(v[u][t]);
        as[u] = N::cfma(a, widen(qv[t]), as[u]);
        az[u] = N::cfma(a, wv[t], az[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < CPI; ++u) {
      as[u] = N::wave_total(as[u]);
      az[u] = N::wave_total(az[u]);
    }
    if (lane == 63) {
#pragma unroll
      for (int u = 0; u < CPI; ++u) {
        const int k = k0 + 4 * u;
        if (c + u < cpw && k < nv) {
          partial[(int64_t)k * chunks + blockIdx.x] = as[u];
          if (w != nullptr) partial[(int64_t)(nv + 1 + k) * chunks + blockIdx.x] = az[u];
        }
      }
    }
  }
  // the extra column nv belongs to the wave whose column list contains it
  const int off = nv - kbase;
  if (off % 4 == 0 && off / 4 < cpw) {
    T as = N::zero(), az = N::zero();
#pragma unroll
    for (int t = 0; t < kPer; ++t) {
      const int64_t i = r0 + t * 64;
      if (sink != nullptr && i < dim) as = N::fma(sink[i], widen(qv[t]), as);
      az = N::cfma(wv[t], wv[t], az);
    }
    as = N::wave_total(as);
    az = N::wave_total(az);
    if (lane == 63) {
      partial[(int64_t)nv * chunks + blockIdx.x] = as;
      if (w != nullptr) partial[(int64_t)(2 * nv + 1) * chunks + blockIdx.x] = az;
    }
  }
}

template <typename T, typename S, int ROWS, int CPI>
__global__ __launch_bounds__(256) void reduce_kernel(Many<const S*> V, int64_t ld, int nv, int cpw, Many<const T*> w,
                                                     Many<const T*> sink, int64_t dim, Many<T*> partial,
                                                     int64_t chunks) {
  const int i = blockIdx.z;
  reduce_body<T, S, ROWS, CPI>(V.p[i], ld, nv, cpw, w.p[i], sink.p[i], dim, partial.p[i], chunks);
}

template <typename T>
__global__ __launch_bounds__(256) void finish_kernel(Many<const T*> partials, int64_t chunks, Many<T*> outs) {
  using N = Num<T>;
  const T* __restrict__ partial = partials.p[blockIdx.z];
  T* __restrict__ out = outs.p[blockIdx.z];
  __shared__ T part[4];
  const int o = blockIdx.x;
  T acc = N::zero();
  for (int64_t c = threadIdx.x; c < chunks; c += 256) acc = N::add(acc, partial[(int64_t)o * chunks + c]);
  acc = N::wave_total(acc);
  if ((threadIdx.x & 63) == 63) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[o] = N::add(N::add(part[0], part[1]), N::add(part[2], part[3]));
}

// Sums of six doubles over the workgroup with one barrier (every thread gets them).
__device__ void block_sum6(double (&v)[6], double (*shared)[6]) {
  for (int q = 0; q < 6; ++q) v[q] = wave_total(v[q]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 63)
    for (int q = 0; q < 6; ++q) shared[wave][q] = v[q];
  __syncthreads();
  for (int q = 0; q < 6; ++q) {
    double t = 0.0;
    for (int u = 0; u < (int)(blockDim.x >> 6); ++u) t += shared[u][q];
    v[q] = t;
  }
}

__device__ __forceinline__ void put(double (&v)[6], int at, double x) { v[at] += x; }
__device__ __forceinline__ void put(double (&v)[6], int at, double2 x) {
  v[at] += x.x;
  v[at + 1] += x.y;
}
__device__ __forceinline__ double take(const double (&v)[6], int at, double) { return v[at]; }
__device__ __forceinline__ double2 take(const double (&v)[6], int at, double2) { return make_double2(v[at], v[at + 1]); }

// Step scalars of step j from the finished reductions red = [s_0..s_{j-1}, |q|^2, sink.q,
// z_0..z_{j-1}, q^H w, |w|^2] and the previous step's scalar block; every workgroup of the
// fused pass computes them (four sums over j numbers, one barrier), so the update rows need
// no coefficient table written by another workgroup.
template <typename T>
struct StepScalars {
  double nu, h0_prev, r, hn, ir, omega, g, ig;
  T c, sv;
};

// Scalars from the sums ss = |s|^2, zz = |z|^2, sz = s^H z over k < j.
template <typename T>
__device__ __forceinline__ StepScalars<T> finish_scalars(double ss, double zz, T sz, T sv, const T* __restrict__ red,
                                                         int j, int fresh, int has_w, const T* __restrict__ sc_prev,
                                                         double breakdown) {
  using N = Num<T>;
  const int nv = j + 1;
  const T* S = red;
  const T* Z = red + nv + 1;
  StepScalars<T> o;
  o.nu = fresh ? 0.0 : N::re(sc_prev[6]);
  o.h0_prev = N::re(sc_prev[5]);
  o.sv = sv;
  o.r = fresh ? 1.0 : sqrt(fmax(N::re(S[j]) - ss, 0.0));
  o.hn = o.nu * o.r;
  o.ir = o.r > 0.0 ? 1.0 / o.r : 0.0;
  if (!fresh && !(o.hn > breakdown)) o.ir = 0.0;  // breakdown: v_j = 0
  o.c = N::zero();
  o.omega = 0.0;
  o.g = 0.0;
  o.ig = 0.0;
  if (has_w) {
    o.c = N::scale(N::sub(Z[j], sz), o.ir);
    o.omega = N::re(Z[nv]);
    o.g = sqrt(fmax(o.omega - zz - N::abs2(o.c), 0.0));
    o.ig = o.g > 0.0 ? 1.0 / o.g : 0.0;
  }
  return o;
}

template <typename T>
__device__ __forceinline__ StepScalars<T> step_scalars(const T* __restrict__ red, const T* __restrict__ sink_v, int j,
                                                       int fresh, int has_w, const T* __restrict__ sc_prev,
                                                       double breakdown, double (*shared)[6]) {
  using N = Num<T>;
  const T* S = red;
  const T* Z = red + j + 2;
  // [|s|^2, |z|^2, s^H z, s . sink_v] (complex sums in two slots; a real basis leaves slots 3, 5 at 0)
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int k = threadIdx.x; k < j; k += blockDim.x) {
    const T s = fresh ? N::zero() : S[k];
    const T z = has_w ? Z[k] : N::zero();
    acc[0] += N::abs2(s);
    acc[1] += N::abs2(z);
    put(acc, 2, N::cfma(s, z, N::zero()));
    if (sink_v != nullptr) put(acc, 4, N::fma(s, sink_v[k], N::zero()));
  }
  block_sum6(acc, shared);
  return finish_scalars<T>(acc[0], acc[1], take(acc, 2, T{}), take(acc, 4, T{}), red, j, fresh, has_w, sc_prev,
                           breakdown);
}

// Hessenberg columns of step j, one wave per row k (WAVES rows per workgroup):
// final H[k, j-1] = H[k, j-1] + nu s_k and tentative H[k, j] = (z_k - (H s)_k)/r with
// (H s)_k = sum_{l >= k-1} H~[k, l] s_l (H~: column j-1 completed), the wave's 64 lanes
// splitting l (all loads of a lane independent).  Workgroup 0 also writes the rows j, j+1,
// sink . v_j and the scalar block of step j:
//   [1/r, 1/g, c/(r g), final H[j, j-1], its h0, h0 of the tentative column j,
//    tentative sub-diagonal nu of column j]
// (h0 = |w|/r, the norm of A v before projection).  `stage` (optional) gets the final column
// j-1 followed by its h0, contiguous for one copy to the host.
template <typename T, int WAVES>
__device__ __forceinline__ void column_part(int blk, T* __restrict__ H, int64_t ldh, int j, int fresh, int has_w,
                                            const T* __restrict__ red, T* __restrict__ sink_v, T* __restrict__ sc_cur,
                                            double* __restrict__ status, T* __restrict__ stage,
                                            const StepScalars<T>& o) {
  using N = Num<T>;
  const int nv = j + 1;
  const T* S = red;
  const T* Z = red + nv + 1;
  T* prev = H + (int64_t)(j - 1) * ldh;  // column j-1 (touched only when !fresh)
  T* col = H + (int64_t)j * ldh;         // column j (written only when has_w)
  const int lane = threadIdx.x & 63;
  const int k = blk * WAVES + (int)(threadIdx.x >> 6);  // uniform per wave
  T hs = N::zero();
  if (!fresh && has_w && k < j) {
    const T sk = S[k];
#pragma unroll 2
    for (int l = max(k - 1, 0) + lane; l < j; l += 64) {
      T h = H[(int64_t)l * ldh + k];
      if (l == j - 1) h = N::fma(N::make(o.nu), sk, h);
      hs = N::fma(h, S[l], hs);
    }
  }
  // The row's only reader is this wave: its reads of column j-1 feed the sum that the
  // writes below wait for, so no barrier is needed.
  hs = N::wave_total(hs);
  if (lane == 63 && k < j) {
    const T s = fresh ? N::zero() : S[k];
    if (!fresh) {
      const T h = N::fma(N::make(o.nu), s, prev[k]);
      prev[k] = h;
      if (stage != nullptr) stage[k] = h;
    }
    if (has_w) col[k] = N::scale(N::sub(Z[k], hs), o.ir);
  }
  if (blk == 0 && threadIdx.x == 0) {
    if (!fresh) {
      prev[j] = N::make(o.hn);
      if (stage != nullptr) {
        stage[j] = N::make(o.hn);
        stage[j + 1] = N::make(o.h0_prev);
      }
    }
    if (has_w) {
      const T hsj = fresh ? N::zero() : N::scale(S[j - 1], o.hn);  // (H s)_j = H[j, j-1] s_{j-1}
      col[j] = N::scale(N::sub(o.c, hsj), o.ir);
      col[j + 1] = N::make(o.g * o.ir);
    }
    if (sink_v != nullptr) sink_v[j] = N::scale(N::sub(S[nv], o.sv), o.ir);
    sc_cur[0] = N::make(o.ir);
    sc_cur[1] = N::make(o.ig);
    sc_cur[2] = N::scale(o.c, o.ir * o.ig);
    sc_cur[3] = N::make(o.hn);
    sc_cur[4] = N::make(o.h0_prev);
    sc_cur[5] = N::make(has_w ? sqrt(fmax(o.omega, 0.0)) * o.ir : 0.0);
    sc_cur[6] = N::make(has_w ? o.g * o.ir : 0.0);
    if (status != nullptr && !fresh) status[0] = o.hn;
  }
}

// 64 RPL rows per workgroup (RPL per lane, 64 apart), WAVES waves splitting the columns,
// partial sums met in LDS:
// with A' = V s and Z' = V z over columns 0..j-1,
//   v_j = q/r - A'/r,  q_{j+1} = w/g - (c/(r g)) q - (Z' - (c/r) A')/g.
// The scalars r, c, g need |s|^2, |z|^2 and s^H z: each wave sums them over its own
// columns while it streams them (the coefficients are wave-uniform), and they meet in LDS
// with the row sums, so the pass starts its loads at once and has one barrier.  After it,
// wave w finishes the rows of slots t = w, w + WAVES, ... of every lane.
template <typename T, typename S, int WAVES, int RPL>
__device__ __forceinline__ void update_part(int blk, S* __restrict__ V, int64_t ld, int j, int fresh,
                                            const T* __restrict__ w, const T* __restrict__ red, int64_t dim,
                                            const T* __restrict__ sc_prev, double breakdown, T* __restrict__ pad,
                                            const int64_t* __restrict__ pad_map) {
  using N = Num<T>;
  __shared__ T pa[WAVES][64 * RPL];
  __shared__ T pb[WAVES][64 * RPL];
  __shared__ double ps[WAVES][4];
  const T* Sv = red;
  const T* Zv = red + j + 2;
  const int has_w = w != nullptr ? 1 : 0;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t base = blk * (int64_t)(64 * RPL) + lane;
  int64_t ic[RPL];
#pragma unroll
  for (int t = 0; t < RPL; ++t) {
    const int64_t i = base + 64 * t;
    ic[t] = i < dim ? i : dim - 1;  // rows past the end read the last row, discarded
  }
  T A[RPL], B[RPL];
#pragma unroll
  for (int t = 0; t < RPL; ++t) A[t] = B[t] = N::zero();
  T sz = N::zero();
  double ss = 0.0, zz = 0.0;
#pragma unroll 4
  for (int k = wave; k < j; k += WAVES) {
    const S* __restrict__ col = V + (int64_t)k * ld;
    const T sk = fresh ? N::zero() : Sv[k];
    const T zk = has_w ? Zv[k] : N::zero();
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
      const T v = widen(col[ic[t]]);
      A[t] = N::fma(sk, v, A[t]);
      B[t] = N::fma(zk, v, B[t]);
    }
    ss += N::abs2(sk);
    zz += N::abs2(zk);
    sz = N::cfma(sk, zk, sz);
  }
#pragma unroll
  for (int t = 0; t < RPL; ++t) {
    pa[wave][t * 64 + lane] = A[t];
    pb[wave][t * 64 + lane] = B[t];
  }
  if (lane == 0) {
    double tt[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    put(tt, 2, sz);
    ps[wave][0] = ss;
    ps[wave][1] = zz;
    ps[wave][2] = tt[2];
    ps[wave][3] = tt[3];
  }
  __syncthreads();
  if (wave >= RPL) return;  // no row slot left for this wave (RPL < WAVES)
  double tss = 0.0, tzz = 0.0, tt[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int u = 0; u < WAVES; ++u) {
    tss += ps[u][0];
    tzz += ps[u][1];
    tt[2] += ps[u][2];
    tt[3] += ps[u][3];
  }
  const StepScalars<T> o =
      finish_scalars<T>(tss, tzz, take(tt, 2, T{}), N::zero(), red, j, fresh, has_w, sc_prev, breakdown);
  for (int t = wave; t < RPL; t += WAVES) {
    const int64_t i = base + 64 * t;
    if (i >= dim) continue;
    T a = N::zero(), bsum = N::zero();
    for (int u = 0; u < WAVES; ++u) {
      a = N::add(a, pa[u][t * 64 + lane]);
      bsum = N::add(bsum, pb[u][t * 64 + lane]);
    }
    S* q = V + (int64_t)j * ld;
    const T qi = widen(q[i]);
    if (!fresh) q[i] = narrow_to<S>(N::sub(N::scale(qi, o.ir), N::scale(a, o.ir)));
    if (has_w) {
      const T ci = N::scale(o.c, o.ir);
      const T b = N::scale(N::sub(bsum, N::mul(ci, a)), o.ig);
      const T next = N::sub(N::sub(N::scale(w[i], o.ig), N::mul(N::scale(ci, o.ig), qi)), b);
      V[(int64_t)(j + 1) * ld + i] = narrow_to<S>(next);
      if (pad != nullptr) {  // the next operator input, scattered into its padded blocks
        const T stored = widen(narrow_to<S>(next));
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          const int64_t m = pad_map[2 * i + k];
          if (m >= 0) pad[m >> 1] = (m & 1) ? N::sub(N::zero(), stored) : stored;
        }
      }
    }
  }
}

// The step's second pass: workgroups [0, col_blocks) write the Hessenberg columns and the
// scalar block, the others update 64 basis rows each.  The scalar block alternates with the
// step parity (sc_prev: step j-1, only read; sc_cur: step j, written by workgroup 0).
template <typename T, typename S, int WAVES, int RPL>
__global__ __launch_bounds__(64 * WAVES) void update_column_kernel(Many<S*> V, int64_t ld, int j, int fresh,
                                                                   Many<const T*> w, Many<T*> H, int64_t ldh,
                                                                   Many<const T*> red, Many<T*> sink_v, Many<T*> sc,
                                                                   Many<double*> status, Many<T*> stage,
                                                                   double breakdown, int64_t dim, int col_blocks,
                                                                   Many<T*> pad, Many<const int64_t*> pad_map) {
  const int i = blockIdx.z;
  const int has_w = w.p[i] != nullptr ? 1 : 0;
  T* sc_prev = sc.p[i] + 8 * ((j + 1) & 1);
  T* sc_cur = sc.p[i] + 8 * (j & 1);
  if ((int)blockIdx.x < col_blocks) {
    __shared__ double shared[WAVES][6];
    const StepScalars<T> o = step_scalars<T>(red.p[i], sink_v.p[i], j, fresh, has_w, sc_prev, breakdown, shared);
    column_part<T, WAVES>((int)blockIdx.x, H.p[i], ldh, j, fresh, has_w, red.p[i], sink_v.p[i], sc_cur, status.p[i],
                          stage.p[i], o);
  } else {
    update_part<T, S, WAVES, RPL>((int)blockIdx.x - col_blocks, V.p[i], ld, j, fresh, w.p[i], red.p[i], dim, sc_prev,
                             breakdown, pad.p[i], pad_map.p[i]);
  }
}

int fail(const std::string& msg) { return c3l_internal_fail(C3L_EINVAL, msg.c_str()); }

int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e == hipSuccess) return C3L_OK;
  return c3l_internal_fail(C3L_EHIP, (std::string(what) + ": " + hipGetErrorString(e)).c_str());
}

}  // namespace

int64_t partial_entries(bool complex_basis, int64_t dim, int32_t max_columns) {
  const int rows = complex_basis ? Shape<double2>::kRows : Shape<double>::kRows;
  return 2 * ((int64_t)max_columns + 2) * ((dim + rows - 1) / rows);
}


template <typename T, typename S>
int step(const Basis<T, S>& b, int j, bool fresh, const T* w, double breakdown, hipStream_t s, Clock* clock) {
  return step_many(&b, 1, j, fresh, &w, breakdown, s, clock);
}

template <typename T, typename S>
int step_many(const Basis<T, S>* bs, int count, int j, bool fresh, const T* const* ws, double breakdown,
              hipStream_t s, Clock* clock) {
  if (bs == nullptr || ws == nullptr || count < 1 || count > kMaxInstances || j < 0 || (j == 0 && !fresh))
    return fail("Krylov orthogonalisation step: invalid argument");
  const Basis<T, S>& b = bs[0];
  const bool has_w = ws[0] != nullptr;
  for (int i = 0; i < count; ++i) {
    const Basis<T, S>& o = bs[i];
    if (o.V == nullptr || o.H == nullptr || o.partial == nullptr || o.red == nullptr ||
        o.sc == nullptr || o.dim <= 0 || o.ld < o.dim || (has_w ? j + 2 > o.ldh : j + 1 > o.ldh) ||
        (ws[i] != nullptr) != has_w || o.dim != b.dim || o.ld != b.ld || o.ldh != b.ldh)
      return fail("Krylov orthogonalisation step: invalid argument");
  }
  constexpr int ROWS = Shape<T>::kRows;
  constexpr int WAVES = Shape<T>::kWaves;
  Many<const S*> Vr{};
  Many<S*> Vw{};
  Many<const T*> w{}, sink{}, partial_r{}, red_r{};
  Many<T*> partial{}, red{}, H{}, sink_v{}, sc{}, stage{};
  Many<double*> status{};
  Many<T*> pad{};
  Many<const int64_t*> pad_map{};
  for (int i = 0; i < count; ++i) {
    const Basis<T, S>& o = bs[i];
    if ((o.pad == nullptr) != (o.pad_map == nullptr) || (o.pad != nullptr) != (bs[0].pad != nullptr))
      return fail("Krylov orthogonalisation step: invalid argument");
    pad.p[i] = o.pad;
    pad_map.p[i] = o.pad_map;
    Vr.p[i] = o.V;
    Vw.p[i] = o.V;
    w.p[i] = ws[i];
    sink.p[i] = o.sink;
    partial.p[i] = o.partial;
    partial_r.p[i] = o.partial;
    red.p[i] = o.red;
    red_r.p[i] = o.red;
    H.p[i] = o.H;
    sink_v.p[i] = o.sink_v;
    sc.p[i] = o.sc;
    stage.p[i] = o.stage;
    status.p[i] = o.status;
  }
  const int nv = j + 1;
  const int64_t chunks = (b.dim + ROWS - 1) / ROWS;
  const int cols = nv + 1;
  const int cpw = (int)std::min<int64_t>(8, std::max<int64_t>(1, (int64_t)cols * chunks / 4096));
  const int groups = (cols + 4 * cpw - 1) / (4 * cpw);
  const double vec = (double)sizeof(T) * (double)b.dim;  // w, sink
  const double col = (double)sizeof(S) * (double)b.dim;  // basis columns
  const unsigned z = (unsigned)count;
  if (clock) clock->begin(0, s);
  constexpr int CPI = Shape<T>::kColsPerIter;
  hipLaunchKernelGGL((reduce_kernel<T, S, ROWS, CPI>), dim3(chunks, groups, z), dim3(256), 0, s, Vr, b.ld, nv, cpw, w, sink,
                     b.dim, partial, chunks);
  if (int rc = check_launch("reduce_kernel")) return rc;
  if (clock) clock->end(0, count * (col * nv + vec * ((has_w ? 1 : 0) + (b.sink ? 1 : 0))), s);
  const int outs = has_w ? 2 * (nv + 1) : nv + 1;
  hipLaunchKernelGGL((finish_kernel<T>), dim3(outs, 1, z), dim3(256), 0, s, partial_r, chunks, red);
  if (int rc = check_launch("finish_kernel")) return rc;
  const int col_blocks = std::max(1, (j + WAVES - 1) / WAVES);
  if (clock) clock->begin(1, s);
  constexpr int RPL = Shape<T>::kRowsPerLane;
  hipLaunchKernelGGL((update_column_kernel<T, S, WAVES, RPL>), dim3(col_blocks + (b.dim + 64 * RPL - 1) / (64 * RPL), 1, z),
                     dim3(64 * WAVES), 0, s, Vw, b.ld, j, fresh ? 1 : 0, w, H, b.ldh, red_r, sink_v, sc, status, stage,
                     breakdown, b.dim, col_blocks, pad, pad_map);
  if (int rc = check_launch("update_column_kernel")) return rc;
  if (clock) clock->end(1, count * (col * (j + 1 + (has_w ? 1 : 0) + (fresh ? 0 : 1)) + vec * (has_w ? 1 : 0)), s);
  return C3L_OK;
}

#define C3L_ORTHO_INSTANTIATE(T, S)                                                                            \
  template int step<T, S>(const Basis<T, S>&, int, bool, const T*, double, hipStream_t, Clock*);             \
  template int step_many<T, S>(const Basis<T, S>*, int, int, bool, const T* const*, double, hipStream_t, Clock*);
C3L_ORTHO_INSTANTIATE(double, double)
C3L_ORTHO_INSTANTIATE(double, float)
C3L_ORTHO_INSTANTIATE(double2, double2)
#undef C3L_ORTHO_INSTANTIATE

}  // namespace c3l_ortho
