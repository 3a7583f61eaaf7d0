// MI355X (gfx950) kernels and C ABI for the C3L coupled-cluster hot path.
//
// FP64 path, with an FP32 (float2) mirror of each plan workspace for the mixed-precision
// JVP.  Data layout in HBM:
//   * Newton/GMRES vectors: real tilde coordinates, one contiguous double array.
//   * Amplitudes / EOM vectors: complex (double2) in the reference's canonical
//     string order (PH rank, then mask value).
//   * Plan workspace: one double2 array holding every padded atom tensor
//     (n^rank entries, first index fastest, antisymmetric within a group) and the
//     padded base blocks scattered from the current input vector.  Twin plans (built
//     from the same arrays) each own a workspace and run in one launch per kernel.
//   * Krylov bases: basis vectors stored contiguously, column k at V + k*ld.
//
// Every kernel is bandwidth/latency bound (GEMV-like contractions, gathers,
// outer-product style wedges); none is GEMM shaped at FP64, so no MFMA here.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/c3l_hip.h"

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

#define C3L_HIP_CHECK(expr)                                                          \
  do {                                                                                \
    hipError_t err__ = (expr);                                                        \
    if (err__ != hipSuccess) {                                                        \
      return fail(C3L_EHIP, std::string(#expr) + ": " + hipGetErrorString(err__));    \
    }                                                                                 \
  } while (0)

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }

// Storage/arithmetic type conversions of the plan kernels (FP64 = double2, FP32 = float2).
template <typename C>
__device__ __forceinline__ C from_d2(double2 v);
template <>
__device__ __forceinline__ double2 from_d2<double2>(double2 v) {
  return v;
}
template <>
__device__ __forceinline__ float2 from_d2<float2>(double2 v) {
  return make_float2((float)v.x, (float)v.y);
}
__device__ __forceinline__ double2 to_d2(double2 v) { return v; }
__device__ __forceinline__ double2 to_d2(float2 v) { return make_double2((double)v.x, (double)v.y); }
template <typename C>
__device__ __forceinline__ C czero();
template <>
__device__ __forceinline__ double2 czero<double2>() {
  return make_double2(0.0, 0.0);
}
template <>
__device__ __forceinline__ float2 czero<float2>() {
  return make_float2(0.f, 0.f);
}

constexpr int kThreads = 256;

int64_t grid_for(int64_t count, int threads = kThreads) {
  int64_t blocks = (count + threads - 1) / threads;
  if (blocks < 1) blocks = 1;
  if (blocks > 1 << 20) blocks = 1 << 20;
  return blocks;
}

// ------------------------------------------------------------------ plan kernels
//
// Every plan kernel runs over blockIdx.z = instance: twin plans (compiled from one
// lowered plan, so identical layouts, descriptors and index arrays) evaluate their
// own workspaces and vectors in one launch.  A single plan is the one-instance case.
constexpr int kMaxTwins = 4;
template <typename P>
struct Twins {
  P p[kMaxTwins];
};
template <typename P>
Twins<P> twin(P a) {
  Twins<P> t{};
  t.p[0] = a;
  return t;
}

template <typename C>
__global__ void scatter_padded_kernel(Twins<const double2*> vs, Twins<C*> pads, const int64_t* __restrict__ src,
                                      const int8_t* __restrict__ sign, int64_t count) {
  const double2* __restrict__ v = vs.p[blockIdx.z];
  C* __restrict__ pad = pads.p[blockIdx.z];
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < count;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = src[e];
    double2 x = make_double2(0.0, 0.0);
    if (s >= 0) {
      x = v[s];
      if (sign[e] < 0) {
        x.x = -x.x;
        x.y = -x.y;
      }
    }
    pad[e] = from_d2<C>(x);
  }
}

__global__ void narrow_kernel(const double2* __restrict__ a, float2* __restrict__ b, int64_t count) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = from_d2<float2>(a[i]);
}

// out[i] = in[i] (FP64 -> FP32) over the listed workspace ranges; grid.y = range.
__global__ void narrow_ranges_kernel(const double2* __restrict__ ws, float2* __restrict__ ws32,
                                     const int64_t* __restrict__ ranges) {
  const int64_t begin = ranges[2 * blockIdx.y], count = ranges[2 * blockIdx.y + 1];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    ws32[begin + i] = from_d2<float2>(ws[begin + i]);
}

struct ContractJob {
  int64_t out_offset;
  int64_t par_offset;
  int32_t par_rank;
  int32_t slot;
  int32_t vec;
  int32_t sign;
};

// sum_q vec[q] par[q*inner] in order q = 0..n-1; the loads are issued eight at a
// time so that one thread has eight in flight instead of one.
template <typename C>
__device__ __forceinline__ C contract_one(const C* __restrict__ par, const C* __restrict__ vec, int64_t inner, int n) {
  C acc = czero<C>();
  int q = 0;
  for (; q + 8 <= n; q += 8) {
    C p[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) p[u] = par[(int64_t)(q + u) * inner];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc = cadd(acc, cmul(vec[q + u], p[u]));
  }
  for (; q < n; ++q) acc = cadd(acc, cmul(vec[q], par[(int64_t)q * inner]));
  return acc;
}

// Sum over aligned groups of 8 lanes with DPP moves, valid in the last lane of each
// group: an inclusive scan inside rows of 16 (row_shr 1, 2, 4, zero filled at the row start).
template <int CTRL>
__device__ __forceinline__ double dpp_shr(double v) {
  const int2 b = __builtin_bit_cast(int2, v);
  const int lo = __builtin_amdgcn_update_dpp(0, b.x, CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, b.y, CTRL, 0xf, 0xf, true);
  return __builtin_bit_cast(double, make_int2(lo, hi));
}
template <int CTRL>
__device__ __forceinline__ float dpp_shr(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xf, 0xf, true));
}
template <typename R>
__device__ __forceinline__ R group8_total(R v) {
  v += dpp_shr<0x111>(v);  // row_shr:1
  v += dpp_shr<0x112>(v);  // row_shr:2
  v += dpp_shr<0x114>(v);  // row_shr:4
  return v;
}
__device__ __forceinline__ double2 group8_total(double2 v) { return make_double2(group8_total(v.x), group8_total(v.y)); }
__device__ __forceinline__ float2 group8_total(float2 v) { return make_float2(group8_total(v.x), group8_total(v.y)); }

// Lanes per output of a contraction job: a job over the stride-1 slot (slot 0) reads n
// consecutive elements per output, so eight lanes share one output (lane l takes the
// terms q = l, l+8, ...) and the wave reads eight runs of consecutive elements instead of
// 64 strided ones; other slots are read coalesced with one output per thread.
__host__ __device__ __forceinline__ int contract_lanes(int slot, int n) { return slot == 0 && n >= 8 ? 8 : 1; }

// grid.y = job; threads cover the n^(R-1) outputs of that job (eight threads per output
// for slot-0 jobs).
template <typename C>
__global__ void contract_kernel(Twins<C*> wss, const C* __restrict__ vectors, const ContractJob* __restrict__ jobs,
                                int n) {
  C* __restrict__ ws = wss.p[blockIdx.z];
  const ContractJob job = jobs[blockIdx.y];
  const int r = job.par_rank - 1;
  int64_t out_size = 1;
  for (int k = 0; k < r; ++k) out_size *= n;
  int64_t inner = 1;
  for (int k = 0; k < job.slot; ++k) inner *= n;  // stride of the contracted slot
  const C* __restrict__ par = ws + job.par_offset;
  const C* __restrict__ vec = vectors + (int64_t)job.vec * n;
  C* __restrict__ out = ws + job.out_offset;
  if (contract_lanes(job.slot, n) == 8) {  // uniform over the launch's job
    // every lane of a group runs the same iterations (groups are whole: 8 | 8 out_size)
    for (int64_t T = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; T < out_size * 8;
         T += (int64_t)gridDim.x * blockDim.x) {
      const int64_t J = T >> 3;
      const int l = (int)(T & 7);
      const C* __restrict__ p = par + J * n;
      C acc = czero<C>();
      for (int q = l; q < n; q += 8) acc = cadd(acc, cmul(vec[q], p[q]));
      acc = group8_total(acc);
      if (l == 7) {
        if (job.sign < 0) {
          acc.x = -acc.x;
          acc.y = -acc.y;
        }
        out[J] = acc;
      }
    }
    return;
  }
  for (int64_t J = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; J < out_size;
       J += (int64_t)gridDim.x * blockDim.x) {
    const int64_t hi = J / inner;
    const int64_t lo = J - hi * inner;
    const int64_t base = hi * n * inner + lo;
    C acc = contract_one(par + base, vec, inner, n);
    if (job.sign < 0) {
      acc.x = -acc.x;
      acc.y = -acc.y;
    }
    out[J] = acc;
  }
}

// Contraction trees, one workgroup each (grid.x = tree): a tree is a root job (whose
// parent atom exists before the launch) followed by every job that depends on it,
// ordered by level; barrier[t] marks the first job of a new level, whose parents the
// workgroup itself has just written.  Trees are independent, so one launch replaces
// one launch per level for the small levels that follow the large first ones.
template <typename C>
__global__ __launch_bounds__(256) void contract_trees_kernel(Twins<C*> wss, const C* __restrict__ vectors,
                                                             const ContractJob* __restrict__ jobs,
                                                             const int32_t* __restrict__ tree_start,
                                                             const uint8_t* __restrict__ barrier, int n) {
  C* __restrict__ ws = wss.p[blockIdx.z];
  for (int t = tree_start[blockIdx.x]; t < tree_start[blockIdx.x + 1]; ++t) {
    if (barrier[t]) __syncthreads();  // the parents written by this workgroup are complete
    const ContractJob job = jobs[t];
    const int r = job.par_rank - 1;
    int64_t out_size = 1;
    for (int k = 0; k < r; ++k) out_size *= n;
    int64_t inner = 1;
    for (int k = 0; k < job.slot; ++k) inner *= n;  // stride of the contracted slot
    const C* __restrict__ par = ws + job.par_offset;
    const C* __restrict__ vec = vectors + (int64_t)job.vec * n;
    C* __restrict__ out = ws + job.out_offset;
    for (int64_t J = threadIdx.x; J < out_size; J += blockDim.x) {
      const int64_t hi = J / inner;
      const int64_t lo = J - hi * inner;
      const int64_t base = hi * n * inner + lo;
      C acc = contract_one(par + base, vec, inner, n);
      if (job.sign < 0) {
        acc.x = -acc.x;
        acc.y = -acc.y;
      }
      out[J] = acc;
    }
  }
}

// ---------------------------------------------------------------- program kernel
//
// One launch evaluates every output block of a program.  Work is cut into
// tiles; a tile belongs to one block and covers `256 >> lanes_log2`
// consecutive canonical coordinates of it.  With lanes_log2 = L > 0 the 2^L
// waves of the workgroup split that block's entries between them (small
// blocks: more workgroups in flight) and their partial sums meet in LDS.
// Because a wave never straddles two entry groups, the entry index and all
// entry/factor metadata are wave-uniform and come through scalar loads.
//
// Factor f of an entry reads ws[offset_f + sum_k idx_k * stride_f[k]]: the
// padded-tensor flat index as a dot product of the coordinate with strides
// prepared on the host (0 for positions the factor does not use).
// Coordinates are packed as four 8-bit indices (n <= 255) up to rank four and as five 6-bit
// indices (n <= 63) at rank five (the S135 response shell).

struct WedgeFactor {
  int64_t offset;
  int32_t stride[5];
};

struct WedgeEntry {
  double2 coeff;
  int32_t first;  // first factor
  int32_t nf;     // factor count
  int32_t pad[2];
};

struct WedgeBlockDesc {
  int64_t count;
  const uint32_t* coord;  // packed indices, one word per coordinate (null: dense block)
  const int64_t* dest;    // output positions (null: dense block)
  const WedgeEntry* entries;
  const WedgeFactor* factors;
  int32_t n_entries;
  int32_t lanes_log2;
  int32_t reps;     // coordinates per thread (two-factor kernel, lanes_log2 == 0 only)
  int32_t coord_bits;  // bits per index of the packed coordinates: 8 up to rank four, 6 at rank five
  int32_t dense_n;  // dense block: coordinate e is the index tuple of flat index e (first index
                    // fastest, base dense_n) and writes position dense_base + e
  int64_t dense_base;
};

constexpr int kMaxReps = 2;
constexpr int kStagedEntries = 64;

template <bool kFive>
__device__ __forceinline__ int64_t factor_index(const WedgeFactor& f, int i0, int i1, int i2, int i3, int i4) {
  int flat = i0 * f.stride[0] + i1 * f.stride[1] + i2 * f.stride[2] + i3 * f.stride[3];
  if constexpr (kFive) flat += i4 * f.stride[4];
  return f.offset + (int64_t)flat;
}

// kFive: the program has rank-five output blocks (a fifth index per coordinate); the other
// programs use the four-index form, which needs fewer registers.
template <bool kTwoFactors, typename C, bool kFive>
__global__ __launch_bounds__(256) void wedge_program_kernel(Twins<const C*> wss, Twins<double2*> outs,
                                                            const WedgeBlockDesc* __restrict__ descs,
                                                            const int2* __restrict__ tiles,
                                                            const double2* __restrict__ diag,
                                                            Twins<const double2*> xs) {
  const C* __restrict__ ws = wss.p[blockIdx.z];
  double2* __restrict__ out = outs.p[blockIdx.z];
  const double2* __restrict__ x = xs.p[blockIdx.z];
  __shared__ double2 part[256];
  __shared__ double2 s_coeff[kStagedEntries];
  __shared__ WedgeFactor s_fac[2 * kStagedEntries];
  const int2 tile = tiles[blockIdx.x];
  const WedgeBlockDesc d = descs[tile.x];
  // two-factor entry metadata is staged in LDS once per workgroup (uniform reads later)
  const bool staged = kTwoFactors && d.n_entries <= kStagedEntries;
  if (staged) {
    for (int q = threadIdx.x; q < d.n_entries; q += blockDim.x) s_coeff[q] = d.entries[q].coeff;
    for (int q = threadIdx.x; q < 2 * d.n_entries; q += blockDim.x) s_fac[q] = d.factors[q];
  }
  __syncthreads();
  const int lg = d.lanes_log2;
  const int per = 256 >> lg;
  const int local = threadIdx.x & (per - 1);
  const int group = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> (8 - lg)));
  const int step = 1 << lg;
  constexpr int R = kTwoFactors ? kMaxReps : 1;
  int i0[R], i1[R], i2[R], i3[R], i4[R];
  bool act[R];
  int64_t e[R];
  C acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    e[r] = (int64_t)tile.y + local + (int64_t)r * per;
    act[r] = r < d.reps && e[r] < d.count;
    i0[r] = i1[r] = i2[r] = i3[r] = i4[r] = 0;
    if (act[r]) {
      if (d.coord != nullptr) {
        const uint32_t c = d.coord[e[r]];
        const int bits = d.coord_bits;
        const uint32_t m = (1u << bits) - 1u;
        i0[r] = c & m;
        i1[r] = (c >> bits) & m;
        i2[r] = (c >> (2 * bits)) & m;
        i3[r] = (c >> (3 * bits)) & m;
        if constexpr (kFive) i4[r] = bits < 8 ? (c >> (4 * bits)) & m : 0u;  // rank-five blocks pack a fifth index
      } else {  // dense: e < dense_n^deg, so the unused higher indices come out 0
        const int n = d.dense_n;
        int f = (int)e[r];
        i0[r] = f % n;
        f /= n;
        i1[r] = f % n;
        f /= n;
        i2[r] = f % n;
        f /= n;
        if constexpr (kFive) {
          i3[r] = f % n;
          i4[r] = f / n;
        } else {
          i3[r] = f;
        }
      }
    }
    acc[r] = czero<C>();
  }
  if (kTwoFactors) {
    // Branch-free body: inactive coordinates read index 0 (a valid address) and are
    // never stored, so the gathers of every coordinate and unrolled entry overlap.
    if (staged) {
#pragma unroll 2
      for (int t = group; t < d.n_entries; t += step) {
        const C c = from_d2<C>(s_coeff[t]);
        const WedgeFactor fa = s_fac[2 * t];
        const WedgeFactor fb = s_fac[2 * t + 1];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const C a = ws[factor_index<kFive>(fa, i0[r], i1[r], i2[r], i3[r], i4[r])];
          const C b = ws[factor_index<kFive>(fb, i0[r], i1[r], i2[r], i3[r], i4[r])];
          acc[r] = cadd(acc[r], cmul(c, cmul(a, b)));
        }
      }
    } else {
#pragma unroll 2
      for (int t = group; t < d.n_entries; t += step) {
        const C c = from_d2<C>(d.entries[t].coeff);
        const WedgeFactor fa = d.factors[2 * t];
        const WedgeFactor fb = d.factors[2 * t + 1];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const C a = ws[factor_index<kFive>(fa, i0[r], i1[r], i2[r], i3[r], i4[r])];
          const C b = ws[factor_index<kFive>(fb, i0[r], i1[r], i2[r], i3[r], i4[r])];
          acc[r] = cadd(acc[r], cmul(c, cmul(a, b)));
        }
      }
    }
  } else if (act[0]) {
    for (int t = group; t < d.n_entries; t += step) {
      const WedgeEntry en = d.entries[t];
      C prod = from_d2<C>(en.coeff);
      for (int f = en.first; f < en.first + en.nf; ++f)
        prod = cmul(prod, ws[factor_index<kFive>(d.factors[f], i0[0], i1[0], i2[0], i3[0], i4[0])]);
      acc[0] = cadd(acc[0], prod);
    }
  }
  if (lg > 0) {  // reps == 1 here
    part[threadIdx.x] = to_d2(acc[0]);
    __syncthreads();
    if (group == 0)
      for (int k = 1; k < step; ++k) acc[0] = cadd(acc[0], from_d2<C>(part[k * per + local]));
  }
  if (group == 0) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (act[r]) {
        const int64_t dd = d.dest != nullptr ? d.dest[e[r]] : d.dense_base + e[r];
        double2 v = to_d2(acc[r]);
        if (diag != nullptr) v = cadd(v, cmul(diag[dd], x[dd]));
        out[dd] = v;
      }
    }
  }
}

// ------------------------------------------------------- matrix-core tile program
//
// FP32 programs: rank-4 output blocks whose entries are products of two atoms that split the
// four positions between them (or of a rank-4 atom and a scalar atom) run on the matrix cores
// instead of the per-coordinate gathers.  A workgroup owns a 16x4x4x4 tile of the padded
// output (16 along position 0).  An entry A * B is taken with A the atom that covers position
// 0 and c the lowest position of B; (pa, pb) are the other two positions.  Wave w owns the
// slab pb = w of the tile and evaluates the entry as one v_mfma_f32_4x4x1_16b_f32 outer
// product per 16 blocks: block b = (position 0 / 4, pa), rows = position 0 % 4, columns = c,
// D[b][i][j] += (coeff A)[b][i] B[b][j], so 16 consecutive lanes read 16 consecutive
// position-0 elements of A (whole 128-byte lines) and no product is wasted.  Complex
// products use the 3M split (Higham): with A' = coeff A, T1 = Re A' Re B, T2 = Im A' Im B,
// T3 = (Re A' + Im A')(Re B + Im B), Re = T1 - T2, Im = T3 - T1 - T2, three real MFMAs (a real
// coefficient, the only kind the CCSD JVP has, scales A by two multiplies instead of four).
// The three choices of c are the three views.  Operand lanes past the tensor read on into
// the next atom or the zeroed tail of the workspace mirror; they feed only outputs outside
// the tensor, which are never written.  The views meet in LDS, and the tile's outputs go out
// 16 consecutive destinations per 16 lanes, with the rank-4 x scalar entries and the diagonal.
// MFMA 4x4x1 (16 blocks) operand layout: lane l holds A[l/4][l%4] (block l/4, row l%4) and
// B[l/4][l%4] (column l%4); accumulator register i of lane l holds D[l/4][i][l%4].

typedef float v4f_t __attribute__((ext_vector_type(4)));

constexpr int kTileViews = 3;
constexpr int kTileMaxRules = 32;
// Below this many entries per block the per-coordinate gathers are as fast (measured: 44
// entries at n = 25, 38 us on the tiles against 36 us gathered; 516 entries at n = 39, 0.87 ms
// against 3.8 ms).
constexpr int kTileMinEntries = 128;
constexpr int kBufferWord3 = 0x00020000;  // raw buffer descriptor word 3 on gfx9-family parts

// Elements past the FP32 workspace that tile operand lanes can reach: at most 15 indices past
// the tensor at position 0 and 3 at the others, times strides of at most n^3 each.
inline int64_t tile_tail_elements(int n) { return 32 * (int64_t)n * n * n + 64; }

// view v: c = v + 1, (pa, pb) the other two of positions 1..3 in increasing order
__host__ __device__ constexpr int view_pa(int v) { return v == 0 ? 2 : 1; }
__host__ __device__ constexpr int view_pb(int v) { return v == 2 ? 2 : 3; }

struct TileRule {  // one operand layout: lane l indexes position pos[0] by l % 16 (A) or l % 4
                   // (B), pos[1] by l / 16; wave w indexes pos[2] by w
  int32_t pos[3];
  uint32_t stride[3];  // bytes per unit of the three positions (0 where the atom does not cover one)
  int32_t b_side;      // 1 for B (l % 4)
};

struct TileEntry {  // coeff * A * B in one view of the FP32 workspace (16 bytes: one scalar load)
  uint32_t a_off, b_off;  // atom offsets (bytes)
  uint32_t lds;           // byte offsets of the A (bits 0-15) and B (bits 16-31) layouts in a wave's table
  float coeff;            // real part; the imaginary parts are a separate table
};

struct TileScalarEntry {  // coeff * A[p0, p1, p2, p3] * s, element offsets
  float2 coeff;
  uint32_t a_off;
  int32_t a_stride[4];
  uint32_t scalar;
};

struct TileBlockDesc {
  int32_t n;
  int32_t per_dim;                    // 4-wide tiles per position
  int32_t view_start[kTileViews + 1];  // entries of view v: [view_start[v], view_start[v + 1])
  int32_t rule_start, rule_end;       // operand layouts of the block
  int32_t scalar_start, scalar_end;   // rank-4 x scalar entries
  const int32_t* dest;                // n^4 (first index fastest): destination of a tuple, -1 if none
};

// Entry tables are kernel arguments so that their wave-uniform reads become scalar loads.
// tiles[blockIdx.x] = (block, g): g = g0 + G0 (t1 + T (t2 + T t3)) with G0 = ceil(n / 16).
template <bool kReal>
__global__ __launch_bounds__(256) void tile_program_mfma_kernel(const float2* __restrict__ ws,
                                                                double2* __restrict__ out,
                                                                const TileBlockDesc* __restrict__ descs,
                                                                const TileRule* __restrict__ rules,
                                                                const TileEntry* __restrict__ entries,
                                                                const float* __restrict__ coeff_im,
                                                                const TileScalarEntry* __restrict__ scalars,
                                                                const int2* __restrict__ tiles,
                                                                const double2* __restrict__ diag,
                                                                const double2* __restrict__ x) {
  __shared__ float2 part[1024];  // the workgroup's 16x4x4x4 outputs, position 0 fastest
  // per wave, layout and lane: byte offset of the lane's operand element from the atom origin
  __shared__ uint32_t lane_off[4][kTileMaxRules][64];
  const int2 tile = tiles[blockIdx.x];
  const TileBlockDesc& d = descs[tile.x];
  const int n = d.n, T = d.per_dim, G0 = (n + 15) / 16;
  const int l = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int P0 = 16 * (tile.y % G0), P1 = 4 * ((tile.y / G0) % T), P2 = 4 * ((tile.y / (G0 * T)) % T),
            P3 = 4 * (tile.y / (G0 * T * T));
  auto origin_of = [&](int k) {  // branch-free selection of the tile origin at position k
    return P0 * (k == 0) + P1 * (k == 1) + P2 * (k == 2) + P3 * (k == 3);
  };
  const uint32_t u15 = l & 15, u3 = l & 3, u4 = l >> 4;
  // the wave's operand offsets for every layout of the block (written and read by this wave
  // only; LDS executes a wave's accesses in order)
  for (int r = d.rule_start; r < d.rule_end; ++r) {
    const TileRule R = rules[r];
    lane_off[wave][r - d.rule_start][l] = (origin_of(R.pos[0]) + (R.b_side ? u3 : u15)) * R.stride[0] +
                                          (origin_of(R.pos[1]) + u4) * R.stride[1] +
                                          (origin_of(R.pos[2]) + wave) * R.stride[2];
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const char* offs = reinterpret_cast<const char*>(&lane_off[wave][0][l]);
  // raw buffer loads: the atom offset rides in the scalar offset, the lane offset in the vector
  // offset; every address is inside the mirror and its tail by construction
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)ws, (short)0, -1, kBufferWord3);
#pragma unroll
  for (int v = 0; v < kTileViews; ++v) {
    v4f_t acc[3] = {v4f_t{0.f, 0.f, 0.f, 0.f}, v4f_t{0.f, 0.f, 0.f, 0.f}, v4f_t{0.f, 0.f, 0.f, 0.f}};
    if (origin_of(view_pb(v)) + wave < n) {  // wave-uniform: a slab past the tensor has no outputs
      auto operands = [&](const TileEntry& E, float2& a, float2& b) {
        const uint32_t ao = *reinterpret_cast<const uint32_t*>(offs + (E.lds & 0xffffu));
        const uint32_t bo = *reinterpret_cast<const uint32_t*>(offs + (E.lds >> 16));
        a = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rsrc, (int)ao, (int)E.a_off, 0));
        b = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rsrc, (int)bo, (int)E.b_off, 0));
      };
      auto mac = [&](const float2 c, const float2 a, const float2 b) {  // c: coefficient
        float ar = c.x * a.x, ai = c.x * a.y;
        if constexpr (!kReal) {
          ar -= c.y * a.y;
          ai += c.y * a.x;
        }
        acc[0] = __builtin_amdgcn_mfma_f32_4x4x1f32(ar, b.x, acc[0], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f32_4x4x1f32(ai, b.y, acc[1], 0, 0, 0);
        acc[2] = __builtin_amdgcn_mfma_f32_4x4x1f32(ar + ai, b.x + b.y, acc[2], 0, 0, 0);
      };
      const int end = d.view_start[v + 1];
      int e = d.view_start[v];
      // batches of eight entries, software-pipelined: the next batch's operand loads are in
      // flight while the MFMAs of the current one run
      constexpr int kBatch = 8;
      struct Batch {
        float2 a[kBatch], b[kBatch], c[kBatch];
      };
      auto issue = [&](int e0, Batch& B) {
#pragma unroll
        for (int q = 0; q < kBatch; ++q) {
          const TileEntry E = entries[e0 + q];
          operands(E, B.a[q], B.b[q]);
          B.c[q] = make_float2(E.coeff, kReal ? 0.f : coeff_im[e0 + q]);
        }
      };
      auto consume = [&](const Batch& B) {
#pragma unroll
        for (int q = 0; q < kBatch; ++q) mac(B.c[q], B.a[q], B.b[q]);
      };
      if (e + kBatch <= end) {
        Batch X, Y;
        issue(e, X);
        e += kBatch;
        while (true) {
          if (e + kBatch > end) {
            consume(X);
            break;
          }
          issue(e, Y);
          e += kBatch;
          consume(X);
          if (e + kBatch > end) {
            consume(Y);
            break;
          }
          issue(e, X);
          e += kBatch;
          consume(Y);
        }
      }
      for (; e < end; ++e) {
        const TileEntry E = entries[e];
        float2 f, g;
        operands(E, f, g);
        mac(make_float2(E.coeff, kReal ? 0.f : coeff_im[e]), f, g);
      }
    }
    // accumulator register i of lane l: position 0 = 4 ((l / 4) % 4) + i, pa = l / 16, c = l % 4,
    // pb = wave; the four waves write disjoint slabs of the workgroup tile
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int loc[4];
      loc[0] = 4 * (int)((l >> 2) & 3) + i;
      loc[v + 1] = (int)u3;
      loc[view_pa(v)] = (int)u4;
      loc[view_pb(v)] = wave;
      const int t = loc[0] + 16 * loc[1] + 64 * loc[2] + 256 * loc[3];
      const float t1 = acc[0][i], t2 = acc[1][i], t3 = acc[2][i];
      float2 val = make_float2(t1 - t2, t3 - t1 - t2);
      if (v > 0) {
        const float2 prev = part[t];
        val.x += prev.x;
        val.y += prev.y;
      }
      part[t] = val;
    }
    __syncthreads();
  }
  // the workgroup's outputs, four per thread (16 consecutive position-0 indices per 16 lanes)
  int32_t dd[4];
  float2 sum[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int t = threadIdx.x + 256 * k;
    const int p0 = P0 + (t & 15), p1 = P1 + ((t >> 4) & 3), p2 = P2 + ((t >> 6) & 3), p3 = P3 + (t >> 8);
    const bool inside = p0 < n && p1 < n && p2 < n && p3 < n;
    dd[k] = inside ? d.dest[p0 + n * (p1 + n * (p2 + n * p3))] : -1;
    sum[k] = part[t];
  }
  // rank-4 x scalar entries, four entries' loads in flight; tuples without a destination read
  // index 0 of the atom
  auto scalar_terms = [&](int e0, int count) {
    float2 cs[4], av[4][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (q >= count) break;
      const TileScalarEntry S = scalars[e0 + q];
      const float2 sv = ws[S.scalar];
      cs[q] = make_float2(S.coeff.x * sv.x - S.coeff.y * sv.y, S.coeff.x * sv.y + S.coeff.y * sv.x);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int t = threadIdx.x + 256 * k;
        const int p0 = P0 + (t & 15), p1 = P1 + ((t >> 4) & 3), p2 = P2 + ((t >> 6) & 3), p3 = P3 + (t >> 8);
        const int64_t idx = dd[k] >= 0 ? (int64_t)p0 * S.a_stride[0] + (int64_t)p1 * S.a_stride[1] +
                                             (int64_t)p2 * S.a_stride[2] + (int64_t)p3 * S.a_stride[3]
                                       : 0;
        av[q][k] = ws[(int64_t)S.a_off + idx];
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (q >= count) break;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        sum[k].x += cs[q].x * av[q][k].x - cs[q].y * av[q][k].y;
        sum[k].y += cs[q].x * av[q][k].y + cs[q].y * av[q][k].x;
      }
    }
  };
  int e = d.scalar_start;
  for (; e + 4 <= d.scalar_end; e += 4) scalar_terms(e, 4);
  if (e < d.scalar_end) scalar_terms(e, d.scalar_end - e);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (dd[k] < 0) continue;
    double2 v = make_double2((double)sum[k].x, (double)sum[k].y);
    if (diag != nullptr) v = cadd(v, cmul(diag[dd[k]], x[dd[k]]));
    out[dd[k]] = v;
  }
}

__global__ void fill_const_kernel(double2* __restrict__ ws, const double2* __restrict__ vectors,
                                  const int64_t* __restrict__ offset, const int32_t* __restrict__ vec, int n) {
  const int k = blockIdx.x;
  for (int q = threadIdx.x; q < n; q += blockDim.x) ws[offset[k] + q] = vectors[(int64_t)vec[k] * n + q];
}

// ------------------------------------------------------------------ tilde maps

template <typename X>
__global__ void tilde_unpack_kernel(const X* __restrict__ x, double2* __restrict__ t,
                                    const int64_t* __restrict__ k0, const int64_t* __restrict__ k1,
                                    const double4* __restrict__ ab, int64_t dim) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < dim; i += (int64_t)gridDim.x * blockDim.x) {
    const double4 c = ab[i];
    const double a = (double)x[k0[i]];
    const double b = (double)x[k1[i]];
    t[i] = make_double2(c.x * a + c.z * b, c.y * a + c.w * b);
  }
}

// y = (float)(a x), the FP32 copy of a scaled FP64 vector.
__global__ void narrow_scale_kernel(const double* __restrict__ x, float* __restrict__ y, double a, int64_t dim) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < dim; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (float)(a * x[i]);
}

__global__ void tilde_pack_kernel(const double2* __restrict__ t, double* __restrict__ x,
                                  const int64_t* __restrict__ src, const double2* __restrict__ w, int64_t real_dim) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < real_dim;
       c += (int64_t)gridDim.x * blockDim.x) {
    const double2 v = t[src[c]];
    const double2 s = w[c];
    x[c] = s.x * v.x - s.y * v.y;
  }
}

// y[c] = Re(w[c] * (alpha x[k0] + beta x[k1])[src[c]] * inv_diag[src[c]]): unpack,
// divide by the rapidity diagonal and pack in one pass over the real coordinates.
__global__ void tilde_precondition_kernel(const double* __restrict__ x, double* __restrict__ y,
                                          const int64_t* __restrict__ k0, const int64_t* __restrict__ k1,
                                          const double4* __restrict__ ab, const int64_t* __restrict__ src,
                                          const double2* __restrict__ w, const double2* __restrict__ inv_diag,
                                          int64_t real_dim) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < real_dim;
       c += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = src[c];
    const double4 co = ab[i];
    const double a = x[k0[i]];
    const double b = x[k1[i]];
    const double2 t = make_double2(co.x * a + co.z * b, co.y * a + co.w * b);
    const double2 v = cmul(cmul(t, inv_diag[i]), w[c]);
    y[c] = v.x;
  }
}

// x[c] = Re(w[c] * t[src[c]] * inv_diag[src[c]]): pack of the diagonally preconditioned
// complex vector, equal to precondition(pack(t)) for tilde-symmetric t.
__global__ void tilde_pack_precondition_kernel(const double2* __restrict__ t, double* __restrict__ x,
                                               const int64_t* __restrict__ src, const double2* __restrict__ w,
                                               const double2* __restrict__ inv_diag, int64_t real_dim) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < real_dim;
       c += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = src[c];
    const double2 v = cmul(cmul(t[i], inv_diag[i]), w[c]);
    x[c] = v.x;
  }
}

__global__ void rscale_kernel(const double* __restrict__ x, double* __restrict__ y, double a, int64_t dim) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < dim; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * x[i];
}

__global__ void raxpby_kernel(double a, const double* __restrict__ x, double b, double* __restrict__ y, int64_t dim) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < dim; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * x[i] + b * y[i];
}

// ------------------------------------------------------------------ Krylov kernels

// Tall-skinny products V^T w / V a over Krylov bases.  A dot launch gives every
// column one wave per chunk of rows (512, 1024 or 2048 rows: shorter chunks for
// few columns keep the GPU full, longer ones amortise the reduction; measured on
// MI355X at 23,150 rows); a one-workgroup-per-column finish kernel adds the chunk
// partials.  An axpy workgroup covers 64 rows; its four waves take every fourth
// column and meet in LDS.
constexpr int kMinChunkRows = 512;

int dot_rows_for(int ncols) { return ncols <= 64 ? 512 : (ncols <= 256 ? 1024 : 2048); }

// partial-sum slots per column: the largest count any chunk length produces
int64_t chunks_for(int64_t dim) { return (dim + kMinChunkRows - 1) / kMinChunkRows; }

template <typename Launch>
int dispatch_rows(int rows, Launch launch) {
  switch (rows) {
    case 512:
      launch(std::integral_constant<int, 512>());
      break;
    case 1024:
      launch(std::integral_constant<int, 1024>());
      break;
    default:
      launch(std::integral_constant<int, 2048>());
      break;
  }
  return 0;
}

struct cplx {
  double x, y;
};
__device__ __forceinline__ cplx zero_of(cplx) { return cplx{0.0, 0.0}; }
__device__ __forceinline__ double zero_of(double) { return 0.0; }
__device__ __forceinline__ cplx add_of(cplx a, cplx b) { return cplx{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ double add_of(double a, double b) { return a + b; }
__device__ __forceinline__ cplx shfl_down_of(cplx v, int off) {
  return cplx{__shfl_down(v.x, off, 64), __shfl_down(v.y, off, 64)};
}
__device__ __forceinline__ double shfl_down_of(double v, int off) { return __shfl_down(v, off, 64); }

template <typename T>
__device__ T block_reduce(T v, T* shared) {
  // 64-wide wavefront shuffle reduction, then across the 4 waves of the block.
  for (int off = 32; off > 0; off >>= 1) v = add_of(v, shfl_down_of(v, off));
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  if (lane == 0) shared[wave] = v;
  __syncthreads();
  T total = zero_of(v);
  if (threadIdx.x == 0) {
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) total = add_of(total, shared[k]);
  }
  return total;
}

// One wave per column over a chunk of ROWS rows (4 columns per workgroup):
// coalesced 64-lane loads, a wave-level reduction and one partial per (column,
// chunk).  With `extra` set, column ncols-1 is `extra` instead of V + (ncols-1)*ld.
template <int ROWS>
__global__ __launch_bounds__(256) void rmultidot_partial_kernel(const double* __restrict__ V, int64_t ld, int ncols,
                                                                const double* __restrict__ extra,
                                                                const double* __restrict__ w, int64_t dim,
                                                                double* __restrict__ partial, int64_t chunks) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int k = blockIdx.y * 4 + wave;
  if (k >= ncols) return;  // whole wave
  const int64_t r0 = blockIdx.x * (int64_t)ROWS;
  const double* __restrict__ col = (extra != nullptr && k == ncols - 1) ? extra : V + (int64_t)k * ld;
  double acc = 0.0;
#pragma unroll 8
  for (int t = 0; t < ROWS / 64; ++t) {
    const int64_t i = r0 + t * 64 + lane;
    if (i < dim) acc += col[i] * w[i];
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) partial[(int64_t)k * chunks + blockIdx.x] = acc;
}

__global__ void rfinish_kernel(const double* __restrict__ partial, int64_t chunks, double* __restrict__ out) {
  __shared__ double shared[kThreads / 64];
  const int k = blockIdx.x;
  double acc = 0.0;
  for (int64_t c = threadIdx.x; c < chunks; c += blockDim.x) acc += partial[(int64_t)k * chunks + c];
  const double total = block_reduce(acc, shared);
  if (threadIdx.x == 0) out[k] = total;
}

// w[i] += scale * sum_k alpha[k] V[k*ld + i]; 64 rows per workgroup, columns split over the 4 waves.
template <typename S>
__global__ __launch_bounds__(256) void rmultiaxpy_kernel(const S* __restrict__ V, int64_t ld, int ncols,
                                                         const double* __restrict__ alpha, double scale,
                                                         double* __restrict__ w, int64_t dim) {
  __shared__ double part[4][64];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t i = blockIdx.x * (int64_t)64 + lane;
  double acc = 0.0;
  if (i < dim) {
#pragma unroll 4
    for (int k = wave; k < ncols; k += 4) acc += alpha[k] * (double)V[(int64_t)k * ld + i];
  }
  part[wave][lane] = acc;
  __syncthreads();
  if (wave == 0 && i < dim) w[i] += scale * (part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane]);
}

// Complex twin of rmultidot_partial_kernel; conj(V) when `conjugate`, and the
// `extra` column (when set) is never conjugated.
template <int ROWS>
__global__ __launch_bounds__(256) void zmultidot_partial_kernel(const double2* __restrict__ V, int64_t ld, int ncols,
                                                                const double2* __restrict__ extra,
                                                                const double2* __restrict__ w, int64_t dim,
                                                                int conjugate, double2* __restrict__ partial,
                                                                int64_t chunks) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int k = blockIdx.y * 4 + wave;
  if (k >= ncols) return;  // whole wave
  const int64_t r0 = blockIdx.x * (int64_t)ROWS;
  const bool ex = extra != nullptr && k == ncols - 1;
  const double2* __restrict__ col = ex ? extra : V + (int64_t)k * ld;
  const double sg = (conjugate && !ex) ? -1.0 : 1.0;
  double ax = 0.0, ay = 0.0;
#pragma unroll 8
  for (int t = 0; t < ROWS / 64; ++t) {
    const int64_t i = r0 + t * 64 + lane;
    if (i < dim) {
      const double2 a = col[i];
      const double2 b = w[i];
      ax += a.x * b.x - sg * a.y * b.y;
      ay += a.x * b.y + sg * a.y * b.x;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    ax += __shfl_down(ax, off, 64);
    ay += __shfl_down(ay, off, 64);
  }
  if (lane == 0) partial[(int64_t)k * chunks + blockIdx.x] = make_double2(ax, ay);
}

__global__ void zfinish_kernel(const double2* __restrict__ partial, int64_t chunks, double2* __restrict__ out) {
  __shared__ cplx shared[kThreads / 64];
  const int k = blockIdx.x;
  cplx acc{0.0, 0.0};
  for (int64_t c = threadIdx.x; c < chunks; c += blockDim.x) {
    const double2 p = partial[(int64_t)k * chunks + c];
    acc.x += p.x;
    acc.y += p.y;
  }
  const cplx total = block_reduce(acc, shared);
  if (threadIdx.x == 0) out[k] = make_double2(total.x, total.y);
}

__global__ __launch_bounds__(256) void zmultiaxpy_kernel(const double2* __restrict__ V, int64_t ld, int ncols,
                                                         const double2* __restrict__ alpha, double scale,
                                                         double2* __restrict__ w, int64_t dim) {
  __shared__ double2 part[4][64];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t i = blockIdx.x * (int64_t)64 + lane;
  double2 acc = make_double2(0.0, 0.0);
  if (i < dim) {
#pragma unroll 4
    for (int k = wave; k < ncols; k += 4) acc = cadd(acc, cmul(alpha[k], V[(int64_t)k * ld + i]));
  }
  part[wave][lane] = acc;
  __syncthreads();
  if (wave == 0 && i < dim) {
    const double2 t = make_double2(part[0][lane].x + part[1][lane].x + part[2][lane].x + part[3][lane].x,
                                   part[0][lane].y + part[1][lane].y + part[2][lane].y + part[3][lane].y);
    w[i] = make_double2(w[i].x + scale * t.x, w[i].y + scale * t.y);
  }
}

// One 64-lane workgroup per shift: column-oriented back substitution on the
// upper-triangular Schur factor held in L2, right-hand side in LDS.
__global__ void shifted_triangular_kernel(const double2* __restrict__ T, int m, const double2* __restrict__ r,
                                          const double2* __restrict__ w, const double2* __restrict__ q,
                                          const double2* __restrict__ shifts, double2* __restrict__ green,
                                          double2* __restrict__ last) {
  extern __shared__ double2 b[];
  const int s = blockIdx.x;
  const double2 z = shifts[s];
  for (int i = threadIdx.x; i < m; i += blockDim.x) b[i] = r[i];
  __syncthreads();
  for (int k = m - 1; k >= 0; --k) {
    const double2 d = T[(int64_t)k * m + k];
    const double2 den = make_double2(z.x - d.x, z.y - d.y);
    const double inv = 1.0 / (den.x * den.x + den.y * den.y);
    const double2 bk = b[k];
    const double2 yk = make_double2((bk.x * den.x + bk.y * den.y) * inv, (bk.y * den.x - bk.x * den.y) * inv);
    __syncthreads();
    if (threadIdx.x == 0) b[k] = yk;
    for (int i = threadIdx.x; i < k; i += blockDim.x) b[i] = cadd(b[i], cmul(T[(int64_t)k * m + i], yk));
    __syncthreads();
  }
  cplx g{0.0, 0.0}, l{0.0, 0.0};
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    const double2 y = b[i];
    const double2 wi = w[i];
    const double2 qi = q[i];
    g.x += wi.x * y.x - wi.y * y.y;
    g.y += wi.x * y.y + wi.y * y.x;
    l.x += qi.x * y.x - qi.y * y.y;
    l.y += qi.x * y.y + qi.y * y.x;
  }
  for (int off = 32; off > 0; off >>= 1) {
    g = add_of(g, shfl_down_of(g, off));
    l = add_of(l, shfl_down_of(l, off));
  }
  if (threadIdx.x == 0) {
    green[s] = make_double2(g.x, g.y);
    last[s] = make_double2(l.x, l.y);
  }
}

}  // namespace

// ------------------------------------------------------------------ plan object

struct PlanBase {
  int32_t kind;
  int64_t offset;
  int64_t count;
  std::vector<int64_t> src;  // host copies; runs of adjacent bases are uploaded together
  std::vector<int8_t> sign;
};

// One scatter launch per run of workspace-adjacent base blocks of a kind.
struct BaseRun {
  int64_t offset;
  int64_t count;
  int64_t* src = nullptr;
  int8_t* sign = nullptr;
};

constexpr int kKinds = 3;

struct KindExec {
  bool built = false;
  std::vector<BaseRun> runs;
  std::vector<void*> owned;
  // inverse of the scatter (c3l_internal_plan_scatter_map): per canonical index two codes
  // (padded offset << 1 | negative), -1 for none; built at its first request
  bool inverse_built = false;
  int64_t inverse_dim = 0;
  int64_t* inverse = nullptr;
  // levels before tree_level run one launch each; the jobs of the later levels run as
  // contraction trees in one launch (tree_level = number of levels: no trees)
  int32_t tree_level = 0;
  int32_t n_trees = 0;
  ContractJob* tree_jobs = nullptr;
  int32_t* tree_start = nullptr;
  uint8_t* tree_barrier = nullptr;
};

// Complex MACs up to which one workgroup evaluates a contraction tree (a rank-2 root
// of n = 25, 625 MACs, and its descendants qualify; a level of rank-3 roots, 15625 MACs
// each, runs as one launch over the whole GPU: a workgroup per such tree leaves most
// CUs idle and took longer than the extra launch).
constexpr int64_t kTreeMacs = 1 << 13;

struct PlanLevel {
  int32_t kind;
  int32_t n_jobs;
  int64_t max_out;
  ContractJob* jobs = nullptr;
  std::vector<ContractJob> host_jobs;
};

struct PlanBlock {
  int32_t program;
  int32_t deg;
  int64_t count;
  int32_t n_entries;
  int32_t n_factors;
  int64_t max_dest = -1;
  bool two_factors = true;
  bool dense = false;  // every index tuple of a padded rank-deg tensor, written at dense_base + flat
  int64_t dense_base = 0;
  // host copies, concatenated into the program's device arrays when it is built
  std::vector<uint32_t> coord;
  std::vector<int64_t> dest;
  std::vector<WedgeEntry> entries;
  std::vector<WedgeFactor> factors;
};

// Device image of one program: every block's arrays concatenated, the block
// descriptors and the tile lists of the two kernel variants.
// Launch shape of one program image.  The defaults are the measured heuristics; the
// autotuner (c3l_plan_autotune) times alternatives on the plan itself and keeps the fastest
// for every plan with the same digest, program and precision.
struct ProgramShape {
  int32_t lanes_max = 2;       // most lanes per coordinate, log2 (small blocks split their entries)
  int32_t lanes_target = 512;  // workgroups per block below which a block splits its entries
  int32_t reps_max = 2;        // coordinates per thread of large two-factor blocks (<= kMaxReps)
  int32_t reps_target = 512;   // workgroups per block below which a block halves its coordinates per thread
  int32_t xcd = 1;             // tiles of a block sliced over the eight XCDs
  int32_t mfma_min = kTileMinEntries;  // FP32: fewest entries of a block on the matrix cores (INT_MAX: none)
  bool operator==(const ProgramShape& o) const {
    return lanes_max == o.lanes_max && lanes_target == o.lanes_target && reps_max == o.reps_max &&
           reps_target == o.reps_target && xcd == o.xcd && mfma_min == o.mfma_min;
  }
};

struct ProgramExec {
  bool built = false;
  bool five = false;       // the program has rank-five output blocks
  ProgramShape shape;      // the shape the current images were built with
  int64_t tune_gen = -1;   // tuning generation at which `shape` was looked up
  int64_t max_dest = -1;
  int32_t tiles_two = 0;
  int32_t tiles_general = 0;
  WedgeBlockDesc* descs = nullptr;
  int2* tiles = nullptr;  // two-factor tiles first, then general tiles
  std::vector<void*> owned;
  std::vector<int32_t> desc_block;  // plan block of every descriptor
  std::vector<int2> host_tiles;
  // FP32 matrix-core image (built at the program's first FP32 launch): rank-4 two-atom blocks
  // go to tile_program_mfma_kernel, the rest keep the gather kernels
  bool mfma_built = false;
  int32_t mfma_tiles = 0;
  bool mfma_real = false;  // every matrix-core entry has a real coefficient
  TileBlockDesc* mfma_descs = nullptr;
  TileRule* mfma_rules = nullptr;
  TileEntry* mfma_entries = nullptr;
  float* mfma_coeff_im = nullptr;
  TileScalarEntry* mfma_scalars = nullptr;
  int2* mfma_tile_list = nullptr;
  int32_t tiles32_two = 0, tiles32_general = 0;
  int2* tiles32 = nullptr;
  std::vector<void*> owned_mfma;  // device arrays of the matrix-core image and tiles32
};

struct PendingTiming {
  int32_t program;
  int32_t instances;  // twin plans evaluated by the timed launch
  hipEvent_t start;
  hipEvent_t stop;
};

constexpr int kPrograms = 9;

struct c3l_plan {
  int32_t n = 0;
  int32_t n_vectors = 0;
  int64_t workspace = 0;
  double2* ws = nullptr;
  double2* vectors = nullptr;
  std::vector<PlanBase> bases;
  std::vector<PlanLevel> levels;
  std::vector<PlanBlock> blocks;
  ProgramExec programs[kPrograms];
  KindExec kinds[kKinds];
  std::vector<void*> owned;
  // mixed precision: kinds loaded and programs evaluated in FP32 (complex64) use ws32,
  // a float2 mirror of the workspace; FP64-produced atoms they read are narrowed into it
  // whenever the FP64 side changed (fixed_version) since the last narrowing.
  int32_t kind_bits[kKinds] = {64, 64, 64};
  int32_t program_bits[kPrograms] = {64, 64, 64, 64, 64, 64, 64, 64, 64};
  float2* ws32 = nullptr;
  int64_t ws32_tail = 0;  // zeroed elements allocated past the FP32 workspace
  float2* vectors32 = nullptr;
  // FP32 programs: rank-4 two-atom blocks on the matrix cores: 0 never, 1 blocks with at least
  // kTileMinEntries entries (below that the gather kernels are as fast), 2 every block
  int32_t mfma = 1;
  int64_t fixed_version = 0;
  int64_t narrowed_version = -1;
  bool ranges_built = false;
  int32_t n_ranges = 0;
  int64_t max_range = 0;
  int64_t* ranges = nullptr;  // device [2 * n_ranges]: (begin, count)
  // optional per-program kernel timing with HIP events
  bool timing = false;
  std::vector<PendingTiming> pending;
  double program_ms[kPrograms] = {};
  int64_t program_launches[kPrograms] = {};
  // FNV-1a digest of everything the plan was built from (sizes, vectors, constant
  // atoms, bases, levels, blocks): plans with equal digests are twins and can be
  // evaluated together in one launch per kernel, each on its own workspace.
  uint64_t digest = 1469598103934665603ull;
};

namespace {

template <typename T>
int upload(c3l_plan* plan, T** dst, const T* src, int64_t count) {
  *dst = nullptr;
  if (count <= 0) return C3L_OK;
  void* ptr = nullptr;
  if (hipMalloc(&ptr, sizeof(T) * (size_t)count) != hipSuccess) return fail(C3L_ENOMEM, "hipMalloc failed in plan upload");
  plan->owned.push_back(ptr);
  C3L_HIP_CHECK(hipMemcpy(ptr, src, sizeof(T) * (size_t)count, hipMemcpyHostToDevice));
  *dst = static_cast<T*>(ptr);
  return C3L_OK;
}

template <typename T>
int upload_owned(std::vector<void*>& owned, T** dst, const T* src, size_t count) {
  *dst = nullptr;
  if (count == 0) return C3L_OK;
  void* ptr = nullptr;
  if (hipMalloc(&ptr, sizeof(T) * count) != hipSuccess) return fail(C3L_ENOMEM, "hipMalloc failed in program build");
  owned.push_back(ptr);
  C3L_HIP_CHECK(hipMemcpy(ptr, src, sizeof(T) * count, hipMemcpyHostToDevice));
  *dst = static_cast<T*>(ptr);
  return C3L_OK;
}

int64_t ipow(int64_t n, int r) {
  int64_t v = 1;
  for (int k = 0; k < r; ++k) v *= n;
  return v;
}

void mix(c3l_plan* plan, const void* data, size_t bytes) {
  const unsigned char* p = static_cast<const unsigned char*>(data);
  uint64_t h = plan->digest;
  for (size_t i = 0; i < bytes; ++i) h = (h ^ p[i]) * 1099511628211ull;
  plan->digest = h;
}

template <typename T>
void mix_value(c3l_plan* plan, T v) {
  mix(plan, &v, sizeof(T));
}

// Splits the contraction levels of `kind` into per-level launches (levels before
// ke.tree_level) and contraction trees: the first level from which every tree (a job
// whose parent is not produced from that level on, with all its dependants) stays
// within kTreeMacs.  Uploads the tree jobs in tree order.
int build_trees(c3l_plan* plan, int32_t kind, KindExec& ke) {
  std::vector<const PlanLevel*> lv;
  for (const PlanLevel& l : plan->levels)
    if (l.kind == kind && l.n_jobs > 0) lv.push_back(&l);
  ke.tree_level = (int32_t)lv.size();
  for (size_t first = 0; first < lv.size(); ++first) {
    struct Node {
      ContractJob job;
      int32_t level;
      int32_t root;
    };
    std::vector<Node> nodes;
    std::map<int64_t, int32_t> producer;  // atom offset -> node that writes it
    std::vector<int64_t> macs;            // per node: MACs of the tree it roots
    bool fits = true;
    for (size_t l = first; l < lv.size() && fits; ++l) {
      for (const ContractJob& j : lv[l]->host_jobs) {
        const auto it = producer.find(j.par_offset);
        const int32_t id = (int32_t)nodes.size();
        const int32_t root = it == producer.end() ? id : nodes[(size_t)it->second].root;
        nodes.push_back(Node{j, (int32_t)l, root});
        macs.push_back(0);
        macs[(size_t)root] += ipow(plan->n, j.par_rank);
        if (macs[(size_t)root] > kTreeMacs) fits = false;
        producer[j.out_offset] = id;
      }
    }
    if (!fits) continue;
    // trees in root order, each ordered by level (nodes are already in level order)
    std::vector<std::vector<int32_t>> members(nodes.size());
    for (size_t i = 0; i < nodes.size(); ++i) members[(size_t)nodes[i].root].push_back((int32_t)i);
    std::vector<ContractJob> jobs;
    std::vector<int32_t> start{0};
    std::vector<uint8_t> barrier;
    for (const std::vector<int32_t>& tree : members) {
      if (tree.empty()) continue;
      for (size_t k = 0; k < tree.size(); ++k) {
        const Node& nd = nodes[(size_t)tree[k]];
        jobs.push_back(nd.job);
        barrier.push_back(k > 0 && nd.level != nodes[(size_t)tree[k - 1]].level ? 1 : 0);
      }
      start.push_back((int32_t)jobs.size());
    }
    int rc = upload_owned(ke.owned, &ke.tree_jobs, jobs.data(), jobs.size());
    if (rc == C3L_OK) rc = upload_owned(ke.owned, &ke.tree_start, start.data(), start.size());
    if (rc == C3L_OK) rc = upload_owned(ke.owned, &ke.tree_barrier, barrier.data(), barrier.size());
    if (rc != C3L_OK) return rc;
    ke.tree_level = (int32_t)first;
    ke.n_trees = (int32_t)start.size() - 1;
    break;
  }
  return C3L_OK;
}

}  // namespace

extern "C" {

const char* c3l_last_error(void) { return g_error.c_str(); }

int c3l_internal_fail(int code, const char* message) { return fail(code, message ? message : ""); }

const char* c3l_version(void) { return "c3l_hip 0.1 gfx950 fp64"; }

int c3l_plan_create(int32_t n, int32_t n_vectors, const double* vectors, int64_t workspace, int32_t n_const,
                    const int32_t* const_atom, const int64_t* const_offset, const int32_t* const_vec,
                    c3l_plan** out) {
  (void)const_atom;
  if (out == nullptr || n <= 0 || n_vectors < 0 || workspace < 0 || n_const < 0)
    return fail(C3L_EINVAL, "c3l_plan_create: invalid argument");
  for (int32_t k = 0; k < n_const; ++k) {
    if (const_vec[k] < 0 || const_vec[k] >= n_vectors || const_offset[k] < 0 || const_offset[k] + n > workspace)
      return fail(C3L_EINVAL, "c3l_plan_create: constant atom out of range");
  }
  c3l_plan* plan = new c3l_plan();
  plan->n = n;
  plan->n_vectors = n_vectors;
  plan->workspace = workspace;
  void* ws = nullptr;
  if (hipMalloc(&ws, sizeof(double2) * (size_t)(workspace > 0 ? workspace : 1)) != hipSuccess) {
    delete plan;
    return fail(C3L_ENOMEM, "c3l_plan_create: workspace allocation failed");
  }
  plan->owned.push_back(ws);
  plan->ws = static_cast<double2*>(ws);
  mix_value(plan, n);
  mix_value(plan, n_vectors);
  mix_value(plan, workspace);
  mix(plan, vectors, sizeof(double2) * (size_t)n_vectors * (size_t)n);
  mix(plan, const_offset, sizeof(int64_t) * (size_t)n_const);
  mix(plan, const_vec, sizeof(int32_t) * (size_t)n_const);
  int rc = upload(plan, &plan->vectors, reinterpret_cast<const double2*>(vectors), (int64_t)n_vectors * n);
  int64_t* d_off = nullptr;
  int32_t* d_vec = nullptr;
  if (rc == C3L_OK) rc = upload(plan, &d_off, const_offset, n_const);
  if (rc == C3L_OK) rc = upload(plan, &d_vec, const_vec, n_const);
  if (rc == C3L_OK && hipMemset(plan->ws, 0, sizeof(double2) * (size_t)(workspace > 0 ? workspace : 1)) != hipSuccess)
    rc = fail(C3L_EHIP, "c3l_plan_create: workspace clear failed");
  if (rc == C3L_OK && n_const > 0) {
    hipLaunchKernelGGL(fill_const_kernel, dim3(n_const), dim3(64), 0, 0, plan->ws, plan->vectors, d_off, d_vec, n);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
      rc = fail(C3L_EHIP, "c3l_plan_create: constant atom fill failed");
  }
  if (rc != C3L_OK) {
    c3l_plan_destroy(plan);
    return rc;
  }
  *out = plan;
  return C3L_OK;
}

int c3l_plan_destroy(c3l_plan* plan) {
  if (plan == nullptr) return C3L_OK;
  for (PendingTiming& p : plan->pending) {
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.stop);
  }
  for (ProgramExec& pe : plan->programs) {
    for (void* p : pe.owned) (void)hipFree(p);
    for (void* p : pe.owned_mfma) (void)hipFree(p);
  }
  for (KindExec& ke : plan->kinds)
    for (void* p : ke.owned) (void)hipFree(p);
  for (void* p : plan->owned) (void)hipFree(p);
  delete plan;
  return C3L_OK;
}

int c3l_plan_set_timing(c3l_plan* plan, int32_t enable) {
  if (plan == nullptr) return fail(C3L_EINVAL, "c3l_plan_set_timing: invalid argument");
  plan->timing = enable != 0;
  for (int k = 0; k < kPrograms; ++k) {
    plan->program_ms[k] = 0.0;
    plan->program_launches[k] = 0;
  }
  return C3L_OK;
}

int c3l_plan_program_time(c3l_plan* plan, int32_t program, double* ms_total, int64_t* launches) {
  if (plan == nullptr || program < 0 || program >= kPrograms)
    return fail(C3L_EINVAL, "c3l_plan_program_time: invalid program");
  // resolve every recorded timing (waits for the recorded work)
  for (PendingTiming& p : plan->pending) {
    C3L_HIP_CHECK(hipEventSynchronize(p.stop));
    float ms = 0.f;
    C3L_HIP_CHECK(hipEventElapsedTime(&ms, p.start, p.stop));
    plan->program_ms[p.program] += ms;
    plan->program_launches[p.program] += p.instances;
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.stop);
  }
  plan->pending.clear();
  if (ms_total) *ms_total = plan->program_ms[program];
  if (launches) *launches = plan->program_launches[program];
  return C3L_OK;
}

int c3l_plan_block_info(c3l_plan* plan, int32_t index, int32_t* program, int32_t* deg, int64_t* count,
                        int32_t* n_entries, int32_t* n_factors) {
  if (plan == nullptr || index < 0 || index >= (int32_t)plan->blocks.size())
    return fail(C3L_EINVAL, "c3l_plan_block_info: invalid block index");
  const PlanBlock& b = plan->blocks[(size_t)index];
  if (program) *program = b.program;
  if (deg) *deg = b.deg;
  if (count) *count = b.count;
  if (n_entries) *n_entries = b.n_entries;
  if (n_factors) *n_factors = b.n_factors;
  return C3L_OK;
}

int32_t c3l_plan_block_count(c3l_plan* plan) { return plan == nullptr ? 0 : (int32_t)plan->blocks.size(); }

int c3l_plan_add_base(c3l_plan* plan, int32_t kind, int64_t offset, int32_t deg, const int64_t* src,
                      const int8_t* sign) {
  if (plan == nullptr || kind < 0 || kind >= kKinds || deg < 0 || deg > 5 || offset < 0)
    return fail(C3L_EINVAL, "c3l_plan_add_base: invalid argument");
  PlanBase b;
  b.kind = kind;
  b.offset = offset;
  b.count = ipow(plan->n, deg);
  if (offset + b.count > plan->workspace) return fail(C3L_EINVAL, "c3l_plan_add_base: block exceeds workspace");
  b.src.assign(src, src + b.count);
  b.sign.assign(sign, sign + b.count);
  mix_value(plan, (int32_t)1);
  mix_value(plan, kind);
  mix_value(plan, offset);
  mix_value(plan, deg);
  mix(plan, src, sizeof(int64_t) * (size_t)b.count);
  mix(plan, sign, (size_t)b.count);
  plan->bases.push_back(std::move(b));
  plan->ranges_built = false;
  KindExec& ke = plan->kinds[kind];
  for (void* p : ke.owned) (void)hipFree(p);
  ke = KindExec();  // rebuilt at the next load
  return C3L_OK;
}

int c3l_plan_add_level(c3l_plan* plan, int32_t kind, int32_t n_jobs, const int64_t* out_offset,
                       const int64_t* par_offset, const int32_t* par_rank, const int32_t* slot, const int32_t* vec,
                       const int32_t* sign) {
  if (plan == nullptr || n_jobs < 0) return fail(C3L_EINVAL, "c3l_plan_add_level: invalid argument");
  std::vector<ContractJob> jobs((size_t)n_jobs);
  int64_t max_out = 1;  // threads the level's largest job needs
  for (int32_t j = 0; j < n_jobs; ++j) {
    if (par_rank[j] < 1 || par_rank[j] > 5 || slot[j] < 0 || slot[j] >= par_rank[j] || vec[j] < 0 ||
        vec[j] >= plan->n_vectors)
      return fail(C3L_EINVAL, "c3l_plan_add_level: job out of range");
    const int64_t outsz = ipow(plan->n, par_rank[j] - 1);
    if (out_offset[j] < 0 || out_offset[j] + outsz > plan->workspace || par_offset[j] < 0 ||
        par_offset[j] + ipow(plan->n, par_rank[j]) > plan->workspace)
      return fail(C3L_EINVAL, "c3l_plan_add_level: job exceeds workspace");
    jobs[(size_t)j] = ContractJob{out_offset[j], par_offset[j], par_rank[j], slot[j], vec[j], sign[j]};
    const int64_t threads = outsz * contract_lanes(slot[j], plan->n);
    if (threads > max_out) max_out = threads;
  }
  mix_value(plan, (int32_t)2);
  mix_value(plan, kind);
  mix(plan, jobs.data(), sizeof(ContractJob) * jobs.size());
  PlanLevel level;
  level.kind = kind;
  level.n_jobs = n_jobs;
  level.max_out = max_out;
  int rc = upload(plan, &level.jobs, jobs.data(), n_jobs);
  if (rc != C3L_OK) return rc;
  level.host_jobs = jobs;
  plan->levels.push_back(level);
  plan->ranges_built = false;
  if (kind >= 0 && kind < kKinds) {
    KindExec& ke = plan->kinds[kind];
    for (void* p : ke.owned) (void)hipFree(p);
    ke = KindExec();  // rebuilt at the next load
  }
  return C3L_OK;
}

}  // extern "C"

namespace {

// Entries and factors of a block being registered (shared by the listed and dense forms).
int parse_block_entries(c3l_plan* plan, PlanBlock& b, int32_t deg, int32_t n_entries, const double* coeff,
                        const int32_t* fstart, int32_t n_factors, const int64_t* foffset, const int32_t* fcode) {
  if (n_entries > 0 && (fstart[0] != 0 || fstart[n_entries] != n_factors))
    return fail(C3L_EINVAL, "c3l_plan_add_block: fstart mismatch");
  b.n_entries = n_entries;
  b.n_factors = n_factors;
  b.factors.resize((size_t)n_factors);
  for (int32_t f = 0; f < n_factors; ++f) {
    const int rank = fcode[f] & 7;
    if (rank > 5 || foffset[f] < 0 || foffset[f] + ipow(plan->n, rank) > plan->workspace)
      return fail(C3L_EINVAL, "c3l_plan_add_block: factor exceeds workspace");
    WedgeFactor wf;
    wf.offset = foffset[f];
    for (int k = 0; k < 5; ++k) wf.stride[k] = 0;
    for (int q = 0; q < rank; ++q) {
      const int pos = (fcode[f] >> (3 + 3 * q)) & 7;
      if (pos >= deg) return fail(C3L_EINVAL, "c3l_plan_add_block: factor position out of range");
      wf.stride[pos] += (int32_t)ipow(plan->n, q);  // atoms store their first index fastest
    }
    b.factors[(size_t)f] = wf;
  }
  b.entries.resize((size_t)n_entries);
  for (int32_t t = 0; t < n_entries; ++t) {
    if (fstart[t + 1] < fstart[t]) return fail(C3L_EINVAL, "c3l_plan_add_block: fstart not monotone");
    WedgeEntry en;
    en.coeff = make_double2(coeff[2 * t], coeff[2 * t + 1]);
    en.first = fstart[t];
    en.nf = fstart[t + 1] - fstart[t];
    en.pad[0] = en.pad[1] = 0;
    if (en.nf != 2) b.two_factors = false;
    b.entries[(size_t)t] = en;
  }
  return C3L_OK;
}

void finish_block(c3l_plan* plan, PlanBlock&& b) {
  plan->blocks.push_back(std::move(b));
  plan->ranges_built = false;
  ProgramExec& pe = plan->programs[plan->blocks.back().program];
  for (void* p : pe.owned) (void)hipFree(p);
  pe = ProgramExec();  // rebuilt at the next apply
}

}  // namespace

extern "C" {

int c3l_plan_add_block(c3l_plan* plan, int32_t program, int32_t deg, int64_t count, const int32_t* coord,
                       const int64_t* dest, int32_t n_entries, const double* coeff, const int32_t* fstart,
                       int32_t n_factors, const int64_t* foffset, const int32_t* fcode) {
  if (plan == nullptr || program < 0 || program >= kPrograms || deg < 0 || deg > 5 || count < 0 ||
      count > INT32_MAX || n_entries < 0 || n_factors < 0)
    return fail(C3L_EINVAL, "c3l_plan_add_block: invalid argument");
  // packed coordinates: four 8-bit indices up to rank four, five 6-bit indices at rank five
  const int bits = deg <= 4 ? 8 : 6;
  if (plan->n > (1 << bits) - 1)
    return fail(C3L_EINVAL, deg <= 4 ? "c3l_plan_add_block: more than 255 sites per group"
                                     : "c3l_plan_add_block: rank-five blocks take at most 63 sites per group");
  PlanBlock b;
  b.program = program;
  b.deg = deg;
  b.count = count;
  int rc = parse_block_entries(plan, b, deg, n_entries, coeff, fstart, n_factors, foffset, fcode);
  if (rc != C3L_OK) return rc;
  b.coord.resize((size_t)count);
  b.dest.assign(dest, dest + count);
  for (int64_t k = 0; k < count; ++k) {
    if (dest[k] < 0) return fail(C3L_EINVAL, "c3l_plan_add_block: negative destination");
    if (dest[k] > b.max_dest) b.max_dest = dest[k];
    uint32_t packed = 0;
    for (int q = 0; q < deg; ++q) {
      const int32_t c = coord[k * deg + q];
      if (c < 0 || c >= plan->n) return fail(C3L_EINVAL, "c3l_plan_add_block: coordinate out of range");
      packed |= (uint32_t)c << (bits * q);
    }
    b.coord[(size_t)k] = packed;
  }
  mix_value(plan, (int32_t)3);
  mix_value(plan, program);
  mix_value(plan, deg);
  mix(plan, b.coord.data(), sizeof(uint32_t) * b.coord.size());
  mix(plan, dest, sizeof(int64_t) * (size_t)count);
  mix(plan, coeff, sizeof(double) * 2 * (size_t)n_entries);
  if (n_entries > 0) mix(plan, fstart, sizeof(int32_t) * ((size_t)n_entries + 1));
  mix(plan, foffset, sizeof(int64_t) * (size_t)n_factors);
  mix(plan, fcode, sizeof(int32_t) * (size_t)n_factors);
  finish_block(plan, std::move(b));
  return C3L_OK;
}

int c3l_plan_add_dense_block(c3l_plan* plan, int32_t program, int32_t deg, int64_t base, int32_t n_entries,
                             const double* coeff, const int32_t* fstart, int32_t n_factors, const int64_t* foffset,
                             const int32_t* fcode) {
  if (plan == nullptr || program < 0 || program >= kPrograms || deg < 0 || deg > 5 || base < 0 || n_entries < 0 ||
      n_factors < 0)
    return fail(C3L_EINVAL, "c3l_plan_add_dense_block: invalid argument");
  if (plan->n > 255) return fail(C3L_EINVAL, "c3l_plan_add_dense_block: more than 255 sites per group");
  const int64_t count = ipow(plan->n, deg);
  if (count > INT32_MAX) return fail(C3L_EINVAL, "c3l_plan_add_dense_block: block too large");
  PlanBlock b;
  b.program = program;
  b.deg = deg;
  b.count = count;
  b.dense = true;
  b.dense_base = base;
  b.max_dest = base + count - 1;
  int rc = parse_block_entries(plan, b, deg, n_entries, coeff, fstart, n_factors, foffset, fcode);
  if (rc != C3L_OK) return rc;
  mix_value(plan, (int32_t)4);
  mix_value(plan, program);
  mix_value(plan, deg);
  mix(plan, &base, sizeof(base));
  mix(plan, coeff, sizeof(double) * 2 * (size_t)n_entries);
  if (n_entries > 0) mix(plan, fstart, sizeof(int32_t) * ((size_t)n_entries + 1));
  mix(plan, foffset, sizeof(int64_t) * (size_t)n_factors);
  mix(plan, fcode, sizeof(int32_t) * (size_t)n_factors);
  finish_block(plan, std::move(b));
  return C3L_OK;
}

int c3l_internal_plan_load_many(c3l_plan* const* plans, int32_t count, int32_t kind, const double* const* v,
                                void* stream);
int c3l_internal_plan_apply_many(c3l_plan* const* plans, int32_t count, int32_t program, const double* diag,
                                 const double* const* x, double* const* out, void* stream);

// 1 when every plan of the list was built from the same arrays (equal digests) and all
// plans are FP64 in kind `kind` and program `program`, else 0.
int c3l_internal_plans_are_twins(c3l_plan* const* plans, int32_t count, int32_t kind, int32_t program) {
  if (plans == nullptr || count < 1 || count > kMaxTwins || kind < 0 || kind >= kKinds || program < 0 ||
      program >= kPrograms)
    return 0;
  for (int32_t i = 0; i < count; ++i) {
    if (plans[i] == nullptr) return 0;
    if (count > 1 && (plans[i]->digest != plans[0]->digest || plans[i]->kind_bits[kind] != 64 ||
                      plans[i]->program_bits[program] != 64))
      return 0;
  }
  return 1;
}

int c3l_plan_load(c3l_plan* plan, int32_t kind, const double* v, void* stream) {
  return c3l_internal_plan_load_many(&plan, 1, kind, &v, stream);
}

// Loads vector v[i] into twin plan plans[i] for i < count with one launch per kernel
// (blockIdx.z = instance).  Twins are checked by digest; several plans load in FP64 only.
int plan_load_many(c3l_plan* const* plans, int32_t count, int32_t kind, const double* const* v, void* stream,
                   bool scatter);

int c3l_internal_plan_load_many(c3l_plan* const* plans, int32_t count, int32_t kind, const double* const* v,
                                void* stream) {
  return plan_load_many(plans, count, kind, v, stream, true);
}

// c3l_internal_plan_load_many without the scatter of the input: the padded base blocks of
// `kind` already hold the input (written by the producer of the vector through the map of
// c3l_internal_plan_scatter_map); only the contraction levels and trees run.
int c3l_internal_plan_load_prepared_many(c3l_plan* const* plans, int32_t count, int32_t kind, void* stream) {
  const double* none[kMaxTwins] = {nullptr, nullptr, nullptr, nullptr};
  double dummy = 0.0;
  for (int32_t i = 0; i < count && i < kMaxTwins; ++i) none[i] = &dummy;  // never read
  return plan_load_many(plans, count, kind, none, stream, false);
}

// Inverse of the scatter of kind `kind` for canonical vectors of `dim` entries: map[2 e + k]
// (k = 0, 1) = (workspace offset << 1) | (1 when the value enters negated) of the padded
// elements that take entry e, -1 for none.  Only kinds whose entries each fill at most two
// padded elements have one (*map = NULL otherwise).  The workspace offsets are relative to
// the plan's workspace (twin plans share the map).
int c3l_internal_plan_scatter_map(c3l_plan* plan, int32_t kind, int64_t dim, const int64_t** map) {
  if (plan == nullptr || map == nullptr || kind < 0 || kind >= kKinds || dim <= 0)
    return fail(C3L_EINVAL, "c3l_internal_plan_scatter_map: invalid argument");
  KindExec& ke = plan->kinds[kind];
  *map = nullptr;
  if (ke.inverse_built) {
    if (ke.inverse_dim == dim) *map = ke.inverse;
    return C3L_OK;
  }
  ke.inverse_built = true;
  ke.inverse_dim = dim;
  std::vector<int64_t> inv((size_t)(2 * dim), -1);
  for (const PlanBase& b : plan->bases) {
    if (b.kind != kind) continue;
    for (int64_t e = 0; e < b.count; ++e) {
      const int64_t src = b.src[(size_t)e];
      if (src < 0) continue;
      if (src >= dim) return C3L_OK;  // not a vector of this size: no map
      const int64_t code = ((b.offset + e) << 1) | (b.sign[(size_t)e] < 0 ? 1 : 0);
      if (inv[(size_t)(2 * src)] < 0)
        inv[(size_t)(2 * src)] = code;
      else if (inv[(size_t)(2 * src + 1)] < 0)
        inv[(size_t)(2 * src + 1)] = code;
      else
        return C3L_OK;  // more than two padded elements per entry: no map
    }
  }
  int rc = upload_owned(ke.owned, &ke.inverse, inv.data(), inv.size());
  if (rc != C3L_OK) return rc;
  *map = ke.inverse;
  return C3L_OK;
}

double* c3l_internal_plan_workspace(c3l_plan* plan) { return plan ? reinterpret_cast<double*>(plan->ws) : nullptr; }

int plan_load_many(c3l_plan* const* plans, int32_t count, int32_t kind, const double* const* v, void* stream,
                   bool scatter) {
  if (plans == nullptr || v == nullptr || count < 1 || count > kMaxTwins)
    return fail(C3L_EINVAL, "c3l_plan_load: invalid argument");
  if (kind < 0 || kind >= kKinds) return fail(C3L_EINVAL, "c3l_plan_load: invalid kind");
  for (int32_t i = 0; i < count; ++i) {
    if (plans[i] == nullptr || v[i] == nullptr) return fail(C3L_EINVAL, "c3l_plan_load: invalid argument");
    if (i > 0 && (plans[i]->digest != plans[0]->digest || plans[i]->kind_bits[kind] != 64))
      return fail(C3L_EINVAL, "c3l_plan_load: plans loaded together must be FP64 twins");
  }
  c3l_plan* plan = plans[0];
  hipStream_t s = static_cast<hipStream_t>(stream);
  KindExec& ke = plan->kinds[kind];
  if (!ke.built) {
    std::vector<const PlanBase*> mine;
    for (const PlanBase& b : plan->bases)
      if (b.kind == kind && b.count > 0) mine.push_back(&b);
    std::sort(mine.begin(), mine.end(), [](const PlanBase* a, const PlanBase* b) { return a->offset < b->offset; });
    for (size_t i = 0; i < mine.size();) {
      size_t j = i + 1;
      while (j < mine.size() && mine[j]->offset == mine[j - 1]->offset + mine[j - 1]->count) ++j;
      std::vector<int64_t> src;
      std::vector<int8_t> sign;
      for (size_t k = i; k < j; ++k) {
        src.insert(src.end(), mine[k]->src.begin(), mine[k]->src.end());
        sign.insert(sign.end(), mine[k]->sign.begin(), mine[k]->sign.end());
      }
      BaseRun run;
      run.offset = mine[i]->offset;
      run.count = (int64_t)src.size();
      int rc = upload_owned(ke.owned, &run.src, src.data(), src.size());
      if (rc == C3L_OK) rc = upload_owned(ke.owned, &run.sign, sign.data(), sign.size());
      if (rc != C3L_OK) return rc;
      ke.runs.push_back(run);
      i = j;
    }
    int rc = build_trees(plan, kind, ke);
    if (rc != C3L_OK) return rc;
    ke.built = true;
  }
  const bool narrow = plan->kind_bits[kind] == 32;
  Twins<const double2*> vv{};
  Twins<double2*> ws{};
  Twins<float2*> ws32{};
  for (int32_t i = 0; i < count; ++i) {
    vv.p[i] = reinterpret_cast<const double2*>(v[i]);
    ws.p[i] = plans[i]->ws;
    ws32.p[i] = plans[i]->ws32;
  }
  const dim3 inst(1, 1, (unsigned)count);
  if (!scatter && narrow) return fail(C3L_EINVAL, "c3l_plan_load: prepared loads are FP64 only");
  for (const BaseRun& r : ke.runs) {
    if (!scatter) break;
    Twins<double2*> pad = ws;
    Twins<float2*> pad32 = ws32;
    for (int32_t i = 0; i < count; ++i) {
      pad.p[i] += r.offset;
      if (narrow) pad32.p[i] += r.offset;
    }
    if (narrow)
      hipLaunchKernelGGL(scatter_padded_kernel<float2>, dim3(grid_for(r.count), 1, count), dim3(kThreads), 0, s, vv,
                         pad32, r.src, r.sign, r.count);
    else
      hipLaunchKernelGGL(scatter_padded_kernel<double2>, dim3(grid_for(r.count), 1, count), dim3(kThreads), 0, s, vv,
                         pad, r.src, r.sign, r.count);
    C3L_HIP_CHECK(hipGetLastError());
  }
  if (!narrow)
    for (int32_t i = 0; i < count; ++i) ++plans[i]->fixed_version;
  int32_t level = 0;
  for (const PlanLevel& l : plan->levels) {
    if (l.kind != kind || l.n_jobs == 0) continue;
    if (level++ >= ke.tree_level) break;
    const dim3 grid(grid_for(l.max_out), l.n_jobs, count);
    if (narrow)
      hipLaunchKernelGGL(contract_kernel<float2>, grid, dim3(kThreads), 0, s, ws32, plan->vectors32, l.jobs, plan->n);
    else
      hipLaunchKernelGGL(contract_kernel<double2>, grid, dim3(kThreads), 0, s, ws, plan->vectors, l.jobs, plan->n);
    C3L_HIP_CHECK(hipGetLastError());
  }
  if (ke.n_trees > 0) {
    const dim3 grid(ke.n_trees, 1, count);
    if (narrow)
      hipLaunchKernelGGL(contract_trees_kernel<float2>, grid, dim3(kThreads), 0, s, ws32, plan->vectors32,
                         ke.tree_jobs, ke.tree_start, ke.tree_barrier, plan->n);
    else
      hipLaunchKernelGGL(contract_trees_kernel<double2>, grid, dim3(kThreads), 0, s, ws, plan->vectors, ke.tree_jobs,
                         ke.tree_start, ke.tree_barrier, plan->n);
    C3L_HIP_CHECK(hipGetLastError());
  }
  return C3L_OK;
}

}  // extern "C"

namespace {


// Lanes per coordinate: split the entries of small blocks over up to four waves
// so that a program launch keeps at least ~2 workgroups per CU busy.
int lanes_log2_for(const PlanBlock& b, const ProgramShape& sh) {
  int lg = 0;
  while (lg < sh.lanes_max && (b.count + (256 >> lg) - 1) / (256 >> lg) < sh.lanes_target && b.n_entries >= (2 << lg))
    ++lg;
  return lg;
}

int build_program(c3l_plan* plan, int32_t program);

// Tuned shapes by (plan digest, program, precision bits), shared by every plan built from the
// same arrays (twins, repeated instances); g_tune_gen changes whenever an entry does.
std::mutex g_tune_mutex;
std::map<std::tuple<uint64_t, int32_t, int32_t>, ProgramShape> g_tuned;
int64_t g_tune_gen = 0;

ProgramShape wanted_shape(c3l_plan* plan, int32_t program) {
  ProgramExec& pe = plan->programs[program];
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  if (pe.tune_gen == g_tune_gen) return pe.shape;
  const auto it = g_tuned.find(std::make_tuple(plan->digest, program, plan->program_bits[program]));
  return it == g_tuned.end() ? ProgramShape{} : it->second;
}

void set_tuned(c3l_plan* plan, int32_t program, const ProgramShape& sh) {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  g_tuned[std::make_tuple(plan->digest, program, plan->program_bits[program])] = sh;
  ++g_tune_gen;
}

// Frees a program's images (after the device has finished every launch that may read them).
int release_mfma_image(ProgramExec& pe) {
  if (!pe.owned_mfma.empty()) {
    C3L_HIP_CHECK(hipDeviceSynchronize());
    for (void* p : pe.owned_mfma) C3L_HIP_CHECK(hipFree(p));
    pe.owned_mfma.clear();
  }
  pe.mfma_built = false;
  pe.mfma_tiles = 0;
  pe.tiles32_two = pe.tiles32_general = 0;
  pe.tiles32 = nullptr;
  return C3L_OK;
}

int release_program_image(ProgramExec& pe) {
  int rc = release_mfma_image(pe);
  if (rc != C3L_OK) return rc;
  if (!pe.owned.empty()) {
    C3L_HIP_CHECK(hipDeviceSynchronize());
    for (void* p : pe.owned) C3L_HIP_CHECK(hipFree(p));
    pe.owned.clear();
  }
  pe.built = false;
  pe.five = false;
  pe.max_dest = -1;
  pe.tiles_two = pe.tiles_general = 0;
  pe.descs = nullptr;
  pe.tiles = nullptr;
  pe.desc_block.clear();
  pe.host_tiles.clear();
  return C3L_OK;
}

// Builds the program's gather image with its current (tuned or default) shape, rebuilding
// it when the tuned shape has changed since it was built.
int ensure_program(c3l_plan* plan, int32_t program) {
  ProgramExec& pe = plan->programs[program];
  const ProgramShape want = wanted_shape(plan, program);
  if (!(want == pe.shape)) {
    const ProgramShape had = pe.shape;
    const bool gather_same = had.lanes_max == want.lanes_max && had.lanes_target == want.lanes_target &&
                             had.reps_max == want.reps_max && had.reps_target == want.reps_target &&
                             had.xcd == want.xcd;
    int rc = gather_same ? release_mfma_image(pe) : release_program_image(pe);
    if (rc != C3L_OK) return rc;
    pe.shape = want;
  }
  {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    pe.tune_gen = g_tune_gen;
  }
  return build_program(plan, program);
}

int build_program(c3l_plan* plan, int32_t program) {
  ProgramExec& pe = plan->programs[program];
  if (pe.built) return C3L_OK;
  const ProgramShape& sh = pe.shape;
  std::vector<size_t> members;
  for (size_t i = 0; i < plan->blocks.size(); ++i)
    if (plan->blocks[i].program == program && plan->blocks[i].count > 0) members.push_back(i);
  // long-running tiles (most entries per wave) first
  std::stable_sort(members.begin(), members.end(), [&](size_t a, size_t b) {
    const PlanBlock& A = plan->blocks[a];
    const PlanBlock& B = plan->blocks[b];
    return (A.n_entries >> lanes_log2_for(A, sh)) > (B.n_entries >> lanes_log2_for(B, sh));
  });
  std::vector<WedgeBlockDesc> descs;
  std::vector<int2> two, general;
  for (size_t i : members) {
    const PlanBlock& b = plan->blocks[i];
    WedgeBlockDesc d;
    d.count = b.count;
    d.n_entries = b.n_entries;
    d.lanes_log2 = lanes_log2_for(b, sh);
    d.reps = 1;
    if (b.two_factors && d.lanes_log2 == 0) {
      // coordinates per thread for large two-factor blocks (more independent gathers in
      // flight; kMaxReps = 2 measured best on MI355X for the n = 25 JVP)
      int reps = std::min(sh.reps_max, kMaxReps);
      while (reps > 1 && (b.count + 256 * reps - 1) / (256 * reps) < sh.reps_target) reps >>= 1;
      d.reps = reps;
    }
    d.dense_n = b.dense ? plan->n : 0;
    d.coord_bits = b.deg <= 4 ? 8 : 6;
    if (b.deg == 5) pe.five = true;
    d.dense_base = b.dense ? b.dense_base : 0;
    uint32_t* coord = nullptr;
    int64_t* dest = nullptr;
    WedgeEntry* entries = nullptr;
    WedgeFactor* factors = nullptr;
    int rc = C3L_OK;
    if (!b.dense) {
      rc = upload_owned(pe.owned, &coord, b.coord.data(), b.coord.size());
      if (rc == C3L_OK) rc = upload_owned(pe.owned, &dest, b.dest.data(), b.dest.size());
    }
    if (rc == C3L_OK) rc = upload_owned(pe.owned, &entries, b.entries.data(), b.entries.size());
    if (rc == C3L_OK) rc = upload_owned(pe.owned, &factors, b.factors.data(), b.factors.size());
    if (rc != C3L_OK) return rc;
    d.coord = coord;
    d.dest = dest;
    d.entries = entries;
    d.factors = factors;
    const int per = (256 >> d.lanes_log2) * d.reps;
    std::vector<int2>& tiles = b.two_factors ? two : general;
    for (int64_t first = 0; first < b.count; first += per) tiles.push_back(make_int2((int)descs.size(), (int)first));
    descs.push_back(d);
    pe.desc_block.push_back((int32_t)i);
    if (b.max_dest > pe.max_dest) pe.max_dest = b.max_dest;
  }
  // Workgroup w runs on XCD w % 8 (round-robin dispatch).  Each block's tiles are cut
  // into 8 contiguous coordinate ranges, range x is queued for XCD x and the queues are
  // interleaved, so every XCD's L2 serves one slice of the block's larger factor tensors.
  auto xcd_order = [](std::vector<int2>& list) {
    constexpr int kXcds = 8;
    std::vector<std::vector<int2>> queue(kXcds);
    for (size_t i = 0; i < list.size();) {
      size_t j = i;
      while (j < list.size() && list[j].x == list[i].x) ++j;
      const size_t count = j - i;
      for (size_t k = 0; k < count; ++k) queue[(k * kXcds) / count].push_back(list[i + k]);
      i = j;
    }
    std::vector<size_t> head(kXcds, 0);
    std::vector<int2> out;
    out.reserve(list.size());
    while (out.size() < list.size()) {
      for (int x = 0; x < kXcds && out.size() < list.size(); ++x) {
        int q = x;
        if (head[q] == queue[q].size()) {  // this queue is drained: take from the fullest one
          size_t best = 0;
          for (int y = 0; y < kXcds; ++y)
            if (queue[y].size() - head[y] > best) best = queue[y].size() - head[y], q = y;
        }
        out.push_back(queue[q][head[q]++]);
      }
    }
    list.swap(out);
  };
  if (sh.xcd) {
    xcd_order(two);
    xcd_order(general);
  }
  pe.tiles_two = (int32_t)two.size();
  pe.tiles_general = (int32_t)general.size();
  std::vector<int2> tiles(two);
  tiles.insert(tiles.end(), general.begin(), general.end());
  int rc = upload_owned(pe.owned, &pe.descs, descs.data(), descs.size());
  if (rc == C3L_OK) rc = upload_owned(pe.owned, &pe.tiles, tiles.data(), tiles.size());
  if (rc != C3L_OK) return rc;
  pe.host_tiles = tiles;
  pe.built = true;
  return C3L_OK;
}

// Splits the entries of a rank-4 two-atom block into the views of tile_program_mfma_kernel
// with the block's operand layouts (byte strides in the FP32 workspace), and collects its
// rank-4 x scalar entries; false when an entry's atoms do not split the four positions
// between them or the block has more layouts than the kernel holds.
bool tile_entries_for(const PlanBlock& b, std::vector<TileRule>& rules, std::vector<TileEntry>& entries,
                      std::vector<float2>& coeffs, int32_t (&view_start)[kTileViews + 1],
                      std::vector<TileScalarEntry>& scalars) {
  auto positions = [](const WedgeFactor& f) {
    int m = 0;
    for (int k = 0; k < 4; ++k)
      if (f.stride[k] != 0) m |= 1 << k;
    return m;
  };
  std::map<std::array<uint32_t, 7>, int32_t> rule_of;
  auto rule = [&](const int (&pos)[3], int b_side, const WedgeFactor& f) -> int32_t {
    std::array<uint32_t, 7> key{};
    for (int k = 0; k < 3; ++k) {
      key[(size_t)k] = (uint32_t)pos[k];
      key[(size_t)k + 3] = 8u * (uint32_t)f.stride[pos[k]];
    }
    key[6] = (uint32_t)b_side;
    auto it = rule_of.find(key);
    if (it != rule_of.end()) return it->second;
    if ((int)rules.size() == kTileMaxRules) return -1;
    rules.push_back(TileRule{{pos[0], pos[1], pos[2]}, {key[3], key[4], key[5]}, b_side});
    return rule_of[key] = (int32_t)rules.size() - 1;
  };
  std::vector<TileEntry> by_view[kTileViews];
  std::vector<float2> coeff_by_view[kTileViews];
  for (int32_t t = 0; t < b.n_entries; ++t) {
    const WedgeEntry& e = b.entries[(size_t)t];
    const WedgeFactor* f[2] = {&b.factors[(size_t)e.first], &b.factors[(size_t)e.first + 1]};
    const int m[2] = {positions(*f[0]), positions(*f[1])};
    if ((m[0] & m[1]) != 0 || (m[0] | m[1]) != 15) return false;
    const float2 coeff = make_float2((float)e.coeff.x, (float)e.coeff.y);
    if (m[0] == 0 || m[1] == 0) {
      const WedgeFactor& a = m[0] == 0 ? *f[1] : *f[0];
      const WedgeFactor& sc = m[0] == 0 ? *f[0] : *f[1];
      TileScalarEntry se{};
      se.coeff = coeff;
      se.a_off = (uint32_t)a.offset;
      for (int k = 0; k < 4; ++k) se.a_stride[k] = a.stride[k];
      se.scalar = (uint32_t)sc.offset;
      scalars.push_back(se);
      continue;
    }
    const int ai = m[0] & 1 ? 0 : 1;  // A covers position 0
    const WedgeFactor& A = *f[ai];
    const WedgeFactor& B = *f[1 - ai];
    const int c = __builtin_ctz((unsigned)m[1 - ai]);
    const int v = c - 1, pa = view_pa(v), pb = view_pb(v);
    const int apos[3] = {0, pa, pb}, bpos[3] = {c, pa, pb};
    const int32_t ra = rule(apos, 0, A), rb = rule(bpos, 1, B);
    if (ra < 0 || rb < 0) return false;
    by_view[v].push_back(TileEntry{8u * (uint32_t)A.offset, 8u * (uint32_t)B.offset,
                                   256u * (uint32_t)ra | (256u * (uint32_t)rb << 16), coeff.x});
    coeff_by_view[v].push_back(coeff);
  }
  for (int v = 0; v < kTileViews; ++v) {
    entries.insert(entries.end(), by_view[v].begin(), by_view[v].end());
    coeffs.insert(coeffs.end(), coeff_by_view[v].begin(), coeff_by_view[v].end());
    view_start[v + 1] = (int32_t)entries.size();
  }
  return true;
}

// The FP32 matrix-core image of a built program (see ProgramExec): rank-4 two-atom blocks
// whose coordinates fill at least a quarter of the 16x4x4x4 workgroup tiles they touch.
int build_program_mfma(c3l_plan* plan, int32_t program) {
  ProgramExec& pe = plan->programs[program];
  if (pe.mfma_built) return C3L_OK;
  const int n = plan->n;
  const int64_t n4 = (int64_t)n * n * n * n;
  const int T = (n + 3) / 4, G0 = (n + 15) / 16;
  const int64_t tiles_per_block = (int64_t)G0 * T * T * T;
  std::vector<char> on_tiles(pe.desc_block.size(), 0);
  std::vector<TileBlockDesc> descs;
  std::vector<TileRule> all_rules;
  std::vector<TileEntry> all_entries;
  std::vector<float> all_im;
  std::vector<TileScalarEntry> all_scalars;
  std::vector<int2> tile_list;
  // byte offsets into the mirror (atom offsets, and within a rank-4 atom) are 32-bit, destinations
  // and tuple indices 31-bit; operand lanes past the tensor need the zeroed tail of the mirror
  const bool fits = plan->mfma != 0 && plan->ws32_tail >= tile_tail_elements(n) &&
                    8 * (plan->workspace + plan->ws32_tail) < (int64_t(1) << 32) && 16 * n4 < (int64_t(1) << 32);
  for (size_t di = 0; fits && di < pe.desc_block.size(); ++di) {
    const PlanBlock& b = plan->blocks[(size_t)pe.desc_block[di]];
    if (b.dense || b.deg != 4 || !b.two_factors || b.n_entries == 0 || b.max_dest >= (int64_t(1) << 31)) continue;
    if (plan->mfma == 1 && b.n_entries < pe.shape.mfma_min) continue;
    std::vector<TileRule> rules;
    std::vector<TileEntry> entries;
    std::vector<float2> coeffs;
    std::vector<TileScalarEntry> scalars;
    TileBlockDesc d{};
    if (!tile_entries_for(b, rules, entries, coeffs, d.view_start, scalars)) continue;
    std::vector<int32_t> dest((size_t)n4, -1);
    std::vector<char> used((size_t)tiles_per_block, 0);
    for (int64_t k = 0; k < b.count; ++k) {
      const uint32_t c = b.coord[(size_t)k];
      const int i0 = c & 255u, i1 = (c >> 8) & 255u, i2 = (c >> 16) & 255u, i3 = c >> 24;
      dest[(size_t)(i0 + (int64_t)n * (i1 + (int64_t)n * (i2 + (int64_t)n * i3)))] = (int32_t)b.dest[(size_t)k];
      used[(size_t)(i0 / 16 + (int64_t)G0 * (i1 / 4 + (int64_t)T * (i2 / 4 + (int64_t)T * (i3 / 4))))] = 1;
    }
    int64_t volume = 0;  // index tuples inside the tensor of the touched tiles
    for (int64_t t = 0; t < tiles_per_block; ++t) {
      if (!used[(size_t)t]) continue;
      const int64_t g0 = t % G0, t1 = (t / G0) % T, t2 = (t / (G0 * T)) % T, t3 = t / ((int64_t)G0 * T * T);
      volume += std::min<int64_t>(16, n - 16 * g0) * std::min<int64_t>(4, n - 4 * t1) *
                std::min<int64_t>(4, n - 4 * t2) * std::min<int64_t>(4, n - 4 * t3);
    }
    if (4 * b.count < volume) continue;  // too sparse for whole tiles
    int32_t* dst = nullptr;
    int rc = upload_owned(pe.owned_mfma, &dst, dest.data(), dest.size());
    if (rc != C3L_OK) return rc;
    d.n = n;
    d.per_dim = T;
    for (int v = 0; v <= kTileViews; ++v) d.view_start[v] += (int32_t)all_entries.size();
    all_entries.insert(all_entries.end(), entries.begin(), entries.end());
    for (const float2& c : coeffs) all_im.push_back(c.y);
    d.rule_start = (int32_t)all_rules.size();
    all_rules.insert(all_rules.end(), rules.begin(), rules.end());
    d.rule_end = (int32_t)all_rules.size();
    d.scalar_start = (int32_t)all_scalars.size();
    all_scalars.insert(all_scalars.end(), scalars.begin(), scalars.end());
    d.scalar_end = (int32_t)all_scalars.size();
    d.dest = dst;
    // workgroups go round-robin to the eight XCDs: give each XCD one contiguous eighth of the
    // block's touched tiles in memory order (neighbouring tiles share atom slices in its L2)
    std::vector<int32_t> seq;
    for (int64_t t = 0; t < tiles_per_block; ++t)
      if (used[(size_t)t]) seq.push_back((int32_t)t);
    constexpr int kXcds = 8;
    const int64_t count = (int64_t)seq.size();
    const int64_t chunk = (count + kXcds - 1) / kXcds;
    for (int64_t q = 0; q < chunk; ++q)
      for (int xcd = 0; xcd < kXcds; ++xcd) {
        const int64_t t = (int64_t)xcd * chunk + q;
        if (t < count) tile_list.push_back(make_int2((int)descs.size(), seq[(size_t)t]));
      }
    descs.push_back(d);
    on_tiles[di] = 1;
  }
  std::vector<int2> two, general;
  for (int32_t k = 0; k < pe.tiles_two + pe.tiles_general; ++k) {
    const int2 t = pe.host_tiles[(size_t)k];
    if (on_tiles[(size_t)t.x]) continue;
    (k < pe.tiles_two ? two : general).push_back(t);
  }
  pe.tiles32_two = (int32_t)two.size();
  pe.tiles32_general = (int32_t)general.size();
  two.insert(two.end(), general.begin(), general.end());
  int rc = upload_owned(pe.owned_mfma, &pe.tiles32, two.data(), two.size());
  if (rc == C3L_OK) rc = upload_owned(pe.owned_mfma, &pe.mfma_descs, descs.data(), descs.size());
  if (rc == C3L_OK) rc = upload_owned(pe.owned_mfma, &pe.mfma_tile_list, tile_list.data(), tile_list.size());
  if (rc == C3L_OK) rc = upload_owned(pe.owned_mfma, &pe.mfma_rules, all_rules.data(), all_rules.size());
  if (rc == C3L_OK) rc = upload_owned(pe.owned_mfma, &pe.mfma_entries, all_entries.data(), all_entries.size());
  if (rc == C3L_OK) rc = upload_owned(pe.owned_mfma, &pe.mfma_coeff_im, all_im.data(), all_im.size());
  if (rc == C3L_OK) rc = upload_owned(pe.owned_mfma, &pe.mfma_scalars, all_scalars.data(), all_scalars.size());
  if (rc != C3L_OK) return rc;
  pe.mfma_tiles = (int32_t)tile_list.size();
  pe.mfma_real = true;
  for (float im : all_im) pe.mfma_real = pe.mfma_real && im == 0.f;
  pe.mfma_built = true;
  return C3L_OK;
}

// Workspace ranges read by FP32 programs that FP64 loads/programs produce (constant
// atoms, amplitude atoms, creation tensors): narrowed into ws32 before an FP32 program
// runs whenever the FP64 side changed.
int narrow_fixed_atoms(c3l_plan* plan, hipStream_t s) {
  if (!plan->ranges_built) {
    std::vector<std::pair<int64_t, int64_t>> produced;  // FP32-produced [begin, end)
    for (const PlanBase& b : plan->bases)
      if (plan->kind_bits[b.kind] == 32) produced.push_back({b.offset, b.offset + b.count});
    for (const PlanLevel& l : plan->levels) {
      if (plan->kind_bits[l.kind] != 32) continue;
      for (const ContractJob& j : l.host_jobs) produced.push_back({j.out_offset, j.out_offset + ipow(plan->n, j.par_rank - 1)});
    }
    std::vector<std::pair<int64_t, int64_t>> need;
    for (const PlanBlock& b : plan->blocks) {
      if (plan->program_bits[b.program] != 32) continue;
      for (const WedgeFactor& f : b.factors) {
        int rank = 0;
        int64_t size = 1;
        for (int k = 0; k < 4; ++k)
          if (f.stride[k] != 0) ++rank;
        size = ipow(plan->n, rank);
        bool mine = false;
        for (const auto& p : produced)
          if (f.offset < p.second && f.offset + size > p.first) mine = true;
        if (!mine) need.push_back({f.offset, f.offset + size});
      }
    }
    std::sort(need.begin(), need.end());
    std::vector<int64_t> flat;
    int64_t max_count = 0;
    for (size_t i = 0; i < need.size();) {
      int64_t b = need[i].first, e = need[i].second;
      size_t j = i + 1;
      while (j < need.size() && need[j].first <= e) e = std::max(e, need[j++].second);
      flat.push_back(b);
      flat.push_back(e - b);
      max_count = std::max(max_count, e - b);
      i = j;
    }
    plan->n_ranges = (int32_t)(flat.size() / 2);
    plan->max_range = max_count;
    int rc = upload(plan, &plan->ranges, flat.data(), (int64_t)flat.size());
    if (rc != C3L_OK) return rc;
    plan->ranges_built = true;
  }
  if (plan->narrowed_version == plan->fixed_version || plan->n_ranges == 0) return C3L_OK;
  hipLaunchKernelGGL(narrow_ranges_kernel, dim3(grid_for(plan->max_range), plan->n_ranges), dim3(kThreads), 0, s,
                     plan->ws, plan->ws32, plan->ranges);
  C3L_HIP_CHECK(hipGetLastError());
  plan->narrowed_version = plan->fixed_version;
  return C3L_OK;
}

}  // namespace

// ---- allocation-free complex multidot (c3l_zmultidot_noalloc) --------------------------
// Its partial sums live in this file-scope device array: one copy per GPU, held until the
// process ends.  In host code `scratch` is only a symbol: the device address of each GPU's
// copy comes from hipGetSymbolAddress, taken once per device with that device current (on
// the entry point's first call there) and kept in scratch_ptr[].
constexpr int64_t kNoAllocScratchDoubles = int64_t(1) << 20;  // 8 MiB per GPU
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {

constexpr int kMaxNoAllocDevices = 64;
std::mutex scratch_mu;
double* scratch_ptr[kMaxNoAllocDevices] = {};

// Device address of the current GPU's copy of `scratch` (looked up once per device).
int current_scratch(double** out) {
  int device = -1;
  C3L_HIP_CHECK(hipGetDevice(&device));
  if (device < 0 || device >= kMaxNoAllocDevices)
    return fail(C3L_EINVAL, "c3l_zmultidot_noalloc: device index beyond the supported range");
  std::lock_guard<std::mutex> guard(scratch_mu);
  if (scratch_ptr[device] == nullptr) {
    void* ptr = nullptr;
    C3L_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (ptr == nullptr) return fail(C3L_EHIP, "c3l_zmultidot_noalloc: no device address for the scratch array");
    scratch_ptr[device] = static_cast<double*>(ptr);
  }
  *out = scratch_ptr[device];
  return C3L_OK;
}

}  // namespace

extern "C" {

int c3l_plan_set_matrix_cores(c3l_plan* plan, int32_t mode) {
  if (plan == nullptr || mode < 0 || mode > 2) return fail(C3L_EINVAL, "c3l_plan_set_matrix_cores: invalid argument");
  if (mode != plan->mfma) {
    plan->mfma = mode;
    for (ProgramExec& pe : plan->programs) {  // rebuilt at the next FP32 launch
      int rc = release_mfma_image(pe);
      if (rc != C3L_OK) return rc;
    }
  }
  return C3L_OK;
}

// Autotuner of the program images' launch shapes: times `reps` launches per candidate on x
// into out (after two untimed ones) by coordinate descent over, in order, the matrix-core
// threshold of an FP32 program, lanes per coordinate and the workgroup target of the lane split,
// coordinates per thread and its workgroup target, and the XCD slicing, starting from the default
// heuristics.  Knobs that change the summation order of a coordinate's entries (matrix cores,
// lane split) replace the best only at a 10 % gain, so that timing noise does not change results;
// the others at 2 %.  The winner is kept for every plan with this plan's digest, program and
// precision (images are rebuilt at their next launch).
int c3l_plan_autotune(c3l_plan* plan, int32_t program, const double* diag, const double* x, double* out, int32_t reps,
                      void* stream, double* ms_default, double* ms_best, int32_t* shape_out) {
  if (plan == nullptr || program < 0 || program >= kPrograms || out == nullptr || reps < 1 || reps > 1000 ||
      (diag == nullptr) != (x == nullptr) || ms_default == nullptr || ms_best == nullptr)
    return fail(C3L_EINVAL, "c3l_plan_autotune: invalid argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  C3L_HIP_CHECK(hipEventCreate(&e0));
  C3L_HIP_CHECK(hipEventCreate(&e1));
  auto measure = [&](const ProgramShape& sh, double* ms) -> int {
    set_tuned(plan, program, sh);
    for (int w = 0; w < 2; ++w) {
      const int rc = c3l_internal_plan_apply_many(&plan, 1, program, diag, &x, &out, stream);
      if (rc != C3L_OK) return rc;
    }
    C3L_HIP_CHECK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) {
      const int rc = c3l_internal_plan_apply_many(&plan, 1, program, diag, &x, &out, stream);
      if (rc != C3L_OK) return rc;
    }
    C3L_HIP_CHECK(hipEventRecord(e1, s));
    C3L_HIP_CHECK(hipEventSynchronize(e1));
    float elapsed = 0.f;
    C3L_HIP_CHECK(hipEventElapsedTime(&elapsed, e0, e1));
    *ms = (double)elapsed / reps;
    return C3L_OK;
  };
  ProgramShape best{};
  double t_best = 0.0;
  int rc = measure(best, &t_best);
  *ms_default = t_best;
  auto descend = [&](int32_t ProgramShape::*knob, std::initializer_list<int32_t> values, double gain) -> int {
    const int32_t start = best.*knob;
    for (int32_t v : values) {
      if (v == start) continue;
      ProgramShape cand = best;
      cand.*knob = v;
      double t = 0.0;
      const int r = measure(cand, &t);
      if (r != C3L_OK) return r;
      if (t < (1.0 - gain) * t_best) {
        best = cand;
        t_best = t;
      }
    }
    return C3L_OK;
  };
  constexpr double kReorders = 0.10, kKeeps = 0.02;
  if (rc == C3L_OK && plan->program_bits[program] == 32 && plan->mfma == 1)
    rc = descend(&ProgramShape::mfma_min, {kTileMinEntries, INT_MAX, 32, 1}, kReorders);
  if (rc == C3L_OK) rc = descend(&ProgramShape::lanes_max, {2, 1, 0}, kReorders);
  if (rc == C3L_OK) rc = descend(&ProgramShape::lanes_target, {512, 256, 1024, 2048}, kReorders);
  if (rc == C3L_OK) rc = descend(&ProgramShape::reps_max, {2, 1}, kKeeps);
  if (rc == C3L_OK) rc = descend(&ProgramShape::reps_target, {512, 256, 1024, 2048}, kKeeps);
  if (rc == C3L_OK) rc = descend(&ProgramShape::xcd, {1, 0}, kKeeps);
  set_tuned(plan, program, best);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != C3L_OK) return rc;
  *ms_best = t_best;
  if (shape_out != nullptr) {
    shape_out[0] = best.lanes_max;
    shape_out[1] = best.lanes_target;
    shape_out[2] = best.reps_max;
    shape_out[3] = best.reps_target;
    shape_out[4] = best.xcd;
    shape_out[5] = best.mfma_min;
  }
  return C3L_OK;
}

int c3l_plan_matrix_core_tiles(c3l_plan* plan, int32_t program, int32_t* tiles) {
  if (plan == nullptr || tiles == nullptr || program < 0 || program >= kPrograms)
    return fail(C3L_EINVAL, "c3l_plan_matrix_core_tiles: invalid argument");
  const ProgramExec& pe = plan->programs[program];
  *tiles = pe.mfma_built ? pe.mfma_tiles : 0;
  return C3L_OK;
}

int c3l_plan_set_precision(c3l_plan* plan, int32_t kind, int32_t program, int32_t bits) {
  if (plan == nullptr || kind < 0 || kind >= kKinds || program < 0 || program >= kPrograms || (bits != 32 && bits != 64))
    return fail(C3L_EINVAL, "c3l_plan_set_precision: invalid argument");
  if (bits == 32 && plan->ws32 == nullptr) {
    void* p = nullptr;
    // a zeroed tail past the mirror: the matrix-core tiles' operand lanes past the tensor read it
    const int64_t tail = tile_tail_elements(plan->n);
    if (hipMalloc(&p, sizeof(float2) * (size_t)(plan->workspace + tail)) != hipSuccess)
      return fail(C3L_ENOMEM, "c3l_plan_set_precision: FP32 workspace allocation failed");
    plan->owned.push_back(p);
    plan->ws32 = static_cast<float2*>(p);
    C3L_HIP_CHECK(hipMemset(plan->ws32 + plan->workspace, 0, sizeof(float2) * (size_t)tail));
    plan->ws32_tail = tail;
    const int64_t nv = (int64_t)plan->n_vectors * plan->n;
    if (nv > 0) {
      if (hipMalloc(&p, sizeof(float2) * (size_t)nv) != hipSuccess)
        return fail(C3L_ENOMEM, "c3l_plan_set_precision: FP32 vector table allocation failed");
      plan->owned.push_back(p);
      plan->vectors32 = static_cast<float2*>(p);
      hipLaunchKernelGGL(narrow_kernel, dim3(grid_for(nv)), dim3(kThreads), 0, 0, plan->vectors, plan->vectors32, nv);
      C3L_HIP_CHECK(hipGetLastError());
      C3L_HIP_CHECK(hipDeviceSynchronize());
    }
  }
  plan->kind_bits[kind] = bits;
  plan->program_bits[program] = bits;
  plan->ranges_built = false;
  plan->narrowed_version = -1;
  return C3L_OK;
}

int c3l_plan_apply(c3l_plan* plan, int32_t program, const double* diag, const double* x, double* out, void* stream) {
  return c3l_internal_plan_apply_many(&plan, 1, program, diag, &x, &out, stream);
}

// Program `program` of twin plans plans[i] on x[i] into out[i], i < count, one launch
// per kernel variant (blockIdx.z = instance).  Several plans: FP64 programs with
// external outputs only, one shared diagonal.
int c3l_internal_plan_apply_many(c3l_plan* const* plans, int32_t count, int32_t program, const double* diag,
                                 const double* const* x, double* const* out, void* stream) {
  if (plans == nullptr || x == nullptr || out == nullptr || count < 1 || count > kMaxTwins || program < 0 ||
      program >= kPrograms)
    return fail(C3L_EINVAL, "c3l_plan_apply: invalid argument");
  for (int32_t i = 0; i < count; ++i) {
    if (plans[i] == nullptr || ((diag == nullptr) != (x[i] == nullptr)) || (out[i] == nullptr && diag != nullptr))
      return fail(C3L_EINVAL, "c3l_plan_apply: invalid argument");
    if (count > 1 && (plans[i]->digest != plans[0]->digest || plans[i]->program_bits[program] != 64 ||
                      out[i] == nullptr))
      return fail(C3L_EINVAL, "c3l_plan_apply: plans applied together must be FP64 twins with external outputs");
  }
  c3l_plan* plan = plans[0];
  int rc = ensure_program(plan, program);
  if (rc != C3L_OK) return rc;
  const ProgramExec& pe = plan->programs[program];
  if (pe.tiles_two + pe.tiles_general == 0) return C3L_OK;
  const bool internal = out[0] == nullptr;  // creation programs write into the workspace
  if (internal && pe.max_dest >= plan->workspace)
    return fail(C3L_EINVAL, "c3l_plan_apply: workspace program writes outside the workspace");
  hipStream_t s = static_cast<hipStream_t>(stream);
  PendingTiming t{program, count, nullptr, nullptr};
  if (plan->timing) {
    C3L_HIP_CHECK(hipEventCreate(&t.start));
    C3L_HIP_CHECK(hipEventCreate(&t.stop));
    C3L_HIP_CHECK(hipEventRecord(t.start, s));
  }
  const double2* dg = reinterpret_cast<const double2*>(diag);
  Twins<double2*> o{};
  Twins<const double2*> xx{}, ws{};
  for (int32_t i = 0; i < count; ++i) {
    o.p[i] = internal ? plan->ws : reinterpret_cast<double2*>(out[i]);
    xx.p[i] = reinterpret_cast<const double2*>(x[i]);
    ws.p[i] = plans[i]->ws;
  }
  const dim3 two(pe.tiles_two, 1, count), general(pe.tiles_general, 1, count);
  if (plan->program_bits[program] == 32) {
    if (internal) return fail(C3L_EINVAL, "c3l_plan_apply: workspace programs run in FP64 only");
    rc = narrow_fixed_atoms(plan, s);
    if (rc != C3L_OK) return rc;
    const Twins<const float2*> ws32 = twin<const float2*>(plan->ws32);
    rc = build_program_mfma(plan, program);
    if (rc != C3L_OK) return rc;
    if (pe.mfma_tiles > 0) {
      if (pe.mfma_real)
        hipLaunchKernelGGL(tile_program_mfma_kernel<true>, dim3(pe.mfma_tiles), dim3(kThreads), 0, s, plan->ws32, o.p[0],
                           pe.mfma_descs, pe.mfma_rules, pe.mfma_entries, pe.mfma_coeff_im, pe.mfma_scalars,
                           pe.mfma_tile_list, dg, xx.p[0]);
      else
        hipLaunchKernelGGL(tile_program_mfma_kernel<false>, dim3(pe.mfma_tiles), dim3(kThreads), 0, s, plan->ws32,
                           o.p[0], pe.mfma_descs, pe.mfma_rules, pe.mfma_entries, pe.mfma_coeff_im, pe.mfma_scalars,
                           pe.mfma_tile_list, dg, xx.p[0]);
      C3L_HIP_CHECK(hipGetLastError());
    }
    if (pe.tiles32_two > 0) {
      if (pe.five)
        hipLaunchKernelGGL((wedge_program_kernel<true, float2, true>), dim3(pe.tiles32_two), dim3(kThreads), 0, s,
                           ws32, o, pe.descs, pe.tiles32, dg, xx);
      else
        hipLaunchKernelGGL((wedge_program_kernel<true, float2, false>), dim3(pe.tiles32_two), dim3(kThreads), 0, s,
                           ws32, o, pe.descs, pe.tiles32, dg, xx);
      C3L_HIP_CHECK(hipGetLastError());
    }
    if (pe.tiles32_general > 0) {
      if (pe.five)
        hipLaunchKernelGGL((wedge_program_kernel<false, float2, true>), dim3(pe.tiles32_general), dim3(kThreads), 0,
                           s, ws32, o, pe.descs, pe.tiles32 + pe.tiles32_two, dg, xx);
      else
        hipLaunchKernelGGL((wedge_program_kernel<false, float2, false>), dim3(pe.tiles32_general), dim3(kThreads), 0,
                           s, ws32, o, pe.descs, pe.tiles32 + pe.tiles32_two, dg, xx);
      C3L_HIP_CHECK(hipGetLastError());
    }
  } else {
    if (pe.tiles_two > 0) {
      if (pe.five)
        hipLaunchKernelGGL((wedge_program_kernel<true, double2, true>), two, dim3(kThreads), 0, s, ws, o, pe.descs,
                           pe.tiles, dg, xx);
      else
        hipLaunchKernelGGL((wedge_program_kernel<true, double2, false>), two, dim3(kThreads), 0, s, ws, o, pe.descs,
                           pe.tiles, dg, xx);
      C3L_HIP_CHECK(hipGetLastError());
    }
    if (pe.tiles_general > 0) {
      if (pe.five)
        hipLaunchKernelGGL((wedge_program_kernel<false, double2, true>), general, dim3(kThreads), 0, s, ws, o,
                           pe.descs, pe.tiles + pe.tiles_two, dg, xx);
      else
        hipLaunchKernelGGL((wedge_program_kernel<false, double2, false>), general, dim3(kThreads), 0, s, ws, o,
                           pe.descs, pe.tiles + pe.tiles_two, dg, xx);
      C3L_HIP_CHECK(hipGetLastError());
    }
    if (internal) ++plan->fixed_version;
  }
  if (plan->timing) {
    C3L_HIP_CHECK(hipEventRecord(t.stop, s));
    plan->pending.push_back(t);
  }
  return C3L_OK;
}

int c3l_tilde_unpack(const double* x, double* t, const int64_t* k0, const int64_t* k1, const double* ab, int64_t dim,
                     void* stream) {
  if (dim < 0 || (dim > 0 && (x == nullptr || t == nullptr || k0 == nullptr || k1 == nullptr || ab == nullptr)))
    return fail(C3L_EINVAL, "c3l_tilde_unpack: invalid argument");
  if (dim == 0) return C3L_OK;
  hipLaunchKernelGGL(tilde_unpack_kernel<double>, dim3(grid_for(dim)), dim3(kThreads), 0, static_cast<hipStream_t>(stream), x,
                     reinterpret_cast<double2*>(t), k0, k1, reinterpret_cast<const double4*>(ab), dim);
  C3L_HIP_CHECK(hipGetLastError());
  return C3L_OK;
}

int c3l_tilde_pack(const double* t, double* x, const int64_t* src, const double* w, int64_t real_dim, void* stream) {
  if (real_dim < 0 || (real_dim > 0 && (x == nullptr || t == nullptr || src == nullptr || w == nullptr)))
    return fail(C3L_EINVAL, "c3l_tilde_pack: invalid argument");
  if (real_dim == 0) return C3L_OK;
  hipLaunchKernelGGL(tilde_pack_kernel, dim3(grid_for(real_dim)), dim3(kThreads), 0, static_cast<hipStream_t>(stream),
                     reinterpret_cast<const double2*>(t), x, src, reinterpret_cast<const double2*>(w), real_dim);
  C3L_HIP_CHECK(hipGetLastError());
  return C3L_OK;
}

int c3l_tilde_pack_precondition(const double* t, double* x, const int64_t* src, const double* w, const double* inv_diag,
                                int64_t real_dim, void* stream) {
  if (real_dim < 0 || (real_dim > 0 && (x == nullptr || t == nullptr || src == nullptr || w == nullptr ||
                                        inv_diag == nullptr)))
    return fail(C3L_EINVAL, "c3l_tilde_pack_precondition: invalid argument");
  if (real_dim == 0) return C3L_OK;
  hipLaunchKernelGGL(tilde_pack_precondition_kernel, dim3(grid_for(real_dim)), dim3(kThreads), 0,
                     static_cast<hipStream_t>(stream), reinterpret_cast<const double2*>(t), x, src,
                     reinterpret_cast<const double2*>(w), reinterpret_cast<const double2*>(inv_diag), real_dim);
  C3L_HIP_CHECK(hipGetLastError());
  return C3L_OK;
}

// FP32 Krylov-basis helpers of the GMRES driver (library-internal).
int c3l_internal_tilde_unpack_f32(const float* x, double* t, const int64_t* k0, const int64_t* k1, const double* ab,
                                  int64_t dim, void* stream) {
  if (dim <= 0 || x == nullptr || t == nullptr || k0 == nullptr || k1 == nullptr || ab == nullptr)
    return fail(C3L_EINVAL, "c3l_internal_tilde_unpack_f32: invalid argument");
  hipLaunchKernelGGL(tilde_unpack_kernel<float>, dim3(grid_for(dim)), dim3(kThreads), 0,
                     static_cast<hipStream_t>(stream), x, reinterpret_cast<double2*>(t), k0, k1,
                     reinterpret_cast<const double4*>(ab), dim);
  C3L_HIP_CHECK(hipGetLastError());
  return C3L_OK;
}

int c3l_internal_narrow_scale(const double* x, float* y, double a, int64_t dim, void* stream) {
  if (dim <= 0 || x == nullptr || y == nullptr) return fail(C3L_EINVAL, "c3l_internal_narrow_scale: invalid argument");
  hipLaunchKernelGGL(narrow_scale_kernel, dim3(grid_for(dim)), dim3(kThreads), 0, static_cast<hipStream_t>(stream), x, y,
                     a, dim);
  C3L_HIP_CHECK(hipGetLastError());
  return C3L_OK;
}

int c3l_internal_rmultiaxpy_f32(const float* V, int64_t ld, int32_t ncols, const double* alpha, double scale,
                                double* w, int64_t dim, void* stream) {
  if (ncols <= 0 || dim <= 0 || ld < dim || V == nullptr || alpha == nullptr || w == nullptr)
    return fail(C3L_EINVAL, "c3l_internal_rmultiaxpy_f32: invalid argument");
  hipLaunchKernelGGL(rmultiaxpy_kernel<float>, dim3((dim + 63) / 64), dim3(kThreads), 0,
                     static_cast<hipStream_t>(stream), V, ld, ncols, alpha, scale, w, dim);
  C3L_HIP_CHECK(hipGetLastError());
  return C3L_OK;
}

int64_t c3l_reduction_chunks(int64_t dim) { return chunks_for(dim); }

int c3l_rmultidot(const double* V, int64_t ld, int32_t ncols, const double* w, int64_t dim, double* scratch,
                  double* out, void* stream) {
  if (ncols <= 0 || dim <= 0 || ld < dim || V == nullptr || w == nullptr || scratch == nullptr || out == nullptr)
    return fail(C3L_EINVAL, "c3l_rmultidot: invalid argument");
  const int rows = dot_rows_for(ncols);
  const int64_t chunks = (dim + rows - 1) / rows;
  hipStream_t s = static_cast<hipStream_t>(stream);
  dispatch_rows(rows, [&](auto R) {
    hipLaunchKernelGGL((rmultidot_partial_kernel<decltype(R)::value>), dim3(chunks, (ncols + 3) / 4), dim3(kThreads), 0,
                       s, V, ld, ncols, (const double*)nullptr, w, dim, scratch, chunks);
  });
  C3L_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(rfinish_kernel, dim3(ncols), dim3(kThreads), 0, s, scratch, chunks, out);
  C3L_HIP_CHECK(hipGetLastError());
  return C3L_OK;
}

// out[k] = V_k . w for k < ncols and out[ncols] = extra . w, one pass over w (library-internal).
int c3l_internal_rmultidot_extra(const double* V, int64_t ld, int32_t ncols, const double* extra, const double* w,
                                 int64_t dim, double* scratch, double* out, void* stream) {
  if (ncols < 0 || dim <= 0 || ld < dim || (ncols > 0 && V == nullptr) || extra == nullptr || w == nullptr ||
      scratch == nullptr || out == nullptr)
    return fail(C3L_EINVAL, "c3l_internal_rmultidot_extra: invalid argument");
  const int total = ncols + 1;
  const int rows = dot_rows_for(total);
  const int64_t chunks = (dim + rows - 1) / rows;
  hipStream_t s = static_cast<hipStream_t>(stream);
  dispatch_rows(rows, [&](auto R) {
    hipLaunchKernelGGL((rmultidot_partial_kernel<decltype(R)::value>), dim3(chunks, (total + 3) / 4), dim3(kThreads), 0,
                       s, V, ld, total, extra, w, dim, scratch, chunks);
  });
  C3L_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(rfinish_kernel, dim3(total), dim3(kThreads), 0, s, scratch, chunks, out);
  C3L_HIP_CHECK(hipGetLastError());
  return C3L_OK;
}

int c3l_rmultiaxpy(const double* V, int64_t ld, int32_t ncols, const double* alpha, double scale, double* w,
                   int64_t dim, void* stream) {
  if (ncols <= 0 || dim <= 0 || ld < dim || V == nullptr || alpha == nullptr || w == nullptr)
    return fail(C3L_EINVAL, "c3l_rmultiaxpy: invalid argument");
  hipLaunchKernelGGL(rmultiaxpy_kernel<double>, dim3((dim + 63) / 64), dim3(kThreads), 0, static_cast<hipStream_t>(stream), V, ld,
                     ncols, alpha, scale, w, dim);
  C3L_HIP_CHECK(hipGetLastError());
  return C3L_OK;
}

int c3l_tilde_precondition(const double* x, double* y, const int64_t* k0, const int64_t* k1, const double* ab,
                           const int64_t* src, const double* w, const double* inv_diag, int64_t real_dim,
                           void* stream) {
  if (real_dim <= 0 || x == nullptr || y == nullptr || k0 == nullptr || k1 == nullptr || ab == nullptr ||
      src == nullptr || w == nullptr || inv_diag == nullptr || x == y)
    return fail(C3L_EINVAL, "c3l_tilde_precondition: invalid argument");
  hipLaunchKernelGGL(tilde_precondition_kernel, dim3(grid_for(real_dim)), dim3(kThreads), 0,
                     static_cast<hipStream_t>(stream), x, y, k0, k1, reinterpret_cast<const double4*>(ab), src,
                     reinterpret_cast<const double2*>(w), reinterpret_cast<const double2*>(inv_diag), real_dim);
  C3L_HIP_CHECK(hipGetLastError());
  return C3L_OK;
}

int c3l_rscale(const double* x, double* y, double a, int64_t dim, void* stream) {
  if (dim <= 0 || x == nullptr || y == nullptr) return fail(C3L_EINVAL, "c3l_rscale: invalid argument");
  hipLaunchKernelGGL(rscale_kernel, dim3(grid_for(dim)), dim3(kThreads), 0, static_cast<hipStream_t>(stream), x, y, a,
                     dim);
  C3L_HIP_CHECK(hipGetLastError());
  return C3L_OK;
}

int c3l_raxpby(double a, const double* x, double b, double* y, int64_t dim, void* stream) {
  if (dim <= 0 || x == nullptr || y == nullptr) return fail(C3L_EINVAL, "c3l_raxpby: invalid argument");
  hipLaunchKernelGGL(raxpby_kernel, dim3(grid_for(dim)), dim3(kThreads), 0, static_cast<hipStream_t>(stream), a, x, b,
                     y, dim);
  C3L_HIP_CHECK(hipGetLastError());
  return C3L_OK;
}

int c3l_zmultidot(const double* V, int64_t ld, int32_t ncols, const double* w, int64_t dim, int32_t conjugate,
                  double* scratch, double* out, void* stream) {
  if (ncols <= 0 || dim <= 0 || ld < dim || V == nullptr || w == nullptr || scratch == nullptr || out == nullptr)
    return fail(C3L_EINVAL, "c3l_zmultidot: invalid argument");
  const int rows = dot_rows_for(ncols);
  const int64_t chunks = (dim + rows - 1) / rows;
  hipStream_t s = static_cast<hipStream_t>(stream);
  dispatch_rows(rows, [&](auto R) {
    hipLaunchKernelGGL((zmultidot_partial_kernel<decltype(R)::value>), dim3(chunks, (ncols + 3) / 4), dim3(kThreads), 0,
                       s, reinterpret_cast<const double2*>(V), ld, ncols, (const double2*)nullptr,
                       reinterpret_cast<const double2*>(w), dim, conjugate, reinterpret_cast<double2*>(scratch), chunks);
  });
  C3L_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(zfinish_kernel, dim3(ncols), dim3(kThreads), 0, s, reinterpret_cast<const double2*>(scratch), chunks,
                     reinterpret_cast<double2*>(out));
  C3L_HIP_CHECK(hipGetLastError());
  return C3L_OK;
}

// c3l_zmultidot with its partial sums in the file-scope `scratch` array instead of a caller
// buffer: no allocation of any kind.  Every partial entry read by the finishing kernel is
// written by the partial kernel of the same call, so nothing needs clearing between calls.
int c3l_zmultidot_noalloc(const double* V, int64_t ld, int32_t ncols, const double* w, int64_t dim,
                          int32_t conjugate, double* out, void* stream) {
  if (ncols <= 0 || dim <= 0 || ld < dim || V == nullptr || w == nullptr || out == nullptr)
    return fail(C3L_EINVAL, "c3l_zmultidot_noalloc: invalid argument");
  const int rows = dot_rows_for(ncols);
  const int64_t chunks = (dim + rows - 1) / rows;
  if ((int64_t)ncols * chunks * 2 > kNoAllocScratchDoubles)
    return fail(C3L_EINVAL, "c3l_zmultidot_noalloc: ncols * ceil(dim / rows) partial sums exceed the " +
                                std::to_string(kNoAllocScratchDoubles / 2) + " complex entries of the scratch array");
  double* partial = nullptr;
  const int rc = current_scratch(&partial);
  if (rc != C3L_OK) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  dispatch_rows(rows, [&](auto R) {
    hipLaunchKernelGGL((zmultidot_partial_kernel<decltype(R)::value>), dim3(chunks, (ncols + 3) / 4), dim3(kThreads), 0,
                       s, reinterpret_cast<const double2*>(V), ld, ncols, (const double2*)nullptr,
                       reinterpret_cast<const double2*>(w), dim, conjugate, reinterpret_cast<double2*>(partial), chunks);
  });
  C3L_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(zfinish_kernel, dim3(ncols), dim3(kThreads), 0, s, reinterpret_cast<const double2*>(partial), chunks,
                     reinterpret_cast<double2*>(out));
  C3L_HIP_CHECK(hipGetLastError());
  return C3L_OK;
}

// out[k] = conj?(V_k) . w for k < ncols and out[ncols] = extra . w (unconjugated), one pass over w
// (library-internal).
int c3l_internal_zmultidot_extra(const double* V, int64_t ld, int32_t ncols, const double* extra, const double* w,
                                 int64_t dim, int32_t conjugate, double* scratch, double* out, void* stream) {
  if (ncols < 0 || dim <= 0 || ld < dim || (ncols > 0 && V == nullptr) || extra == nullptr || w == nullptr ||
      scratch == nullptr || out == nullptr)
    return fail(C3L_EINVAL, "c3l_internal_zmultidot_extra: invalid argument");
  const int total = ncols + 1;
  const int rows = dot_rows_for(total);
  const int64_t chunks = (dim + rows - 1) / rows;
  hipStream_t s = static_cast<hipStream_t>(stream);
  dispatch_rows(rows, [&](auto R) {
    hipLaunchKernelGGL((zmultidot_partial_kernel<decltype(R)::value>), dim3(chunks, (total + 3) / 4), dim3(kThreads), 0,
                       s, reinterpret_cast<const double2*>(V), ld, total, reinterpret_cast<const double2*>(extra),
                       reinterpret_cast<const double2*>(w), dim, conjugate, reinterpret_cast<double2*>(scratch), chunks);
  });
  C3L_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(zfinish_kernel, dim3(total), dim3(kThreads), 0, s, reinterpret_cast<const double2*>(scratch), chunks,
                     reinterpret_cast<double2*>(out));
  C3L_HIP_CHECK(hipGetLastError());
  return C3L_OK;
}

int c3l_zmultiaxpy(const double* V, int64_t ld, int32_t ncols, const double* alpha, double scale, double* w,
                   int64_t dim, void* stream) {
  if (ncols <= 0 || dim <= 0 || ld < dim || V == nullptr || alpha == nullptr || w == nullptr)
    return fail(C3L_EINVAL, "c3l_zmultiaxpy: invalid argument");
  hipLaunchKernelGGL(zmultiaxpy_kernel, dim3((dim + 63) / 64), dim3(kThreads), 0, static_cast<hipStream_t>(stream),
                     reinterpret_cast<const double2*>(V), ld, ncols, reinterpret_cast<const double2*>(alpha), scale,
                     reinterpret_cast<double2*>(w), dim);
  C3L_HIP_CHECK(hipGetLastError());
  return C3L_OK;
}

int c3l_shifted_triangular_green(const double* T, int32_t m, const double* r, const double* w, const double* q,
                                 const double* shifts, int32_t n_shifts, double* green, double* last, void* stream) {
  if (m <= 0 || m > 4096 || n_shifts <= 0 || T == nullptr || r == nullptr || w == nullptr || q == nullptr ||
      shifts == nullptr || green == nullptr || last == nullptr)
    return fail(C3L_EINVAL, "c3l_shifted_triangular_green: invalid argument");
  hipLaunchKernelGGL(shifted_triangular_kernel, dim3(n_shifts), dim3(64), sizeof(double2) * m,
                     static_cast<hipStream_t>(stream), reinterpret_cast<const double2*>(T), m,
                     reinterpret_cast<const double2*>(r), reinterpret_cast<const double2*>(w),
                     reinterpret_cast<const double2*>(q), reinterpret_cast<const double2*>(shifts),
                     reinterpret_cast<double2*>(green), reinterpret_cast<double2*>(last));
  C3L_HIP_CHECK(hipGetLastError());
  return C3L_OK;
}

}  // extern "C"
