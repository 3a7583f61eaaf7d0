// Native Krylov drivers: restarted GMRES for the Newton correction and the
// shift-invariant Arnoldi resolvent, both running their whole loop in C++ so
// that the host only sees one call per linear solve / per Krylov basis.
//
// Both orthogonalise with the one-reduction delayed-reorthogonalisation step of
// ortho.hip: the operator is applied to the once-projected vector, and one
// reduction plus one update pass per step complete the previous Hessenberg
// column and project the new vector.  Column j-1 is therefore final after step j.
//
// GMRES: left-preconditioned restarted GMRES; the update pass writes the final
// Hessenberg column into mapped host memory once per iteration for the Givens
// update on the host (SciPy's tolerance
// control: target max(atol, rtol ||b||) on the true residual, checked after every
// restart, with an adaptively rescaled inner tolerance on the preconditioned
// residual).  The last column of a cycle is completed by a step without operator.
//
// Arnoldi: the Hessenberg matrix stays on the device; every shift z runs GMRES on
// the shared basis: its Givens QR of (z I - H) advances by up to eight columns per
// launch (eight lanes per shift, O(m) work per column) and the sink projection of its solution is
// updated incrementally, so the Green function and the exact residual of every
// shift are available at every step without any dense factorisation.  Up to four
// bases on twin plans advance in lockstep (one launch per kernel for all of them),
// and one host thread can advance several such runs, each on its own stream; the
// device stop flags reach the host through pinned memory and are polled, not waited
// for.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/c3l_hip.h"
#include "ortho.h"

// Shared error slot and FP32-basis helpers of the library (defined in c3l_hip.hip).
extern "C" int c3l_internal_fail(int code, const char* message);
extern "C" int c3l_internal_tilde_unpack_f32(const float* x, double* t, const int64_t* k0, const int64_t* k1,
                                             const double* ab, int64_t dim, void* stream);
extern "C" int c3l_internal_narrow_scale(const double* x, float* y, double a, int64_t dim, void* stream);
extern "C" int c3l_internal_rmultiaxpy_f32(const float* V, int64_t ld, int32_t ncols, const double* alpha,
                                           double scale, double* w, int64_t dim, void* stream);
// Twin plans (equal digests) loaded / applied together, one launch per kernel.
extern "C" int c3l_internal_plans_are_twins(c3l_plan* const* plans, int32_t count, int32_t kind, int32_t program);
extern "C" int c3l_internal_plan_load_prepared_many(c3l_plan* const* plans, int32_t count, int32_t kind, void* stream);
extern "C" int c3l_internal_plan_scatter_map(c3l_plan* plan, int32_t kind, int64_t dim, const int64_t** map);
extern "C" double* c3l_internal_plan_workspace(c3l_plan* plan);
extern "C" int c3l_internal_plan_load_many(c3l_plan* const* plans, int32_t count, int32_t kind,
                                           const double* const* v, void* stream);
extern "C" int c3l_internal_plan_apply_many(c3l_plan* const* plans, int32_t count, int32_t program,
                                            const double* diag, const double* const* x, double* const* out,
                                            void* stream);

namespace {

int dfail(int code, const std::string& msg) { return c3l_internal_fail(code, msg.c_str()); }

// Optional HIP-event timing of the two bandwidth passes of every orthogonalisation
// step inside the drivers (families: 0 Arnoldi reduction, 1 Arnoldi update,
// 2 GMRES reduction, 3 GMRES update), with the algorithmic bytes of every timed
// launch.  Shared by all host threads.
struct KrylovTimer {
  std::mutex mu;
  bool enabled = false;
  struct Pending {
    int which;
    double bytes;
    hipEvent_t start, stop;
  };
  std::vector<Pending> pending;
  double ms[4] = {0, 0, 0, 0};
  double bytes[4] = {0, 0, 0, 0};
  int64_t launches[4] = {0, 0, 0, 0};
};

KrylovTimer& ktimer() {
  static KrylovTimer t;
  return t;
}

// Brackets one pass of an orthogonalisation step with events on its stream:
// family `base` + pass (0 reduction, 1 update).
struct DriverClock : c3l_ortho::Clock {
  int base;
  hipEvent_t start = nullptr;
  explicit DriverClock(int b) : base(b) {}
  void begin(int, hipStream_t stream) override {
    start = nullptr;
    if (!ktimer().enabled) return;
    if (hipEventCreate(&start) != hipSuccess) {
      start = nullptr;
      return;
    }
    if (hipEventRecord(start, stream) != hipSuccess) {
      (void)hipEventDestroy(start);
      start = nullptr;
    }
  }
  void end(int pass, double bytes, hipStream_t stream) override {
    if (start == nullptr) return;
    hipEvent_t stop = nullptr;
    if (hipEventCreate(&stop) != hipSuccess || hipEventRecord(stop, stream) != hipSuccess) {
      if (stop != nullptr) (void)hipEventDestroy(stop);
      (void)hipEventDestroy(start);
      start = nullptr;
      return;
    }
    std::lock_guard<std::mutex> guard(ktimer().mu);
    ktimer().pending.push_back({base + pass, bytes, start, stop});
    start = nullptr;
  }
};

#define D_CHECK(expr)                                                                       \
  do {                                                                                      \
    hipError_t e__ = (expr);                                                                \
    if (e__ != hipSuccess) return dfail(C3L_EHIP, std::string(#expr) + ": " + hipGetErrorString(e__)); \
  } while (0)

#define D_CALL(expr)            \
  do {                          \
    int rc__ = (expr);          \
    if (rc__ != C3L_OK) return rc__; \
  } while (0)

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// ---------------------------------------------------------------------------
// Multi-shift GMRES on one shared Arnoldi basis.
//
// For shift z the (j+2)-entry column j of (z I~ - H) is a_k = [k == j] z - H[k, j].
// The earlier Givens rotations act on it as the recurrence
//   x_{k+1} = -conj(s_k) x_k + c_k a_{k+1},   R[k, j] = c_k x_k + s_k a_{k+1}.
// The solution is tracked only through its sink projection:
//   pi_j = (sink.v_j - sum_{k<j} R[k, j] pi_k) / R[j, j],  G += gamma_j pi_j,
// where gamma_j is the rotated right-hand side entry (GMRES with incremental
// solution update, x_m = x_{m-1} + gamma_m p_m, p = V R^{-1}).  The GMRES
// residual |gamma~| / beta is exact for every shift at every step.
// Per-shift state (row s, length m_max): rot_c, rot_s, pi.
__device__ __forceinline__ double2 shfl2(double2 v, int lane) {
  return make_double2(__shfl(v.x, lane, 64), __shfl(v.y, lane, 64));
}
__device__ __forceinline__ double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }

// Multi-shift GMRES over a batch of nb <= kShiftBatch completed columns c0..c0+nb-1:
// kShiftBatch lanes per shift, lane b of a group owning column c0+b.  Each lane applies
// the shift's rotations 0..c0-1 (computed by earlier launches) to its column in order,
// accumulating sum_k R[k, c0+b] pi_k on the way; the group's first lane then applies the
// rotations of the batch itself, column after column, and forms each new rotation, pi
// and Green increment.  The residual and Green value after every column go to
// resid_hist / green_hist ([b][shift]) so that the stop test can take the state of the
// exact stop column.
constexpr int kShiftBatch = 8;
constexpr int kMaxGroupBases = 64;  // bases of one group run (c3l_arnoldi_group_begin)
constexpr int kShiftsPerBlock = 256 / kShiftBatch;
constexpr int kTileK = 32;  // rotations staged in LDS per tile

__device__ __forceinline__ void multishift_batch_body(const double2* __restrict__ H, int c0, int nb, int m_max,
                                                      const double2* __restrict__ shifts, int n_shifts,
                                                      const double2* __restrict__ sink_v, double beta,
                                                      double* __restrict__ rot_c, double2* __restrict__ rot_s,
                                                      double2* __restrict__ pi, double2* __restrict__ gamma,
                                                      double2* __restrict__ green, double* __restrict__ resid,
                                                      double* __restrict__ resid_hist,
                                                      double2* __restrict__ green_hist,
                                                      const double* __restrict__ status, int galerkin) {
  if (status[2] != 0.0) return;  // uniform: nothing after the stop
  const int lane = threadIdx.x & 63;
  const int mine = lane % kShiftBatch;
  const int s_raw = blockIdx.x * kShiftsPerBlock + (int)(threadIdx.x / kShiftBatch);
  const bool valid = s_raw < n_shifts;
  const int s = valid ? s_raw : n_shifts - 1;  // groups past the end repeat the last shift and write nothing
  const double2 z = shifts[s];
  const int64_t row = (int64_t)s * m_max;
  const int64_t ldh = m_max + 1;
  const int cm = c0 + (mine < nb ? mine : 0);  // lanes past nb repeat column c0 (unused)
  const double2* __restrict__ hc = H + (int64_t)cm * ldh;
  double2 xm = make_double2(-hc[0].x + (cm == 0 ? z.x : 0.0), -hc[0].y + (cm == 0 ? z.y : 0.0));
  double2 am = make_double2(0.0, 0.0);
  // The rotations of the block's shifts and the lanes' H entries pass through LDS in tiles
  // of kTileK steps: one round of coalesced loads serves kTileK links of the chain (rows
  // padded by one entry so that the eight shifts of a wave read different banks).
  static_assert(kShiftBatch * kTileK == 256, "one H entry per thread and tile");
  __shared__ double s_c[kShiftsPerBlock][kTileK + 1];
  __shared__ double2 s_s[kShiftsPerBlock][kTileK + 1];
  __shared__ double2 s_p[kShiftsPerBlock][kTileK + 1];
  __shared__ double2 s_h[kShiftBatch][kTileK + 1];
  const int local = (int)(threadIdx.x / kShiftBatch);
  for (int t0 = 0; t0 < c0; t0 += kTileK) {
    const int len = min(kTileK, c0 - t0);
    __syncthreads();  // the previous tile is consumed
    for (int e = threadIdx.x; e < kShiftsPerBlock * kTileK; e += 256) {
      const int sh = e / kTileK, kk = e % kTileK;
      if (kk < len) {
        const int se = min((int)blockIdx.x * kShiftsPerBlock + sh, n_shifts - 1);
        const int64_t at = (int64_t)se * m_max + t0 + kk;
        s_c[sh][kk] = rot_c[at];
        s_s[sh][kk] = rot_s[at];
        s_p[sh][kk] = pi[at];
      }
    }
    {
      const int b = (int)(threadIdx.x / kTileK), kk = (int)(threadIdx.x % kTileK);
      const int cb = c0 + (b < nb ? b : 0);
      if (kk < len) s_h[b][kk] = H[(int64_t)cb * ldh + t0 + kk + 1];
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < len; ++kk) {
      const int k = t0 + kk;
      const double c = s_c[local][kk];
      const double2 sn = s_s[local][kk];
      const double2 pk = s_p[local][kk];
      const double2 h = s_h[mine][kk];
      const double2 a1 = make_double2(-h.x + (k + 1 == cm ? z.x : 0.0), -h.y + (k + 1 == cm ? z.y : 0.0));
      const double2 R = zadd(make_double2(c * xm.x, c * xm.y), zmul(sn, a1));
      am = zadd(am, zmul(R, pk));
      xm = zadd(zmul(make_double2(-sn.x, sn.y), xm), make_double2(c * a1.x, c * a1.y));
    }
  }
  // the batch's own Hessenberg entries H[c0+r+1, c0+b] (r <= b) and sink projections,
  // loaded once by the block instead of one dependent load after another by the tail
  __shared__ double2 s_hb[kShiftBatch][kShiftBatch];
  __shared__ double2 s_sv[kShiftBatch];
  if (threadIdx.x < kShiftBatch * kShiftBatch) {
    const int b = (int)(threadIdx.x / kShiftBatch), r = (int)(threadIdx.x % kShiftBatch);
    if (b < nb && r <= b) s_hb[b][r] = H[(int64_t)(c0 + b) * ldh + c0 + r + 1];
  } else if (threadIdx.x < kShiftBatch * kShiftBatch + kShiftBatch) {
    const int b = (int)threadIdx.x - kShiftBatch * kShiftBatch;
    if (b < nb) s_sv[b] = sink_v[c0 + b];
  }
  __syncthreads();
  // the group's first lane collects every column's rotated entry and partial sum
  const int first = lane - mine;
  double2 x[kShiftBatch], acc[kShiftBatch];
#pragma unroll
  for (int b = 0; b < kShiftBatch; ++b) {
    x[b] = shfl2(xm, first + b);
    acc[b] = shfl2(am, first + b);
  }
  if (mine != 0 || !valid) return;
  double cr[kShiftBatch];
  double2 sr[kShiftBatch], pr[kShiftBatch];
  double2 g = gamma[s], G = green[s];
  double res = 0.0;
#pragma unroll
  for (int b = 0; b < kShiftBatch; ++b) {
    if (b >= nb) break;
    const int cb = c0 + b;
    double2 xb = x[b], ab = acc[b];
    for (int bb = 0; bb < b; ++bb) {  // rotations of this batch
      const int k = c0 + bb;
      const double2 hk = s_hb[b][bb];  // H[k+1, cb]
      const double2 a1 = make_double2(-hk.x + (k + 1 == cb ? z.x : 0.0), -hk.y + (k + 1 == cb ? z.y : 0.0));
      const double2 R = zadd(make_double2(cr[bb] * xb.x, cr[bb] * xb.y), zmul(sr[bb], a1));
      ab = zadd(ab, zmul(R, pr[bb]));
      xb = zadd(zmul(make_double2(-sr[bb].x, sr[bb].y), xb), make_double2(cr[bb] * a1.x, cr[bb] * a1.y));
    }
    const double2 bsub = make_double2(-s_hb[b][b].x, -s_hb[b][b].y);
    const double ax = hypot(xb.x, xb.y);
    const double bx = hypot(bsub.x, bsub.y);
    double cn;
    double2 snn, r;
    if (ax == 0.0) {
      cn = 0.0;
      snn = make_double2(1.0, 0.0);
      r = bsub;
    } else {
      const double nrm = hypot(ax, bx);
      const double2 u = make_double2(xb.x / ax, xb.y / ax);
      cn = ax / nrm;
      snn = zmul(u, make_double2(bsub.x / nrm, -bsub.y / nrm));
      r = make_double2(u.x * nrm, u.y * nrm);
    }
    rot_c[row + cb] = cn;
    rot_s[row + cb] = snn;
    cr[b] = cn;
    sr[b] = snn;
    // pi_cb = (sink.v_cb - sum_k R[k, cb] pi_k) / R[cb, cb]
    const double2 num = make_double2(s_sv[b].x - ab.x, s_sv[b].y - ab.y);
    const double den = r.x * r.x + r.y * r.y;
    const double2 p = den > 0.0 ? make_double2((num.x * r.x + num.y * r.y) / den, (num.y * r.x - num.x * r.y) / den)
                                : make_double2(0.0, 0.0);
    pi[row + cb] = p;
    pr[b] = p;
    const double2 gj = make_double2(cn * g.x, cn * g.y);
    // Galerkin (FOM) residual |h_{c+1,c}| |y_c| / beta of the square system: the last
    // coordinate is g / x_b before this rotation, i.e. the GMRES residual divided by cn.
    const double fom = ax > 0.0 ? bx * hypot(g.x, g.y) / (ax * beta) : INFINITY;
    g = zmul(make_double2(-snn.x, snn.y), g);  // residual entry after this rotation
    G = zadd(G, zmul(gj, p));
    res = galerkin ? fom : hypot(g.x, g.y) / beta;
    resid_hist[(int64_t)b * n_shifts + s] = res;
    green_hist[(int64_t)b * n_shifts + s] = G;
  }
  gamma[s] = g;
  green[s] = G;
  resid[s] = res;
}


__global__ void init_shift_state_kernel(double2* __restrict__ gamma, double2* __restrict__ green, double beta,
                                        int n_shifts, double* __restrict__ status) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n_shifts; s += gridDim.x * blockDim.x) {
    gamma[s] = make_double2(beta, 0.0);
    green[s] = make_double2(0.0, 0.0);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    status[0] = 0.0;
    status[1] = INFINITY;
    status[2] = 0.0;
    status[3] = 0.0;
  }
}

__global__ void fill_kernel(double* __restrict__ a, double v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) a[i] = v;
}

// Stop test after a batch (one workgroup).  For the first column c of the batch with a
// nonfinite or broken-down sub-diagonal H[c+1, c], or (c+1 >= min_dim) with every shift's
// residual <= tolerance: status = [hn, worst, flag (3 nonfinite, 2 breakdown, 1 converged),
// c+1], and green/resid take the values after that column.  Without a stop, status[1] is
// the worst residual after the last column and status[3] its count.
__device__ __forceinline__ void arnoldi_batch_status_body(const double* __restrict__ resid_hist,
                                                          const double2* __restrict__ green_hist,
                                                          double2* __restrict__ green, double* __restrict__ resid,
                                                          int n, double* __restrict__ status,
                                                          const double2* __restrict__ H, int m_max, int c0, int nb,
                                                          int min_dim, double tolerance, double breakdown) {
  __shared__ double part[kShiftBatch][16];
  __shared__ double sub[kShiftBatch];
  __shared__ int stop_b;
  if (status[2] != 0.0) return;  // uniform: frozen after an earlier stop
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t ldh = m_max + 1;
  if ((int)threadIdx.x < nb) sub[threadIdx.x] = H[(int64_t)(c0 + threadIdx.x) * ldh + c0 + threadIdx.x + 1].x;
  double m[kShiftBatch];
#pragma unroll
  for (int b = 0; b < kShiftBatch; ++b) m[b] = 0.0;
  // Four rows of every column per thread and pass, all loads unconditional: indices past
  // the end repeat the last shift (a duplicate cannot change a maximum) and columns past
  // nb repeat column nb-1 (their maxima are never read).
  constexpr int kRows = 4;
  for (int i0 = threadIdx.x; i0 < n; i0 += kRows * (int)blockDim.x) {
    double v[kRows][kShiftBatch];
#pragma unroll
    for (int r = 0; r < kRows; ++r)
#pragma unroll
      for (int b = 0; b < kShiftBatch; ++b)
        v[r][b] = resid_hist[(int64_t)min(b, nb - 1) * n + min(i0 + r * (int)blockDim.x, n - 1)];
#pragma unroll
    for (int r = 0; r < kRows; ++r)
#pragma unroll
      for (int b = 0; b < kShiftBatch; ++b) {
        const double x = v[r][b];
        m[b] = (x != x) ? INFINITY : (x > m[b] ? x : m[b]);  // NaN counts as unconverged
      }
  }
#pragma unroll
  for (int b = 0; b < kShiftBatch; ++b) {
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_down(m[b], off, 64);
      m[b] = o > m[b] ? o : m[b];
    }
    if (lane == 0) part[b][wave] = m[b];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    stop_b = -1;
    const int waves = (int)(blockDim.x >> 6);
    for (int b = 0; b < nb; ++b) {
      double worst = 0.0;
      for (int u = 0; u < waves; ++u) worst = part[b][u] > worst ? part[b][u] : worst;
      const int c = c0 + b;
      const double hn = sub[b];
      status[0] = hn;
      status[1] = worst;
      status[3] = (double)(c + 1);
      double flag = 0.0;
      if (!(hn == hn) || hn == INFINITY) flag = 3.0;
      else if (hn <= breakdown) flag = 2.0;
      else if (c + 1 >= min_dim && worst <= tolerance) flag = 1.0;
      if (flag != 0.0) {
        status[2] = flag;
        stop_b = b;
        break;
      }
    }
  }
  __syncthreads();
  const int sb = stop_b;
  if (sb < 0 || sb == nb - 1) return;  // the shift state already belongs to the last column
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    green[i] = green_hist[(int64_t)sb * n + i];
    resid[i] = resid_hist[(int64_t)sb * n + i];
  }
}

// Device state of one Arnoldi basis's shifts; the batch kernels run over several
// bases of equal shape at once (blockIdx.z = basis).
struct ShiftState {
  const double2* H;
  const double2* sink_v;
  double beta;
  double* rot_c;
  double2* rot_s;
  double2* pi;
  double2* gamma;
  double2* green;
  double* resid;
  double* resid_hist;
  double2* green_hist;
  double* status;
  int galerkin;  // 1: stop on the Galerkin (FOM) residual, 0: on the GMRES residual
};
using ShiftStates = c3l_ortho::Many<ShiftState>;

__global__ __launch_bounds__(256) void multishift_batch_kernel(ShiftStates states, int c0, int nb, int m_max,
                                                               const double2* __restrict__ shifts, int n_shifts) {
  const ShiftState& t = states.p[blockIdx.z];
  multishift_batch_body(t.H, c0, nb, m_max, shifts, n_shifts, t.sink_v, t.beta, t.rot_c, t.rot_s, t.pi, t.gamma,
                        t.green, t.resid, t.resid_hist, t.green_hist, t.status, t.galerkin);
}

__global__ __launch_bounds__(1024) void arnoldi_batch_status_kernel(ShiftStates states, int n, int m_max, int c0,
                                                                    int nb, int min_dim, double tolerance,
                                                                    double breakdown) {
  const ShiftState& t = states.p[blockIdx.z];
  arnoldi_batch_status_body(t.resid_hist, t.green_hist, t.green, t.resid, n, t.status, t.H, m_max, c0, nb, min_dim,
                            tolerance, breakdown);
}

bool out_invalid(const int32_t* m_out, const int32_t* breakdown_out, const double* worst_out) {
  return m_out == nullptr || breakdown_out == nullptr || worst_out == nullptr;
}

int64_t blocks_for(int64_t n, int t = 256) {
  int64_t b = (n + t - 1) / t;
  return b < 1 ? 1 : (b > 65536 ? 65536 : b);
}

// Workspace layouts in doubles.
// GMRES (restart R, real basis of rd rows):
//   coef:    red [2(R+2)] | sc [16] | H [R(R+1)] | y [R] | stage [R+2]
//   scratch: partial   (also the norms' partials)
// Arnoldi (m_max columns, complex basis of dim rows, n_shifts shifts), complex entries doubled:
//   coef:    red [2(m+2)] | sc [16] | sink_v [m+1]
//   scratch: w [dim] | partial | green_hist [kShiftBatch n_shifts] | resid_hist (doubles)
struct Layout {
  int64_t red, sc, H, y, stage, sink_v, coef_total;
  int64_t w, partial, green_hist, resid_hist, scratch_total;
};

Layout gmres_layout(int64_t rd, int32_t R) {
  Layout L{};
  L.red = 0;
  L.sc = L.red + 2 * ((int64_t)R + 2);
  L.H = L.sc + 16;
  L.y = L.H + (int64_t)R * (R + 1);
  L.stage = L.y + R;
  L.coef_total = L.stage + R + 2;
  L.w = 0;
  L.partial = 0;
  L.scratch_total = L.partial + c3l_ortho::partial_entries(false, rd, R);
  return L;
}

Layout arnoldi_layout(int64_t dim, int32_t m, int32_t n_shifts) {
  Layout L{};
  L.red = 0;
  L.sc = L.red + 2 * 2 * ((int64_t)m + 2);
  L.sink_v = L.sc + 2 * 16;
  L.coef_total = L.sink_v + 2 * ((int64_t)m + 1);
  L.w = 0;
  L.partial = 2 * dim;
  L.green_hist = L.partial + 2 * c3l_ortho::partial_entries(true, dim, m);
  L.resid_hist = L.green_hist + 2 * (int64_t)kShiftBatch * n_shifts;
  L.scratch_total = L.resid_hist + (int64_t)kShiftBatch * n_shifts;
  return L;
}

}  // namespace

namespace {

struct Gmres {
  c3l_plan* plan;
  const c3l_tilde_maps* tm;
  const double* jvp_diag;
  double* t_work;
  double* out_work;
  double* scratch;
  double* coef;
  void* stream;
  int64_t rd;
  int32_t* jvp_calls;
  int64_t ldv = 0;  // basis leading dimension (entries)

  int apply(const double* v, double* out) {
    D_CALL(c3l_tilde_unpack(v, t_work, tm->k0, tm->k1, tm->ab, tm->dim, stream));
    D_CALL(c3l_plan_load(plan, 1, t_work, stream));
    D_CALL(c3l_plan_apply(plan, 2, jvp_diag, t_work, out_work, stream));
    D_CALL(c3l_tilde_pack(out_work, out, tm->src, tm->w, rd, stream));
    if (jvp_calls) ++*jvp_calls;
    return C3L_OK;
  }
  int unpack(const double* v) { return c3l_tilde_unpack(v, t_work, tm->k0, tm->k1, tm->ab, tm->dim, stream); }
  int unpack(const float* v) { return c3l_internal_tilde_unpack_f32(v, t_work, tm->k0, tm->k1, tm->ab, tm->dim, stream); }
  // basis column = a x
  int store_scaled(const double* x, double* col, double a) { return c3l_rscale(x, col, a, rd, stream); }
  int store_scaled(const double* x, float* col, double a) { return c3l_internal_narrow_scale(x, col, a, rd, stream); }
  // x += V y
  int add_combination(const double* V, int used, const double* y, double* x) {
    return c3l_rmultiaxpy(V, ldv > 0 ? ldv : rd, used, y, 1.0, x, rd, stream);
  }
  int add_combination(const float* V, int used, const double* y, double* x) {
    return c3l_internal_rmultiaxpy_f32(V, ldv > 0 ? ldv : rd, used, y, 1.0, x, rd, stream);
  }
  // out = M^{-1} J v with the preconditioner fused into the pack of J v
  template <typename S>
  int apply_preconditioned(const S* v, double* out) {
    D_CALL(unpack(v));
    D_CALL(c3l_plan_load(plan, 1, t_work, stream));
    D_CALL(c3l_plan_apply(plan, 2, jvp_diag, t_work, out_work, stream));
    D_CALL(c3l_tilde_pack_precondition(out_work, out, tm->src, tm->w, tm->inv_diag, rd, stream));
    if (jvp_calls) ++*jvp_calls;
    return C3L_OK;
  }
  int precondition(const double* v, double* out) {
    return c3l_tilde_precondition(v, out, tm->k0, tm->k1, tm->ab, tm->src, tm->w, tm->inv_diag, rd, stream);
  }
  int norm(const double* v, double* result) {
    D_CALL(c3l_rmultidot(v, rd, 1, v, rd, scratch, coef, stream));
    double h = 0.0;
    D_CHECK(hipMemcpyAsync(&h, coef, sizeof(double), hipMemcpyDeviceToHost, static_cast<hipStream_t>(stream)));
    D_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    *result = std::sqrt(h > 0.0 ? h : 0.0);
    return C3L_OK;
  }
};

// Two pinned host slots and their events: the host reads Hessenberg column c from slot
// c % 2 while the GPU already works on the next step.  The slots are mapped, coherent host
// memory: the update pass of step c + 1 writes the column straight into its slot (no copy
// launch), and the event recorded after that pass tells the host the slot is complete.
struct HostColumns {
  double* host = nullptr;
  double* dev = nullptr;  // the same slots as the device addresses them
  hipEvent_t done[2] = {nullptr, nullptr};
  int64_t stride = 0;
  int allocate(size_t doubles) {
    if (hipHostMalloc(reinterpret_cast<void**>(&host), sizeof(double) * doubles,
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
      return dfail(C3L_EHIP, "hipHostMalloc of the Hessenberg column slots failed");
    if (hipHostGetDevicePointer(reinterpret_cast<void**>(&dev), host, 0) != hipSuccess)
      return dfail(C3L_EHIP, "hipHostGetDevicePointer of the Hessenberg column slots failed");
    return C3L_OK;
  }
  ~HostColumns() {
    for (hipEvent_t e : done)
      if (e != nullptr) {
        (void)hipEventSynchronize(e);  // no update pass may still write the slots when they are freed
        (void)hipEventDestroy(e);
      }
    if (host != nullptr) (void)hipHostFree(host);
  }
};

template <typename Store>
int gmres_run(c3l_plan* plan, const c3l_tilde_maps* tm, const double* jvp_diag, double* t_work, double* out_work,
              const double* b, double* x, Store* V, double* w, double* tmp, double* scratch, double* coef,
              double rtol, double atol_in, int32_t restart, int32_t cycles, int32_t* info, int32_t* iterations,
              double* true_rel, int32_t* jvp_calls, void* stream) {
  if (plan == nullptr || tm == nullptr || b == nullptr || x == nullptr || V == nullptr || w == nullptr ||
      tmp == nullptr || scratch == nullptr || coef == nullptr || restart <= 0 || cycles <= 0 || tm->real_dim <= 0 ||
      !(rtol > 0.0) || !(atol_in >= 0.0) || info == nullptr || iterations == nullptr || true_rel == nullptr)
    return dfail(C3L_EINVAL, "c3l_gmres: invalid argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t rd = tm->real_dim;
  const int64_t ldv = c3l_krylov_basis_ld(rd);
  Gmres g{plan, tm, jvp_diag, t_work, out_work, scratch, coef, stream, rd, jvp_calls, ldv};
  hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(rd)), dim3(256), 0, s, x, 0.0, rd);
  D_CHECK(hipGetLastError());
  double bnrm = 0.0;
  D_CALL(g.norm(b, &bnrm));
  *iterations = 0;
  *info = 1;
  *true_rel = 0.0;
  if (bnrm == 0.0) {
    *info = 0;
    return C3L_OK;
  }
  D_CALL(g.precondition(b, tmp));
  double mbnrm = 0.0;
  D_CALL(g.norm(tmp, &mbnrm));
  // SciPy's tolerance control: the outer target is max(atol, rtol ||b||) on the true
  // residual; the inner test acts on the preconditioned residual with a tolerance
  // rescaled after every restart (tightened when the inner test passed but the true
  // residual did not, relaxed when the cycle ran out).
  const double eps = 2.220446049250313e-16;
  const double atol = std::max(atol_in, rtol * bnrm);
  double ptol_factor = 1.0;
  double ptol = mbnrm * std::min(ptol_factor, atol / bnrm);
  if (bnrm < atol) {  // x = 0 already meets an absolute target
    D_CHECK(hipMemcpyAsync(w, b, sizeof(double) * rd, hipMemcpyDeviceToDevice, s));
    *info = 0;
    *true_rel = 1.0;
    return C3L_OK;
  }
  const int R = restart;
  const Layout L = gmres_layout(rd, R);
  c3l_ortho::Basis<double, Store> basis{V, ldv, rd, coef + L.H, R + 1, nullptr, nullptr, scratch + L.partial,
                                 coef + L.red, coef + L.sc, nullptr, coef + L.stage};
  DriverClock clock(2);
  HostColumns pinned;
  pinned.stride = R + 2;
  D_CALL(pinned.allocate((size_t)(2 * pinned.stride)));
  for (hipEvent_t& e : pinned.done) D_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  std::vector<double> H((size_t)(R + 1) * R), cs(R), sn(R), S(R + 1), y(R);
  auto h = [&](int i, int j) -> double& { return H[(size_t)i * R + j]; };
  D_CHECK(hipMemcpyAsync(w, b, sizeof(double) * rd, hipMemcpyDeviceToDevice, s));  // r = b - A 0
  double rn = bnrm;
  for (int32_t cycle = 0; cycle < cycles; ++cycle) {
    D_CALL(g.precondition(w, tmp));
    double beta = 0.0;
    D_CALL(g.norm(tmp, &beta));
    if (!(beta > 0.0) || !std::isfinite(beta)) return dfail(C3L_EINVAL, "c3l_gmres: singular preconditioned residual");
    D_CALL(g.store_scaled(tmp, V, 1.0 / beta));
    std::fill(H.begin(), H.end(), 0.0);
    std::fill(S.begin(), S.end(), 0.0);
    S[0] = beta;
    int used = 0;
    bool breakdown = false, stop = false;
    double presid = 0.0;
    // Givens update of completed column c from its pinned slot; true when the cycle stops.
    auto absorb = [&](int c) -> int {
      const double* col = pinned.host + (int64_t)(c % 2) * pinned.stride;
      D_CHECK(hipEventSynchronize(pinned.done[c % 2]));
      for (int k = 0; k <= c; ++k) h(k, c) = col[k];
      const double hn = col[c + 1];
      const double h0 = col[c + 2];  // |A v_c| before projection
      if (!std::isfinite(hn)) return dfail(C3L_EINVAL, "c3l_gmres: nonfinite Krylov vector");
      if (hn <= eps * h0) {  // exact solution indicator
        h(c + 1, c) = 0.0;
        breakdown = true;
      } else {
        h(c + 1, c) = hn;
      }
      for (int k = 0; k < c; ++k) {
        const double a = h(k, c), bb = h(k + 1, c);
        h(k, c) = cs[k] * a + sn[k] * bb;
        h(k + 1, c) = -sn[k] * a + cs[k] * bb;
      }
      const double a = h(c, c), bb = h(c + 1, c);
      const double r = std::hypot(a, bb);
      cs[c] = r == 0.0 ? 1.0 : a / r;
      sn[c] = r == 0.0 ? 0.0 : bb / r;
      h(c, c) = r;
      h(c + 1, c) = 0.0;
      S[c + 1] = -sn[c] * S[c];
      S[c] = cs[c] * S[c];
      presid = std::fabs(S[c + 1]);
      ++*iterations;
      used = c + 1;
      stop = presid <= ptol || breakdown;
      return C3L_OK;
    };
    // Step j applies the operator to the once-projected v_j (none at j = R) and completes
    // column j-1, which its update pass writes into pinned slot (j-1) % 2; the host absorbs
    // column j-2 meanwhile, so it never waits for the step it has just queued (and has read
    // slot (j-2) % 2 before step j+1, which rewrites it, is queued).
    int absorbed = 0;
    for (int j = 0; j <= R && !stop; ++j) {
      const bool has_w = j < R;
      if (has_w) D_CALL(g.apply_preconditioned(V + (int64_t)j * ldv, w));
      const int c = j - 1;
      basis.stage = j == 0 ? nullptr : pinned.dev + (int64_t)(c % 2) * pinned.stride;
      D_CALL(c3l_ortho::step(basis, j, j == 0, has_w ? w : nullptr, 0.0, s, &clock));
      if (j == 0) continue;
      D_CHECK(hipEventRecord(pinned.done[c % 2], s));
      if (c >= 1) {
        D_CALL(absorb(c - 1));
        absorbed = c;
      }
    }
    while (!stop && absorbed < R) {  // columns still in flight at the end of the cycle
      D_CALL(absorb(absorbed));
      ++absorbed;
    }
    if (h(used - 1, used - 1) == 0.0) S[used - 1] = 0.0;
    for (int k = used - 1; k >= 0; --k) {
      double acc = S[k];
      for (int l = k + 1; l < used; ++l) acc -= h(k, l) * y[l];
      y[k] = h(k, k) == 0.0 ? 0.0 : acc / h(k, k);
    }
    D_CHECK(hipMemcpyAsync(coef + L.y, y.data(), sizeof(double) * used, hipMemcpyHostToDevice, s));
    D_CALL(g.add_combination(V, used, coef + L.y, x));
    D_CALL(g.apply(x, w));
    D_CALL(c3l_raxpby(1.0, b, -1.0, w, rd, stream));
    D_CALL(g.norm(w, &rn));  // synchronises: y may be reused by the next cycle
    if (rn <= atol || breakdown) break;
    ptol_factor = presid <= ptol ? std::max(eps, 0.25 * ptol_factor) : std::min(1.0, 1.5 * ptol_factor);
    ptol = presid * std::min(ptol_factor, atol / rn);
  }
  *true_rel = rn / bnrm;
  *info = rn <= atol ? 0 : cycles;
  return C3L_OK;
}

// ---- GMRES for a group of independent systems in lockstep (c3l_gmres_group) --------------
//
// `count` systems of equal real dimension rd share one union plan (their contraction plans
// side by side, one launch per program for all of them) and one set of union tilde maps
// (dimension count * rd).  Every vector is the systems' vectors side by side; the Krylov basis
// column j holds column j of every system (system i's basis is V + i*rd with leading
// dimension count*rd).  Each system runs its own restarted GMRES (gmres_run's algorithm:
// SciPy's tolerance control, delayed-reorthogonalisation CGS, host Givens one step behind the
// device) in lockstep with the others: one union Jacobian product per step serves them all,
// the orthogonalisation runs over the systems still iterating (up to four per launch), a
// system that meets its inner test leaves the cycle, and one that meets its target leaves
// the solve.  Cycles end together: the union product of the updated steps gives every
// system's true residual at once.
struct GroupSystem {
  c3l_ortho::Basis<double, double> b64;
  c3l_ortho::Basis<double, float> b32;
  double* scratch = nullptr;
  double* coef = nullptr;
  bool done = false, stop = false, breakdown = false, trivial = false;
  double bnrm = 0.0, atol = 0.0, ptol_factor = 1.0, ptol = 0.0, presid = 0.0, rn = 0.0;
  int used = 0, absorbed = 0;
  std::vector<double> H, cs, sn, S, y;
};

template <typename Store>
int gmres_group_run(c3l_plan* plan, const c3l_tilde_maps* tm, int32_t count, const double* jvp_diag, double* t_work,
                    double* out_work, const double* b, double* x, Store* V, double* w, double* tmp, double* scratch,
                    double* coef, const double* rtol, const double* atol_in, const int32_t* active, int32_t restart,
                    int32_t cycles, int32_t* info, int32_t* iterations, double* true_rel, int32_t* jvp_calls,
                    void* stream) {
  if (plan == nullptr || tm == nullptr || b == nullptr || x == nullptr || V == nullptr || w == nullptr ||
      tmp == nullptr || scratch == nullptr || coef == nullptr || rtol == nullptr || atol_in == nullptr ||
      active == nullptr || info == nullptr || iterations == nullptr || true_rel == nullptr || jvp_calls == nullptr ||
      count < 1 || restart <= 0 || cycles <= 0 || tm->real_dim <= 0 || tm->real_dim % count != 0)
    return dfail(C3L_EINVAL, "c3l_gmres_group: invalid argument");
  for (int32_t i = 0; i < count; ++i)
    if (active[i] && (!(rtol[i] > 0.0) || !(atol_in[i] >= 0.0)))
      return dfail(C3L_EINVAL, "c3l_gmres_group: invalid tolerance");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t rd = tm->real_dim / count;
  const int64_t ldv = tm->real_dim;
  const int R = restart;
  const Layout L = gmres_layout(rd, R);
  const double eps = 2.220446049250313e-16;
  Gmres g{plan, tm, jvp_diag, t_work, out_work, scratch, coef, stream, tm->real_dim, nullptr};
  std::vector<GroupSystem> sys((size_t)count);
  for (int32_t i = 0; i < count; ++i) {
    GroupSystem& q = sys[(size_t)i];
    q.scratch = scratch + i * L.scratch_total;
    q.coef = coef + i * L.coef_total;
    q.b64 = c3l_ortho::Basis<double, double>{reinterpret_cast<double*>(V) + i * rd, ldv, rd, q.coef + L.H, R + 1,
                                              nullptr, nullptr, q.scratch + L.partial, q.coef + L.red,
                                              q.coef + L.sc, nullptr, q.coef + L.stage};
    q.b32 = c3l_ortho::Basis<double, float>{reinterpret_cast<float*>(V) + i * rd, ldv, rd, q.coef + L.H, R + 1,
                                             nullptr, nullptr, q.scratch + L.partial, q.coef + L.red,
                                             q.coef + L.sc, nullptr, q.coef + L.stage};
    q.H.assign((size_t)(R + 1) * R, 0.0);
    q.cs.assign(R, 0.0);
    q.sn.assign(R, 0.0);
    q.S.assign(R + 1, 0.0);
    q.y.assign(R, 0.0);
    q.done = !active[i];
    iterations[i] = 0;
    info[i] = active[i] ? 1 : 0;
    true_rel[i] = 0.0;
    jvp_calls[i] = 0;
  }
  // per-system norms of a union vector: one reduction per live system, one copy, one wait
  std::vector<double> norms((size_t)count, 0.0);
  auto segment_norms = [&](const double* v, std::vector<int32_t> const& which) -> int {
    for (int32_t i : which)
      D_CALL(c3l_rmultidot(v + i * rd, rd, 1, v + i * rd, rd, sys[(size_t)i].scratch,
                           sys[(size_t)i].coef + L.y, stream));
    for (int32_t i : which)
      D_CHECK(hipMemcpyAsync(&norms[(size_t)i], sys[(size_t)i].coef + L.y, sizeof(double), hipMemcpyDeviceToHost, s));
    D_CHECK(hipStreamSynchronize(s));
    for (int32_t i : which) norms[(size_t)i] = std::sqrt(norms[(size_t)i] > 0.0 ? norms[(size_t)i] : 0.0);
    return C3L_OK;
  };
  auto live_list = [&]() {
    std::vector<int32_t> out;
    for (int32_t i = 0; i < count; ++i)
      if (!sys[(size_t)i].done) out.push_back(i);
    return out;
  };
  hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(ldv)), dim3(256), 0, s, x, 0.0, ldv);
  D_CHECK(hipGetLastError());
  std::vector<int32_t> live = live_list();
  D_CALL(segment_norms(b, live));
  for (int32_t i : live) {
    sys[(size_t)i].bnrm = norms[(size_t)i];
    if (sys[(size_t)i].bnrm == 0.0) {
      sys[(size_t)i].done = true;
      info[i] = 0;
    }
  }
  live = live_list();
  D_CALL(g.precondition(b, tmp));
  D_CALL(segment_norms(tmp, live));
  D_CHECK(hipMemcpyAsync(w, b, sizeof(double) * ldv, hipMemcpyDeviceToDevice, s));  // r = b - A 0
  for (int32_t i : live) {
    GroupSystem& q = sys[(size_t)i];
    q.atol = std::max(atol_in[i], rtol[i] * q.bnrm);
    q.ptol = norms[(size_t)i] * std::min(q.ptol_factor, q.atol / q.bnrm);
    q.rn = q.bnrm;
    if (q.bnrm < q.atol) {  // x = 0 already meets an absolute target
      q.done = q.trivial = true;
      info[i] = 0;
      true_rel[i] = 1.0;
    }
  }
  // pinned Hessenberg columns: two slots per system
  HostColumns pinned;
  pinned.stride = R + 2;
  D_CALL(pinned.allocate((size_t)(2 * pinned.stride * count)));
  std::vector<hipEvent_t> events((size_t)2 * count, nullptr);
  struct EventGuard {
    std::vector<hipEvent_t>& e;
    ~EventGuard() {
      for (hipEvent_t ev : e)
        if (ev != nullptr) {
          (void)hipEventSynchronize(ev);
          (void)hipEventDestroy(ev);
        }
    }
  } guard{events};
  for (hipEvent_t& e : events) D_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  auto slot = [&](int32_t i, int c) { return pinned.host + ((int64_t)i * 2 + (c % 2)) * pinned.stride; };
  auto slot_dev = [&](int32_t i, int c) { return pinned.dev + ((int64_t)i * 2 + (c % 2)) * pinned.stride; };
  auto absorb = [&](int32_t i, int c) -> int {
    GroupSystem& q = sys[(size_t)i];
    const double* col = slot(i, c);
    D_CHECK(hipEventSynchronize(events[(size_t)i * 2 + (c % 2)]));
    auto h = [&](int r, int k) -> double& { return q.H[(size_t)r * R + k]; };
    for (int k = 0; k <= c; ++k) h(k, c) = col[k];
    const double hn = col[c + 1];
    const double h0 = col[c + 2];
    if (!std::isfinite(hn)) return dfail(C3L_EINVAL, "c3l_gmres_group: nonfinite Krylov vector");
    if (hn <= eps * h0) {
      h(c + 1, c) = 0.0;
      q.breakdown = true;
    } else {
      h(c + 1, c) = hn;
    }
    for (int k = 0; k < c; ++k) {
      const double a = h(k, c), bb = h(k + 1, c);
      h(k, c) = q.cs[k] * a + q.sn[k] * bb;
      h(k + 1, c) = -q.sn[k] * a + q.cs[k] * bb;
    }
    const double a = h(c, c), bb = h(c + 1, c);
    const double r = std::hypot(a, bb);
    q.cs[c] = r == 0.0 ? 1.0 : a / r;
    q.sn[c] = r == 0.0 ? 0.0 : bb / r;
    h(c, c) = r;
    h(c + 1, c) = 0.0;
    q.S[c + 1] = -q.sn[c] * q.S[c];
    q.S[c] = q.cs[c] * q.S[c];
    q.presid = std::fabs(q.S[c + 1]);
    ++iterations[i];
    q.used = c + 1;
    q.stop = q.presid <= q.ptol || q.breakdown;
    return C3L_OK;
  };
  // orthogonalisation step over a list of systems, at most four per launch
  auto step_over = [&](std::vector<int32_t> const& which, int j, bool has_w) -> int {
    for (size_t k = 0; k < which.size(); k += c3l_ortho::kMaxInstances) {
      const int n = (int)std::min<size_t>(c3l_ortho::kMaxInstances, which.size() - k);
      const double* wz[c3l_ortho::kMaxInstances];
      c3l_ortho::Basis<double, Store> bs[c3l_ortho::kMaxInstances];
      for (int u = 0; u < n; ++u) {
        const int32_t i = which[k + (size_t)u];
        wz[u] = has_w ? w + i * rd : nullptr;
        if constexpr (std::is_same<Store, float>::value) bs[u] = sys[(size_t)i].b32;
        else bs[u] = sys[(size_t)i].b64;
      }
      D_CALL(c3l_ortho::step_many(bs, n, j, j == 0, wz, 0.0, s, nullptr));
    }
    return C3L_OK;
  };
  for (int32_t cycle = 0; cycle < cycles; ++cycle) {
    live = live_list();
    if (live.empty()) break;
    D_CALL(g.precondition(w, tmp));
    D_CALL(segment_norms(tmp, live));
    for (int32_t i : live) {
      GroupSystem& q = sys[(size_t)i];
      const double beta = norms[(size_t)i];
      if (!(beta > 0.0) || !std::isfinite(beta))
        return dfail(C3L_EINVAL, "c3l_gmres_group: singular preconditioned residual");
      if constexpr (std::is_same<Store, float>::value)
        D_CALL(c3l_internal_narrow_scale(tmp + i * rd, q.b32.V, 1.0 / beta, rd, stream));
      else
        D_CALL(c3l_rscale(tmp + i * rd, q.b64.V, 1.0 / beta, rd, stream));
      std::fill(q.H.begin(), q.H.end(), 0.0);
      std::fill(q.S.begin(), q.S.end(), 0.0);
      q.S[0] = beta;
      q.used = 0;
      q.absorbed = 0;
      q.stop = q.breakdown = false;
      q.presid = 0.0;
    }
    for (int j = 0; j <= R; ++j) {
      std::vector<int32_t> iterating;
      for (int32_t i : live)
        if (!sys[(size_t)i].stop) iterating.push_back(i);
      if (iterating.empty()) break;
      const bool has_w = j < R;
      if (has_w) {
        D_CALL(g.apply_preconditioned(V + (int64_t)j * ldv, w));
        for (int32_t i : iterating) ++jvp_calls[i];
      }
      const int c = j - 1;
      for (int32_t i : iterating) {  // the update pass writes column c into its pinned slot
        GroupSystem& q = sys[(size_t)i];
        q.b32.stage = j == 0 ? nullptr : slot_dev(i, c);
        q.b64.stage = q.b32.stage;
      }
      D_CALL(step_over(iterating, j, has_w));
      if (j == 0) continue;
      for (int32_t i : iterating) D_CHECK(hipEventRecord(events[(size_t)i * 2 + (c % 2)], s));
      if (c >= 1)
        for (int32_t i : iterating) {
          D_CALL(absorb(i, c - 1));
          sys[(size_t)i].absorbed = c;
        }
    }
    for (int32_t i : live) {  // columns still in flight at the end of the cycle
      GroupSystem& q = sys[(size_t)i];
      while (!q.stop && q.absorbed < R) {
        D_CALL(absorb(i, q.absorbed));
        ++q.absorbed;
      }
    }
    for (int32_t i : live) {
      GroupSystem& q = sys[(size_t)i];
      auto h = [&](int r, int k) -> double { return q.H[(size_t)r * R + k]; };
      const int used = q.used;
      if (used == 0) continue;
      if (h(used - 1, used - 1) == 0.0) q.S[used - 1] = 0.0;
      for (int k = used - 1; k >= 0; --k) {
        double acc = q.S[k];
        for (int l = k + 1; l < used; ++l) acc -= h(k, l) * q.y[l];
        q.y[k] = h(k, k) == 0.0 ? 0.0 : acc / h(k, k);
      }
      D_CHECK(hipMemcpyAsync(q.coef + L.y, q.y.data(), sizeof(double) * used, hipMemcpyHostToDevice, s));
      if constexpr (std::is_same<Store, float>::value)
        D_CALL(c3l_internal_rmultiaxpy_f32(q.b32.V, ldv, used, q.coef + L.y, 1.0, x + i * rd, rd, stream));
      else
        D_CALL(c3l_rmultiaxpy(q.b64.V, ldv, used, q.coef + L.y, 1.0, x + i * rd, rd, stream));
    }
    D_CALL(g.apply(x, w));
    for (int32_t i : live) ++jvp_calls[i];
    D_CALL(c3l_raxpby(1.0, b, -1.0, w, ldv, stream));
    D_CALL(segment_norms(w, live));  // synchronises: y may be reused by the next cycle
    for (int32_t i : live) {
      GroupSystem& q = sys[(size_t)i];
      q.rn = norms[(size_t)i];
      if (q.rn <= q.atol || q.breakdown) {
        q.done = true;
        continue;
      }
      q.ptol_factor = q.presid <= q.ptol ? std::max(eps, 0.25 * q.ptol_factor) : std::min(1.0, 1.5 * q.ptol_factor);
      q.ptol = q.presid * std::min(q.ptol_factor, q.atol / q.rn);
    }
  }
  for (int32_t i = 0; i < count; ++i) {
    GroupSystem& q = sys[(size_t)i];
    if (!active[i] || q.bnrm == 0.0 || q.trivial) continue;
    true_rel[i] = q.rn / q.bnrm;
    info[i] = q.rn <= q.atol ? 0 : cycles;
  }
  return C3L_OK;
}

}  // namespace

// One batched Arnoldi resolvent in flight (c3l_arnoldi_begin / _drive / _end): the
// per-basis device pointers, the lockstep position and the bases still running.
// Stop flags travel to pinned host memory with an event per shift batch.  The host
// reads a batch's flags as soon as its event has completed (polled once per step) and
// waits for it only when the next batch is enqueued, so it never drains its stream
// while other resolvents have work queued.
struct c3l_arnoldi_run {
  std::vector<c3l_plan*> plans;
  std::vector<c3l_ortho::Basis<double2>> basis;
  std::vector<ShiftState> state;
  std::vector<double*> w;  // sigma output (scratch) per basis
  std::vector<double*> V;  // basis columns per basis
  std::vector<int32_t> live;
  const double* diag = nullptr;
  const double2* shifts = nullptr;
  double* status = nullptr;
  int64_t dim = 0;
  int64_t ld = 0;  // basis leading dimension (complex entries), c3l_krylov_basis_ld(dim)
  int32_t count = 0, m_max = 0, n_shifts = 0, min_dim = 0;
  double tolerance = 0.0, breakdown = 0.0;
  hipStream_t stream = nullptr;
  int32_t j = 0, c0 = 0;
  // the update pass of step j scatters q_{j+1} into each plan's padded input blocks, so the
  // sigma of step j + 1 skips its scatter (plans whose input entries fill at most two
  // padded elements each)
  bool fused_scatter = false;
  double* pinned = nullptr;  // two slots of 4 * count doubles
  hipEvent_t copied[2] = {nullptr, nullptr};
  int32_t batches = 0;     // status copies enqueued so far
  int32_t flags_read = 0;  // status copies read so far
  DriverClock clock{0};
  // Group runs (c3l_arnoldi_group_begin): bases share plan slots.  Basis i belongs to slot
  // slot_of[i] at row offset row_of[i] (complex entries) of the slot's vectors; the slot's
  // columns have slot_ld complex entries and its sigma output is slot_w.  Empty otherwise.
  std::vector<int32_t> slot_of;
  std::vector<int64_t> row_of;
  std::vector<c3l_plan*> slot_plan;
  std::vector<double*> slot_V, slot_w;
  int64_t slot_ld = 0;

  bool done() const { return live.empty() || j > m_max; }

  ~c3l_arnoldi_run() {
    // stop-flag copies into `pinned` may still be queued when a run ends early
    if (stream != nullptr) (void)hipStreamSynchronize(stream);
    for (hipEvent_t& e : copied)
      if (e != nullptr) (void)hipEventDestroy(e);
    if (pinned != nullptr) (void)hipHostFree(pinned);
  }

  // Reads every status copy whose event has completed, and waits for the oldest one
  // when both slots are in use; drops the bases whose stop flag is set.
  int poll() {
    while (flags_read < batches) {
      const int slot = flags_read & 1;
      if (batches - flags_read < 2) {
        const hipError_t e = hipEventQuery(copied[slot]);
        if (e == hipErrorNotReady) break;
        D_CHECK(e);
      }
      D_CHECK(hipEventSynchronize(copied[slot]));
      const double* h = pinned + (size_t)slot * 4 * count;
      std::vector<int32_t> next;
      for (int32_t i : live)
        if (h[(size_t)i * 4 + 2] == 0.0) next.push_back(i);
      live.swap(next);
      ++flags_read;
    }
    return C3L_OK;
  }

  // Step j of a group run: one sigma launch over the slots that still have a live basis,
  // the orthogonalisation and shift updates over the live bases, four per launch.
  int advance_group() {
    const bool has_w = j < m_max;
    std::vector<int32_t> slots;
    for (int32_t i : live)
      if (std::find(slots.begin(), slots.end(), slot_of[(size_t)i]) == slots.end()) slots.push_back(slot_of[(size_t)i]);
    if (has_w) {
      c3l_plan* lp[c3l_ortho::kMaxInstances];
      const double* vj[c3l_ortho::kMaxInstances];
      double* wj[c3l_ortho::kMaxInstances];
      const int ns = (int)slots.size();
      for (int u = 0; u < ns; ++u) {
        lp[u] = slot_plan[(size_t)slots[(size_t)u]];
        vj[u] = slot_V[(size_t)slots[(size_t)u]] + (int64_t)2 * j * slot_ld;
        wj[u] = slot_w[(size_t)slots[(size_t)u]];
      }
      D_CALL(c3l_internal_plan_load_many(lp, ns, 2, vj, stream));
      D_CALL(c3l_internal_plan_apply_many(lp, ns, 3, diag, vj, wj, stream));
    }
    const bool batch = j > 0 && (j - c0 == kShiftBatch || j == m_max);
    const int nb = j - c0;
    for (size_t k0 = 0; k0 < live.size(); k0 += c3l_ortho::kMaxInstances) {
      const int k = (int)std::min<size_t>(c3l_ortho::kMaxInstances, live.size() - k0);
      const double2* wz[c3l_ortho::kMaxInstances];
      c3l_ortho::Basis<double2> lb[c3l_ortho::kMaxInstances];
      ShiftStates ls{};
      for (int u = 0; u < k; ++u) {
        const int32_t i = live[k0 + (size_t)u];
        lb[u] = basis[(size_t)i];
        wz[u] = has_w ? reinterpret_cast<const double2*>(slot_w[(size_t)slot_of[(size_t)i]]) + row_of[(size_t)i] : nullptr;
        ls.p[u] = state[(size_t)i];
      }
      D_CALL(c3l_ortho::step_many(lb, k, j, j == 0, wz, breakdown, stream, &clock));
      if (batch) {
        const int64_t shift_blocks = (n_shifts + kShiftsPerBlock - 1) / kShiftsPerBlock;
        hipLaunchKernelGGL(multishift_batch_kernel, dim3(shift_blocks, 1, k), dim3(256), 0, stream, ls, c0, nb, m_max,
                           shifts, n_shifts);
        D_CHECK(hipGetLastError());
        hipLaunchKernelGGL(arnoldi_batch_status_kernel, dim3(1, 1, k), dim3(1024), 0, stream, ls, n_shifts, m_max, c0,
                           nb, min_dim, tolerance, breakdown);
        D_CHECK(hipGetLastError());
      }
    }
    if (batch) {
      c0 = j;
      const int slot = batches & 1;
      D_CHECK(hipMemcpyAsync(pinned + (size_t)slot * 4 * count, status, 4 * sizeof(double) * (size_t)count,
                             hipMemcpyDeviceToHost, stream));
      D_CHECK(hipEventRecord(copied[slot], stream));
      ++batches;
    }
    ++j;
    return poll();
  }

  // Step j: sigma on the once-projected v_j of every live basis (none at j = m_max),
  // one orthogonalisation step, and every kShiftBatch columns the shift update.
  int advance() {
    if (!slot_of.empty()) return advance_group();
    const int k = (int)live.size();
    const bool has_w = j < m_max;
    c3l_plan* lp[c3l_ortho::kMaxInstances];
    const double* vj[c3l_ortho::kMaxInstances];
    double* wj[c3l_ortho::kMaxInstances];
    const double2* wz[c3l_ortho::kMaxInstances];
    c3l_ortho::Basis<double2> lb[c3l_ortho::kMaxInstances];
    ShiftStates ls{};
    for (int u = 0; u < k; ++u) {
      const int32_t i = live[(size_t)u];
      lp[u] = plans[(size_t)i];
      vj[u] = V[(size_t)i] + (int64_t)2 * j * ld;
      wj[u] = w[(size_t)i];
      wz[u] = has_w ? reinterpret_cast<const double2*>(wj[u]) : nullptr;
      lb[u] = basis[(size_t)i];
      ls.p[u] = state[(size_t)i];
    }
    if (has_w) {
      if (fused_scatter && j > 0)
        D_CALL(c3l_internal_plan_load_prepared_many(lp, k, 2, stream));
      else
        D_CALL(c3l_internal_plan_load_many(lp, k, 2, vj, stream));
      D_CALL(c3l_internal_plan_apply_many(lp, k, 3, diag, vj, wj, stream));
    }
    D_CALL(c3l_ortho::step_many(lb, k, j, j == 0, wz, breakdown, stream, &clock));
    if (j > 0 && (j - c0 == kShiftBatch || j == m_max)) {
      const int nb = j - c0;
      const int64_t shift_blocks = (n_shifts + kShiftsPerBlock - 1) / kShiftsPerBlock;
      hipLaunchKernelGGL(multishift_batch_kernel, dim3(shift_blocks, 1, k), dim3(256), 0, stream, ls, c0, nb, m_max,
                         shifts, n_shifts);
      D_CHECK(hipGetLastError());
      hipLaunchKernelGGL(arnoldi_batch_status_kernel, dim3(1, 1, k), dim3(1024), 0, stream, ls, n_shifts, m_max, c0,
                         nb, min_dim, tolerance, breakdown);
      D_CHECK(hipGetLastError());
      c0 = j;
      const int slot = batches & 1;
      D_CHECK(hipMemcpyAsync(pinned + (size_t)slot * 4 * count, status, 4 * sizeof(double) * (size_t)count,
                             hipMemcpyDeviceToHost, stream));
      D_CHECK(hipEventRecord(copied[slot], stream));
      ++batches;
    }
    ++j;
    return poll();
  }
};

namespace {

// Sets up `count` bases of c3l_arnoldi_batch on `stream` (validation, per-basis pointers,
// shift-state initialisation).
int arnoldi_setup(c3l_arnoldi_run* r, c3l_plan* const* plans, int32_t count, const double* diag, double* V, double* H,
                  int64_t dim, int32_t m_max, const double* shifts, int32_t n_shifts, const double* sink,
                  const double* beta, double* rot_c, double* rot_s, double* pi, double* gamma, double* green,
                  double* resid, double* scratch, double* coef, double* status, double tolerance, int32_t min_dim,
                  double breakdown_tol, int32_t criterion, void* stream) {
  if (criterion != 0 && criterion != 1) return dfail(C3L_EINVAL, "c3l_arnoldi: criterion must be 0 or 1");
  if (plans == nullptr || count < 1 || count > c3l_ortho::kMaxInstances || V == nullptr || H == nullptr || dim <= 0 ||
      m_max <= 0 || n_shifts <= 0 || shifts == nullptr || sink == nullptr || beta == nullptr || rot_c == nullptr ||
      rot_s == nullptr || pi == nullptr || gamma == nullptr || green == nullptr || resid == nullptr ||
      scratch == nullptr || coef == nullptr || status == nullptr)
    return dfail(C3L_EINVAL, "c3l_arnoldi: invalid argument");
  for (int32_t i = 0; i < count; ++i)
    if (plans[i] == nullptr || !(beta[i] > 0.0)) return dfail(C3L_EINVAL, "c3l_arnoldi: invalid argument");
  if (!c3l_internal_plans_are_twins(plans, count, 2, 3))
    return dfail(C3L_EINVAL, "c3l_arnoldi: the bases' plans must be FP64 twins (built from the same arrays)");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const Layout L = arnoldi_layout(dim, m_max, n_shifts);
  // per-basis strides (doubles) of the caller's consecutive instance blocks; basis columns are
  // padded to c3l_krylov_basis_ld(dim) complex entries so that each starts on a cache line
  const int64_t ld = c3l_krylov_basis_ld(dim);
  const int64_t sV = 2 * ((int64_t)m_max + 1) * ld, sH = 2 * (int64_t)m_max * ((int64_t)m_max + 1);
  const int64_t sRot = (int64_t)n_shifts * m_max, sShift = n_shifts;
  r->plans.assign(plans, plans + count);
  r->count = count;
  r->diag = diag;
  r->shifts = reinterpret_cast<const double2*>(shifts);
  r->status = status;
  r->dim = dim;
  r->ld = ld;
  r->m_max = m_max;
  r->n_shifts = n_shifts;
  r->min_dim = min_dim;
  r->tolerance = tolerance;
  r->breakdown = breakdown_tol;
  r->stream = s;
  for (int32_t i = 0; i < count; ++i) {
    double* sc = scratch + i * L.scratch_total;
    double* cf = coef + i * L.coef_total;
    double2* Hz = reinterpret_cast<double2*>(H + i * sH);
    double2* sink_v = reinterpret_cast<double2*>(cf + L.sink_v);
    r->basis.push_back(c3l_ortho::Basis<double2>{reinterpret_cast<double2*>(V + i * sV), ld, dim, Hz,
                                                 (int64_t)m_max + 1, sink_v,
                                                 reinterpret_cast<const double2*>(sink + 2 * i * dim),
                                                 reinterpret_cast<double2*>(sc + L.partial),
                                                 reinterpret_cast<double2*>(cf + L.red),
                                                 reinterpret_cast<double2*>(cf + L.sc), status + 4 * i, nullptr});
    r->state.push_back(ShiftState{Hz, sink_v, beta[i], rot_c + i * sRot,
                                  reinterpret_cast<double2*>(rot_s + 2 * i * sRot),
                                  reinterpret_cast<double2*>(pi + 2 * i * sRot),
                                  reinterpret_cast<double2*>(gamma + 2 * i * sShift),
                                  reinterpret_cast<double2*>(green + 2 * i * sShift), resid + i * sShift,
                                  sc + L.resid_hist, reinterpret_cast<double2*>(sc + L.green_hist), status + 4 * i,
                                  (int)criterion});
    r->w.push_back(sc + L.w);
    r->V.push_back(V + i * sV);
    r->live.push_back(i);
    hipLaunchKernelGGL(init_shift_state_kernel, dim3(blocks_for(n_shifts)), dim3(256), 0, s, r->state.back().gamma,
                       r->state.back().green, beta[i], n_shifts, status + 4 * i);
    D_CHECK(hipGetLastError());
  }
  // twins share their padded-block layout, so one map serves every basis
  const int64_t* map = nullptr;
  D_CALL(c3l_internal_plan_scatter_map(plans[0], 2, dim, &map));
  if (map != nullptr) {
    r->fused_scatter = true;
    for (int32_t i = 0; i < count; ++i) {
      r->basis[(size_t)i].pad = reinterpret_cast<double2*>(c3l_internal_plan_workspace(plans[i]));
      r->basis[(size_t)i].pad_map = map;
    }
  }
  D_CHECK(hipHostMalloc(reinterpret_cast<void**>(&r->pinned), 2 * 4 * sizeof(double) * (size_t)count));
  for (hipEvent_t& e : r->copied) D_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return C3L_OK;
}

// Sets up the bases of c3l_arnoldi_group_begin: slot s (0 <= s < n_slots) holds per_slot bases;
// basis i = s * per_slot + p reads rows [p dim, (p+1) dim) of the slot's vectors.
int arnoldi_group_setup(c3l_arnoldi_run* r, c3l_plan* const* slot_plans, int32_t n_slots, int32_t per_slot,
                        const double* diag, double* V, double* W, double* H, int64_t dim, int32_t m_max,
                        const double* shifts, int32_t n_shifts, const double* sink, const double* beta,
                        double* rot_c, double* rot_s, double* pi, double* gamma, double* green, double* resid,
                        double* scratch, double* coef, double* status, double tolerance, int32_t min_dim,
                        double breakdown_tol, void* stream) {
  const int32_t count = n_slots * per_slot;
  if (slot_plans == nullptr || n_slots < 1 || n_slots > c3l_ortho::kMaxInstances || per_slot < 1 ||
      count > kMaxGroupBases || V == nullptr || W == nullptr || H == nullptr || dim <= 0 || m_max <= 0 ||
      n_shifts <= 0 || shifts == nullptr || sink == nullptr || beta == nullptr || rot_c == nullptr ||
      rot_s == nullptr || pi == nullptr || gamma == nullptr || green == nullptr || resid == nullptr ||
      scratch == nullptr || coef == nullptr || status == nullptr)
    return dfail(C3L_EINVAL, "c3l_arnoldi_group_begin: invalid argument");
  for (int32_t i = 0; i < n_slots; ++i)
    if (slot_plans[i] == nullptr) return dfail(C3L_EINVAL, "c3l_arnoldi_group_begin: invalid argument");
  for (int32_t i = 0; i < count; ++i)
    if (!(beta[i] > 0.0)) return dfail(C3L_EINVAL, "c3l_arnoldi_group_begin: every right-hand side must be nonzero");
  if (!c3l_internal_plans_are_twins(slot_plans, n_slots, 2, 3))
    return dfail(C3L_EINVAL, "c3l_arnoldi_group_begin: the slots' plans must be FP64 twins (built from the same arrays)");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const Layout L = arnoldi_layout(dim, m_max, n_shifts);
  const int64_t ld = (int64_t)per_slot * dim;                      // complex entries per slot column
  const int64_t sV = 2 * ((int64_t)m_max + 1) * ld;                 // doubles per slot basis block
  const int64_t sH = 2 * (int64_t)m_max * ((int64_t)m_max + 1);
  const int64_t sRot = (int64_t)n_shifts * m_max, sShift = n_shifts;
  r->plans.clear();
  r->count = count;
  r->diag = diag;
  r->shifts = reinterpret_cast<const double2*>(shifts);
  r->status = status;
  r->dim = dim;
  r->m_max = m_max;
  r->n_shifts = n_shifts;
  r->min_dim = min_dim;
  r->tolerance = tolerance;
  r->breakdown = breakdown_tol;
  r->stream = s;
  r->slot_ld = ld;
  for (int32_t q = 0; q < n_slots; ++q) {
    r->slot_plan.push_back(slot_plans[q]);
    r->slot_V.push_back(V + q * sV);
    r->slot_w.push_back(W + 2 * q * ld);
  }
  for (int32_t i = 0; i < count; ++i) {
    const int32_t q = i / per_slot, p = i % per_slot;
    double* sc = scratch + i * L.scratch_total;
    double* cf = coef + i * L.coef_total;
    double2* Hz = reinterpret_cast<double2*>(H + i * sH);
    double2* sink_v = reinterpret_cast<double2*>(cf + L.sink_v);
    double2* Vi = reinterpret_cast<double2*>(V + q * sV) + (int64_t)p * dim;
    r->basis.push_back(c3l_ortho::Basis<double2>{Vi, ld, dim, Hz, (int64_t)m_max + 1, sink_v,
                                                 reinterpret_cast<const double2*>(sink) + (int64_t)i * dim,
                                                 reinterpret_cast<double2*>(sc + L.partial),
                                                 reinterpret_cast<double2*>(cf + L.red),
                                                 reinterpret_cast<double2*>(cf + L.sc), status + 4 * i, nullptr});
    r->state.push_back(ShiftState{Hz, sink_v, beta[i], rot_c + i * sRot,
                                  reinterpret_cast<double2*>(rot_s + 2 * i * sRot),
                                  reinterpret_cast<double2*>(pi + 2 * i * sRot),
                                  reinterpret_cast<double2*>(gamma + 2 * i * sShift),
                                  reinterpret_cast<double2*>(green + 2 * i * sShift), resid + i * sShift,
                                  sc + L.resid_hist, reinterpret_cast<double2*>(sc + L.green_hist), status + 4 * i, 0});
    r->slot_of.push_back(q);
    r->row_of.push_back((int64_t)p * dim);
    r->w.push_back(W + 2 * (q * ld + (int64_t)p * dim));
    r->V.push_back(reinterpret_cast<double*>(Vi));
    r->live.push_back(i);
    hipLaunchKernelGGL(init_shift_state_kernel, dim3(blocks_for(n_shifts)), dim3(256), 0, s, r->state.back().gamma,
                       r->state.back().green, beta[i], n_shifts, status + 4 * i);
    D_CHECK(hipGetLastError());
  }
  D_CHECK(hipHostMalloc(reinterpret_cast<void**>(&r->pinned), 2 * 4 * sizeof(double) * (size_t)count));
  for (hipEvent_t& e : r->copied) D_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return C3L_OK;
}

}  // namespace

extern "C" {

int c3l_arnoldi_group_begin(c3l_plan* const* slot_plans, int32_t n_slots, int32_t per_slot, const double* diag,
                            double* V, double* W, double* H, int64_t dim, int32_t m_max, const double* shifts,
                            int32_t n_shifts, const double* sink, const double* beta, double* rot_c, double* rot_s,
                            double* pi, double* gamma, double* green, double* resid, double* scratch, double* coef,
                            double* status, double tolerance, int32_t min_dim, double breakdown_tol, void* stream,
                            c3l_arnoldi_run** out) {
  if (out == nullptr) return dfail(C3L_EINVAL, "c3l_arnoldi_group_begin: invalid argument");
  *out = nullptr;
  c3l_arnoldi_run* r = new c3l_arnoldi_run();
  const int rc = arnoldi_group_setup(r, slot_plans, n_slots, per_slot, diag, V, W, H, dim, m_max, shifts, n_shifts,
                                     sink, beta, rot_c, rot_s, pi, gamma, green, resid, scratch, coef, status,
                                     tolerance, min_dim, breakdown_tol, stream);
  if (rc != C3L_OK) {
    delete r;
    return rc;
  }
  *out = r;
  return C3L_OK;
}

int64_t c3l_krylov_basis_ld(int64_t rows) { return rows <= 0 ? 0 : (rows + 63) / 64 * 64; }

int c3l_gmres(c3l_plan* plan, const c3l_tilde_maps* tm, const double* jvp_diag, double* t_work, double* out_work,
              const double* b, double* x, void* V, int32_t basis_bits, double* w, double* tmp, double* scratch,
              double* coef, double rtol, double atol, int32_t restart, int32_t cycles, int32_t* info,
              int32_t* iterations, double* true_rel, int32_t* jvp_calls, void* stream) {
  if (basis_bits == 64)
    return gmres_run(plan, tm, jvp_diag, t_work, out_work, b, x, static_cast<double*>(V), w, tmp, scratch, coef, rtol,
                     atol, restart, cycles, info, iterations, true_rel, jvp_calls, stream);
  if (basis_bits == 32)
    return gmres_run(plan, tm, jvp_diag, t_work, out_work, b, x, static_cast<float*>(V), w, tmp, scratch, coef, rtol,
                     atol, restart, cycles, info, iterations, true_rel, jvp_calls, stream);
  return dfail(C3L_EINVAL, "c3l_gmres: basis_bits must be 32 or 64");
}

int c3l_gmres_group(c3l_plan* plan, const c3l_tilde_maps* tm, int32_t count, const double* jvp_diag, double* t_work,
                    double* out_work, const double* b, double* x, void* V, int32_t basis_bits, double* w, double* tmp,
                    double* scratch, double* coef, const double* rtol, const double* atol, const int32_t* active,
                    int32_t restart, int32_t cycles, int32_t* info, int32_t* iterations, double* true_rel,
                    int32_t* jvp_calls, void* stream) {
  if (basis_bits == 64)
    return gmres_group_run(plan, tm, count, jvp_diag, t_work, out_work, b, x, static_cast<double*>(V), w, tmp, scratch,
                           coef, rtol, atol, active, restart, cycles, info, iterations, true_rel, jvp_calls, stream);
  if (basis_bits == 32)
    return gmres_group_run(plan, tm, count, jvp_diag, t_work, out_work, b, x, static_cast<float*>(V), w, tmp, scratch,
                           coef, rtol, atol, active, restart, cycles, info, iterations, true_rel, jvp_calls, stream);
  return dfail(C3L_EINVAL, "c3l_gmres_group: basis_bits must be 32 or 64");
}

int c3l_krylov_workspace(int32_t complex_basis, int64_t dim, int32_t max_columns, int32_t n_shifts,
                         int64_t* scratch_doubles, int64_t* coef_doubles) {
  if (dim <= 0 || max_columns <= 0 || n_shifts < 0 || scratch_doubles == nullptr || coef_doubles == nullptr)
    return dfail(C3L_EINVAL, "c3l_krylov_workspace: invalid argument");
  const Layout L = complex_basis ? arnoldi_layout(dim, max_columns, n_shifts) : gmres_layout(dim, max_columns);
  *scratch_doubles = L.scratch_total;
  *coef_doubles = L.coef_total;
  return C3L_OK;
}

int c3l_krylov_set_timing(int32_t enable) {
  KrylovTimer& t = ktimer();
  std::lock_guard<std::mutex> guard(t.mu);
  t.enabled = enable != 0;
  for (int k = 0; k < 4; ++k) t.ms[k] = t.bytes[k] = 0.0, t.launches[k] = 0;
  return C3L_OK;
}

int c3l_krylov_timing(int32_t family, double* ms_total, int64_t* launches, double* bytes) {
  if (family < 0 || family > 3) return dfail(C3L_EINVAL, "c3l_krylov_timing: invalid family");
  KrylovTimer& t = ktimer();
  std::lock_guard<std::mutex> guard(t.mu);
  for (KrylovTimer::Pending& p : t.pending) {
    D_CHECK(hipEventSynchronize(p.stop));
    float ms = 0.f;
    D_CHECK(hipEventElapsedTime(&ms, p.start, p.stop));
    t.ms[p.which] += ms;
    t.bytes[p.which] += p.bytes;
    t.launches[p.which] += 1;
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.stop);
  }
  t.pending.clear();
  if (ms_total) *ms_total = t.ms[family];
  if (launches) *launches = t.launches[family];
  if (bytes) *bytes = t.bytes[family];
  return C3L_OK;
}

int c3l_arnoldi(c3l_plan* plan, const double* diag, double* V, double* H, int64_t dim, int32_t m_max,
                const double* shifts, int32_t n_shifts, const double* sink, double beta, double* rot_c,
                double* rot_s, double* pi, double* gamma, double* green, double* resid, double* scratch,
                double* coef, double* status, double tolerance, int32_t min_dim, double breakdown_tol,
                int32_t criterion, int32_t* m_out, int32_t* breakdown_out, double* worst_out, void* stream) {
  if (plan == nullptr || out_invalid(m_out, breakdown_out, worst_out))
    return dfail(C3L_EINVAL, "c3l_arnoldi: invalid argument");
  c3l_arnoldi_run* run = new c3l_arnoldi_run();
  const int rc = arnoldi_setup(run, &plan, 1, diag, V, H, dim, m_max, shifts, n_shifts, sink, &beta, rot_c, rot_s, pi,
                               gamma, green, resid, scratch, coef, status, tolerance, min_dim, breakdown_tol,
                               criterion, stream);
  if (rc != C3L_OK) {
    delete run;
    return rc;
  }
  const int rc_drive = c3l_arnoldi_drive(&run, 1);
  const int rc_end = c3l_arnoldi_end(run, m_out, breakdown_out, worst_out);
  return rc_drive != C3L_OK ? rc_drive : rc_end;
}

int c3l_arnoldi_batch(c3l_plan* const* plans, int32_t count, const double* diag, double* V, double* H, int64_t dim,
                      int32_t m_max, const double* shifts, int32_t n_shifts, const double* sink, const double* beta,
                      double* rot_c, double* rot_s, double* pi, double* gamma, double* green, double* resid,
                      double* scratch, double* coef, double* status, double tolerance, int32_t min_dim,
                      double breakdown_tol, int32_t* m_out, int32_t* breakdown_out, double* worst_out, void* stream) {
  c3l_arnoldi_run* run = nullptr;
  D_CALL(c3l_arnoldi_begin(plans, count, diag, V, H, dim, m_max, shifts, n_shifts, sink, beta, rot_c, rot_s, pi, gamma,
                           green, resid, scratch, coef, status, tolerance, min_dim, breakdown_tol, stream, &run));
  const int rc = c3l_arnoldi_drive(&run, 1);
  const int rc_end = c3l_arnoldi_end(run, m_out, breakdown_out, worst_out);
  return rc != C3L_OK ? rc : rc_end;
}

int c3l_arnoldi_begin(c3l_plan* const* plans, int32_t count, const double* diag, double* V, double* H, int64_t dim,
                      int32_t m_max, const double* shifts, int32_t n_shifts, const double* sink, const double* beta,
                      double* rot_c, double* rot_s, double* pi, double* gamma, double* green, double* resid,
                      double* scratch, double* coef, double* status, double tolerance, int32_t min_dim,
                      double breakdown_tol, void* stream, c3l_arnoldi_run** out) {
  if (out == nullptr) return dfail(C3L_EINVAL, "c3l_arnoldi_begin: invalid argument");
  *out = nullptr;
  c3l_arnoldi_run* r = new c3l_arnoldi_run();
  const int rc = arnoldi_setup(r, plans, count, diag, V, H, dim, m_max, shifts, n_shifts, sink, beta, rot_c, rot_s,
                               pi, gamma, green, resid, scratch, coef, status, tolerance, min_dim, breakdown_tol, 0,
                               stream);
  if (rc != C3L_OK) {
    delete r;
    return rc;
  }
  *out = r;
  return C3L_OK;
}

int c3l_arnoldi_drive(c3l_arnoldi_run* const* runs, int32_t n_runs) {
  if (runs == nullptr || n_runs < 0) return dfail(C3L_EINVAL, "c3l_arnoldi_drive: invalid argument");
  for (int32_t r = 0; r < n_runs; ++r)
    if (runs[r] == nullptr) return dfail(C3L_EINVAL, "c3l_arnoldi_drive: invalid argument");
  // one step of every unfinished resolvent in turn, so that all streams stay fed
  for (bool any = true; any;) {
    any = false;
    for (int32_t r = 0; r < n_runs; ++r) {
      if (runs[r]->done()) continue;
      D_CALL(runs[r]->advance());
      any = true;
    }
  }
  return C3L_OK;
}

int c3l_arnoldi_end(c3l_arnoldi_run* run, int32_t* m_out, int32_t* breakdown_out, double* worst_out) {
  if (run == nullptr) return dfail(C3L_EINVAL, "c3l_arnoldi_end: invalid argument");
  std::unique_ptr<c3l_arnoldi_run> owned(run);
  if (m_out == nullptr || breakdown_out == nullptr || worst_out == nullptr)
    return dfail(C3L_EINVAL, "c3l_arnoldi_end: invalid argument");
  // the final flags (frozen after each basis's stop), read after the stream has drained
  std::vector<double> host((size_t)run->count * 4);
  D_CHECK(hipMemcpyAsync(host.data(), run->status, sizeof(double) * host.size(), hipMemcpyDeviceToHost, run->stream));
  D_CHECK(hipStreamSynchronize(run->stream));
  for (int32_t i = 0; i < run->count; ++i)
    if (host[(size_t)i * 4 + 2] == 3.0) return dfail(C3L_EINVAL, "c3l_arnoldi: nonfinite Arnoldi recurrence");
  for (int32_t i = 0; i < run->count; ++i) {
    const double* h = host.data() + (size_t)i * 4;
    m_out[i] = (int32_t)h[3];
    breakdown_out[i] = h[2] == 2.0 ? 1 : 0;
    worst_out[i] = h[1];
  }
  return C3L_OK;
}

}  // extern "C"
