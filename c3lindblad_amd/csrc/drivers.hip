// Native Krylov drivers: restarted GMRES for the Newton correction and the
// shift-invariant Arnoldi resolvent, both running their whole loop in C++ so
// that the host only sees one call per linear solve / per Krylov basis.
//
// GMRES: left-preconditioned restarted GMRES with two-pass classical
// Gram-Schmidt on the device; the (j+2)-number Hessenberg column is copied to
// the host once per iteration for the Givens update (SciPy semantics: inner
// tolerance ||M b|| * rtol, true residual test at every restart).
//
// Arnoldi: two-pass CGS on the device; the Hessenberg matrix stays on the
// device; for every shift z a Givens QR of (z I - H) is advanced by one
// column per step (one thread per shift), giving the exact FOM residual
// |h_{m+1,m} e_m^T (z - H_m)^{-1} beta e_1| / beta = rho_GMRES(z) / c_m(z)
// at O(m) cost per shift and step.  One 16-byte transfer per step carries
// the new sub-diagonal entry and the worst shift residual to the host.

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/c3l_hip.h"

// Shared error slot of the library (defined in c3l_hip.hip).
extern "C" int c3l_internal_fail(int code, const char* message);

namespace {

int dfail(int code, const std::string& msg) { return c3l_internal_fail(code, msg.c_str()); }

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

// H[k, j] = c1[k] + c2[k] (k <= j); H[j+1, j] = sqrt(s); status[0] = H[j+1, j].
__global__ void finish_column_kernel(const double2* __restrict__ coef, int j, int m_max, double2* __restrict__ H,
                                     double* __restrict__ status) {
  const double2* c1 = coef;
  const double2* c2 = coef + (m_max + 1);
  double2* col = H + (int64_t)j * (m_max + 1);
  for (int k = threadIdx.x; k <= j; k += blockDim.x) col[k] = make_double2(c1[k].x + c2[k].x, c1[k].y + c2[k].y);
  if (threadIdx.x == 0) {
    const double s = coef[2 * (m_max + 1)].x;
    const double hn = sqrt(s > 0.0 ? s : 0.0);
    col[j + 1] = make_double2(hn, 0.0);
    status[0] = hn;
  }
}

// V_next = w / hn (or 0 after a breakdown), hn read from device memory.
__global__ void normalize_kernel(const double2* __restrict__ w, double2* __restrict__ v, const double* __restrict__ hn,
                                 double breakdown, int64_t dim) {
  const double h = hn[0];
  const double a = h > breakdown ? 1.0 / h : 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < dim; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = make_double2(a * w[i].x, a * w[i].y);
}

// One thread per shift: advance the Givens QR of (z I - H) by column j.
// rot_c [m_max][n_shifts] (real cosines), rot_s [m_max][n_shifts] (complex sines),
// rho [n_shifts] running product of |s_k|, resid [n_shifts] FOM relative residual.
__global__ void shifted_givens_kernel(const double2* __restrict__ H, int j, int m_max,
                                      const double2* __restrict__ shifts, int n_shifts, double* __restrict__ rot_c,
                                      double2* __restrict__ rot_s, double* __restrict__ rho, double* __restrict__ resid) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_shifts) return;
  const double2 z = shifts[s];
  const double2* col = H + (int64_t)j * (m_max + 1);
  double2 x = make_double2(-col[0].x, -col[0].y);
  if (j == 0) {
    x.x += z.x;
    x.y += z.y;
  }
  for (int k = 0; k < j; ++k) {
    double2 y = make_double2(-col[k + 1].x, -col[k + 1].y);
    if (k + 1 == j) {
      y.x += z.x;
      y.y += z.y;
    }
    const double c = rot_c[(int64_t)k * n_shifts + s];
    const double2 sn = rot_s[(int64_t)k * n_shifts + s];
    // x_new (discarded) = c x + sn y ; y_new = -conj(sn) x + c y
    const double2 cs = make_double2(sn.x, -sn.y);
    const double2 t = zmul(cs, x);
    x = make_double2(c * y.x - t.x, c * y.y - t.y);
  }
  const double2 b = make_double2(-col[j + 1].x, -col[j + 1].y);
  const double ax = hypot(x.x, x.y);
  const double bx = hypot(b.x, b.y);
  double c;
  double2 sn;
  if (ax == 0.0) {
    c = 0.0;
    sn = make_double2(1.0, 0.0);
  } else {
    const double nrm = hypot(ax, bx);
    c = ax / nrm;
    // sn = (x/|x|) conj(b) / nrm
    const double2 u = make_double2(x.x / ax, x.y / ax);
    sn = zmul(u, make_double2(b.x / nrm, -b.y / nrm));
  }
  rot_c[(int64_t)j * n_shifts + s] = c;
  rot_s[(int64_t)j * n_shifts + s] = sn;
  const double r = rho[s] * hypot(sn.x, sn.y);
  rho[s] = r;
  resid[s] = c > 0.0 ? r / c : INFINITY;
}

__global__ void fill_kernel(double* __restrict__ a, double v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) a[i] = v;
}

// out[0] = max_i a[i] (NaN propagates as +inf)
__global__ void max_kernel(const double* __restrict__ a, int n, double* __restrict__ out) {
  __shared__ double part[256];
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double v = a[i];
    m = (v != v) ? INFINITY : (v > m ? v : m);
  }
  part[threadIdx.x] = m;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] = part[threadIdx.x] > part[threadIdx.x + w] ? part[threadIdx.x] : part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = part[0];
}

int64_t blocks_for(int64_t n, int t = 256) {
  int64_t b = (n + t - 1) / t;
  return b < 1 ? 1 : (b > 65536 ? 65536 : b);
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

  int apply(const double* v, double* out) {
    D_CALL(c3l_tilde_unpack(v, t_work, tm->k0, tm->k1, tm->ab, tm->dim, stream));
    D_CALL(c3l_plan_load(plan, 1, t_work, stream));
    D_CALL(c3l_plan_apply(plan, 2, jvp_diag, t_work, out_work, stream));
    D_CALL(c3l_tilde_pack(out_work, out, tm->src, tm->w, rd, stream));
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

}  // namespace

extern "C" {

int c3l_gmres(c3l_plan* plan, const c3l_tilde_maps* tm, const double* jvp_diag, double* t_work, double* out_work,
              const double* b, double* x, double* V, double* w, double* tmp, double* scratch, double* coef,
              double rtol, int32_t restart, int32_t cycles, int32_t* info, int32_t* iterations, double* true_rel,
              int32_t* jvp_calls, void* stream) {
  if (plan == nullptr || tm == nullptr || b == nullptr || x == nullptr || V == nullptr || w == nullptr ||
      tmp == nullptr || scratch == nullptr || coef == nullptr || restart <= 0 || cycles <= 0 || tm->real_dim <= 0 ||
      !(rtol > 0.0) || info == nullptr || iterations == nullptr || true_rel == nullptr)
    return dfail(C3L_EINVAL, "c3l_gmres: invalid argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t rd = tm->real_dim;
  Gmres g{plan, tm, jvp_diag, t_work, out_work, scratch, coef, stream, rd, jvp_calls};
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
  const double ptol = mbnrm * rtol;
  const double atol = rtol * bnrm;
  const int R = restart;
  std::vector<double> H((size_t)(R + 1) * R), cs(R), sn(R), S(R + 1), y(R), col(2 * (R + 1) + 1);
  auto h = [&](int i, int j) -> double& { return H[(size_t)i * R + j]; };
  for (int32_t cycle = 0; cycle < cycles; ++cycle) {
    if (cycle == 0) {
      D_CHECK(hipMemcpyAsync(w, b, sizeof(double) * rd, hipMemcpyDeviceToDevice, s));
    } else {
      D_CALL(g.apply(x, w));
      D_CALL(c3l_raxpby(1.0, b, -1.0, w, rd, stream));
      double rn = 0.0;
      D_CALL(g.norm(w, &rn));
      if (rn < atol) {
        *info = 0;
        break;
      }
    }
    D_CALL(g.precondition(w, V));
    double beta = 0.0;
    D_CALL(g.norm(V, &beta));
    if (beta == 0.0) {
      *info = 0;
      break;
    }
    D_CALL(c3l_rscale(V, V, 1.0 / beta, rd, stream));
    std::fill(H.begin(), H.end(), 0.0);
    std::fill(S.begin(), S.end(), 0.0);
    S[0] = beta;
    int used = 0;
    bool done = false;
    for (int j = 0; j < R; ++j) {
      D_CALL(g.apply(V + (int64_t)j * rd, tmp));
      D_CALL(g.precondition(tmp, w));
      double* c1 = coef;
      double* c2 = coef + (R + 1);
      double* sq = coef + 2 * (R + 1);
      D_CALL(c3l_rmultidot(V, rd, j + 1, w, rd, scratch, c1, stream));
      D_CALL(c3l_rmultiaxpy(V, rd, j + 1, c1, -1.0, w, rd, stream));
      D_CALL(c3l_rmultidot(V, rd, j + 1, w, rd, scratch, c2, stream));
      D_CALL(c3l_rmultiaxpy(V, rd, j + 1, c2, -1.0, w, rd, stream));
      D_CALL(c3l_rmultidot(w, rd, 1, w, rd, scratch, sq, stream));
      D_CHECK(hipMemcpyAsync(col.data(), coef, sizeof(double) * (2 * (R + 1) + 1), hipMemcpyDeviceToHost, s));
      D_CHECK(hipStreamSynchronize(s));
      for (int k = 0; k <= j; ++k) h(k, j) = col[k] + col[R + 1 + k];
      const double sqv = col[2 * (R + 1)];
      const double hn = std::sqrt(sqv > 0.0 ? sqv : 0.0);
      if (!std::isfinite(hn)) return dfail(C3L_EINVAL, "c3l_gmres: nonfinite Krylov vector");
      h(j + 1, j) = hn;
      if (hn > 0.0) D_CALL(c3l_rscale(w, V + (int64_t)(j + 1) * rd, 1.0 / hn, rd, stream));
      for (int k = 0; k < j; ++k) {
        const double a = h(k, j), bb = h(k + 1, j);
        h(k, j) = cs[k] * a + sn[k] * bb;
        h(k + 1, j) = -sn[k] * a + cs[k] * bb;
      }
      const double a = h(j, j), bb = h(j + 1, j);
      const double r = std::hypot(a, bb);
      cs[j] = r == 0.0 ? 1.0 : a / r;
      sn[j] = r == 0.0 ? 0.0 : bb / r;
      h(j, j) = r;
      h(j + 1, j) = 0.0;
      S[j + 1] = -sn[j] * S[j];
      S[j] = cs[j] * S[j];
      ++*iterations;
      used = j + 1;
      if (std::fabs(S[j + 1]) <= ptol || hn == 0.0) {
        done = true;
        break;
      }
    }
    for (int k = used - 1; k >= 0; --k) {
      double acc = S[k];
      for (int l = k + 1; l < used; ++l) acc -= h(k, l) * y[l];
      y[k] = acc / h(k, k);
    }
    D_CHECK(hipMemcpyAsync(coef, y.data(), sizeof(double) * used, hipMemcpyHostToDevice, s));
    D_CALL(c3l_rmultiaxpy(V, rd, used, coef, 1.0, x, rd, stream));
    D_CHECK(hipStreamSynchronize(s));  // y is a host stack buffer reused by the next cycle
    (void)done;
  }
  D_CALL(g.apply(x, w));
  D_CALL(c3l_raxpby(1.0, b, -1.0, w, rd, stream));
  double rn = 0.0;
  D_CALL(g.norm(w, &rn));
  *true_rel = rn / bnrm;
  if (*true_rel <= rtol * (1.0 + 1e-8)) *info = 0;
  return C3L_OK;
}

int c3l_arnoldi(c3l_plan* plan, const double* diag, double* V, double* H, int64_t dim, int32_t m_max,
                const double* shifts, int32_t n_shifts, double* rot_c, double* rot_s, double* rho, double* resid,
                double* scratch, double* coef, double* status, double tolerance, int32_t min_dim,
                double breakdown_tol, int32_t* m_out, int32_t* breakdown_out, double* worst_out, void* stream) {
  if (plan == nullptr || V == nullptr || H == nullptr || dim <= 0 || m_max <= 0 || n_shifts <= 0 || shifts == nullptr ||
      rot_c == nullptr || rot_s == nullptr || rho == nullptr || resid == nullptr || scratch == nullptr ||
      coef == nullptr || status == nullptr || m_out == nullptr || breakdown_out == nullptr || worst_out == nullptr)
    return dfail(C3L_EINVAL, "c3l_arnoldi: invalid argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  double2* Vz = reinterpret_cast<double2*>(V);
  double2* Hz = reinterpret_cast<double2*>(H);
  double2* cz = reinterpret_cast<double2*>(coef);
  double2* w = reinterpret_cast<double2*>(scratch);  // first dim complex entries of scratch hold w
  double* red = scratch + 2 * dim;                   // reduction partials after w
  hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(n_shifts)), dim3(256), 0, s, rho, 1.0, n_shifts);
  D_CHECK(hipGetLastError());
  double host[2] = {0.0, 0.0};
  int32_t m = 0;
  int32_t broke = 0;
  double worst = INFINITY;
  for (int32_t j = 0; j < m_max; ++j) {
    double* vj = V + (int64_t)2 * j * dim;
    D_CALL(c3l_plan_load(plan, 2, vj, stream));
    D_CALL(c3l_plan_apply(plan, 3, diag, vj, reinterpret_cast<double*>(w), stream));
    double* wd = reinterpret_cast<double*>(w);
    D_CALL(c3l_zmultidot(V, dim, j + 1, wd, dim, 1, red, coef, stream));
    D_CALL(c3l_zmultiaxpy(V, dim, j + 1, coef, -1.0, wd, dim, stream));
    D_CALL(c3l_zmultidot(V, dim, j + 1, wd, dim, 1, red, coef + 2 * (m_max + 1), stream));
    D_CALL(c3l_zmultiaxpy(V, dim, j + 1, coef + 2 * (m_max + 1), -1.0, wd, dim, stream));
    D_CALL(c3l_zmultidot(wd, dim, 1, wd, dim, 1, red, coef + 4 * (m_max + 1), stream));
    hipLaunchKernelGGL(finish_column_kernel, dim3(1), dim3(256), 0, s, cz, j, m_max, Hz, status);
    D_CHECK(hipGetLastError());
    hipLaunchKernelGGL(normalize_kernel, dim3(blocks_for(dim)), dim3(256), 0, s, w, Vz + (int64_t)(j + 1) * dim, status,
                       breakdown_tol, dim);
    D_CHECK(hipGetLastError());
    hipLaunchKernelGGL(shifted_givens_kernel, dim3(blocks_for(n_shifts)), dim3(256), 0, s, Hz, j, m_max,
                       reinterpret_cast<const double2*>(shifts), n_shifts, rot_c, reinterpret_cast<double2*>(rot_s), rho,
                       resid);
    D_CHECK(hipGetLastError());
    hipLaunchKernelGGL(max_kernel, dim3(1), dim3(256), 0, s, resid, n_shifts, status + 1);
    D_CHECK(hipGetLastError());
    D_CHECK(hipMemcpyAsync(host, status, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    D_CHECK(hipStreamSynchronize(s));
    m = j + 1;
    worst = host[1];
    if (!std::isfinite(host[0])) return dfail(C3L_EINVAL, "c3l_arnoldi: nonfinite Arnoldi recurrence");
    if (host[0] <= breakdown_tol) {
      broke = 1;
      break;
    }
    if (m >= min_dim && worst <= tolerance) break;
  }
  *m_out = m;
  *breakdown_out = broke;
  *worst_out = worst;
  return C3L_OK;
}

}  // extern "C"
