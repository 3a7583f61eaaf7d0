// One-reduction Gram-Schmidt step for the native Krylov drivers (library-internal).
//
// Classical Gram-Schmidt with delayed reorthogonalisation: the basis vector
// v_j enters the operator while it has been projected only once (call it q),
// and ONE reduction pass over the basis yields both its reorthogonalisation
// coefficients s = V^H q and the projection z = V^H w of w = A q; ONE update
// pass then writes the final v_j = (q - V s) / r and the once-projected next
// vector.  The Hessenberg column j-1 is completed at the same time
// (H[:j, j-1] += nu s, H[j, j-1] = nu r) and the new column j is corrected for
// A v_j = (w - A V s) / r = (w - V_{j+1} H s) / r.  Two passes over the basis
// per step instead of the four of two-pass CGS.

#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace c3l_ortho {

// Device buffers of one Krylov basis, all owned by the caller.  T is the arithmetic
// type (double or double2), S the storage type of the basis columns (T, or float for
// a real basis kept in FP32).
template <typename T, typename S = T>
struct Basis {
  S* V;            // columns 0..m (column j at V + j*ld)
  int64_t ld;      // >= dim
  int64_t dim;     // rows
  T* H;            // Hessenberg columns: column j at H + j*ldh, rows 0..j+1
  int64_t ldh;     // >= max columns + 1
  T* sink_v;       // sink . v_j per column (nullable)
  const T* sink;   // dotted unconjugated with every basis vector (nullable)
  T* partial;      // reduction partials: 2*(ldh+1)*chunks entries
  T* red;          // finished reductions: 2*(ldh+1) entries
  T* sc;           // step scalars, two alternating blocks of 8 (16 entries)
  double* status;  // status[0] = final H[j, j-1] when set (nullable)
  T* stage;        // when set: final column j-1 rows 0..j, then its h0 (j+2 entries; nullable)
  // When set, the update pass also scatters q_{j+1} into the padded input blocks of the
  // operator's plan: entry i goes to pad[m >> 1] (negated when m & 1) for each code m >= 0 of
  // pad_map[2 i], pad_map[2 i + 1], so the next operator application skips its own scatter.
  T* pad = nullptr;
  const int64_t* pad_map = nullptr;
};

// Scratch sizes in units of T (entries) for a basis of `dim` rows and at most
// `max_columns` columns (Hessenberg leading dimension max_columns + 1).
int64_t partial_entries(bool complex_basis, int64_t dim, int32_t max_columns);

// Optional HIP-event brackets around the two bandwidth passes of a step
// (pass 0: the reduction kernel, pass 1: the update kernel), with the
// algorithmic bytes of the pass.
struct Clock {
  virtual void begin(int pass, hipStream_t stream) = 0;
  virtual void end(int pass, double bytes, hipStream_t stream) = 0;
  virtual ~Clock() = default;
};

// Step j: with `w` (= A q_j, q_j in column j) finalise column j-1 and v_j and form
// the tentative column j and the once-projected q_{j+1} in column j+1; with
// w == nullptr only finalise column j-1 and v_j.  `fresh`: q_j is already final
// (first step), so no reorthogonalisation and no column j-1.  `breakdown`: a
// final H[j, j-1] at or below it zeroes v_j.
template <typename T, typename S>
int step(const Basis<T, S>& b, int j, bool fresh, const T* w, double breakdown, hipStream_t stream, Clock* clock);

// Per-instance pointers of a kernel argument: launches over several bases of equal
// shape index them with blockIdx.z.
constexpr int kMaxInstances = 4;
template <typename P>
struct Many {
  P p[kMaxInstances];
};

// Step j of `count` bases of equal dim, ld and ldh at once (one launch per kernel,
// blockIdx.z = basis); w[i] belongs to basis i and is null for all or for none.
template <typename T, typename S>
int step_many(const Basis<T, S>* b, int count, int j, bool fresh, const T* const* w, double breakdown,
              hipStream_t stream, Clock* clock);

}  // namespace c3l_ortho
