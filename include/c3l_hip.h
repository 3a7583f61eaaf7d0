/*
 * c3l_hip.h -- C ABI of the MI355X (gfx950) coupled-cluster hot path of C3L.
 *
 * All vectors are device pointers borrowed for the duration of a
 * stream-ordered call; complex numbers are interleaved (re, im) doubles.
 * Plan handles own their device copies of index maps, coefficients and
 * atom workspaces and are released by c3l_plan_destroy.  Every entry point
 * returns 0 on success or a C3L_E* code; c3l_last_error() gives a
 * thread-local message.  Streams are passed as void* (a hipStream_t).
 *
 * Reference interfaces replaced (file:line in haoen2021/C3Lindblad):
 *   c3l_plan_*          ComponentStaticProblem.compile / component_static_residual /
 *                       component_static_linearization / component_static_linearized_jvp
 *                       (src/c3l/cc/component_tensor.py:1276,1822,1371,1552) and
 *                       ComponentEOMProblem.compile / component_eom_context /
 *                       component_eom_sigma (component_tensor.py:2012,2234,2288);
 *                       the JAX twins Engine.residual/jvp (cc/_static_engine.py:88-107)
 *                       and JaxFullS135ResponseOperator._device_action (cc/s135_jax.py:941-977).
 *   c3l_tilde_unpack/pack  TildeMap.unpack/pack (src/c3l/cc/symmetry.py:67-107);
 *                       JAX unpack/pack (cc/_static_engine.py:119-126).
 *   c3l_rmultidot / c3l_rmultiaxpy    GMRES CGS2 inside solve_component_ccsd_jax
 *                       (cc/_static_gmres.py:12-66).
 *   c3l_zmultidot / c3l_zmultiaxpy    Arnoldi CGS2 (src/c3l/green/krylov.py:361-370;
 *                       green/resident_arnoldi.py:82-114).
 *   c3l_shifted_triangular_green      per-shift reduced resolvent
 *                       (green/krylov.py:66-93,257-281; cc/component_jax.py:998-1009).
 */
#ifndef C3L_HIP_H
#define C3L_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  C3L_OK = 0,
  C3L_EINVAL = 1,   /* bad shape, index or argument */
  C3L_EHIP = 2,     /* HIP runtime error */
  C3L_ENOMEM = 3,   /* device allocation failed */
};

typedef struct c3l_plan c3l_plan;

/* Last error message of the calling thread ("" when none). */
const char* c3l_last_error(void);
/* Library version string. */
const char* c3l_version(void);

/* ---- contraction plans (residual, JVP, EOM sigma) ---------------------------------
 * A plan owns one complex workspace of `workspace` elements holding every atom
 * (padded tensor of n^rank entries at atom_offset[a]) plus the padded base blocks
 * added with c3l_plan_add_base.  `vectors` is the [n_vectors x n] complex table of
 * creator/annihilator vectors.  Constant creator atoms are filled at creation:
 * atom const_atom[k] <- vectors[const_vec[k]].
 */
int c3l_plan_create(int32_t n, int32_t n_vectors, const double* vectors, int64_t workspace,
                    int32_t n_const, const int32_t* const_atom, const int64_t* const_offset,
                    const int32_t* const_vec, c3l_plan** out);
int c3l_plan_destroy(c3l_plan* plan);

/* Register padded base block `block` of atom kind `kind` (0=T, 1=D, 2=X): padded
 * element e (of n^deg) equals sign[e] * v[src[e]] of the loaded complex vector v,
 * or 0 when src[e] < 0.  The block lives at workspace offset `offset`. */
int c3l_plan_add_base(c3l_plan* plan, int32_t kind, int64_t offset, int32_t deg, const int64_t* src,
                      const int8_t* sign);

/* Register one contraction level of `kind`: job j builds atom at out_offset[j]
 * (rank par_rank[j]-1) = sign[j] * sum_q vectors[vec[j]][q] * parent[.., q@slot[j], ..]
 * with the parent at par_offset[j] (rank par_rank[j]).  Levels run in the order added. */
int c3l_plan_add_level(c3l_plan* plan, int32_t kind, int32_t n_jobs, const int64_t* out_offset,
                       const int64_t* par_offset, const int32_t* par_rank, const int32_t* slot,
                       const int32_t* vec, const int32_t* sign);

/* Register one output block of program `program` (0 <= program < 9: residual, reference,
 * jvp, sigma, creation, observable, source_addition, source_removal, residual_low): `count`
 * coordinates of degree `deg` <= 5 (coord[count*deg] local indices < n <= 255, dest[count]
 * output positions) and `n_entries` wedge entries: entry t = coeff[t] * prod_{f in [fstart[t],
 * fstart[t+1])} ws[foffset[f] + flat_f], flat_f = sum_s coord[pos_s] n^s (first index fastest)
 * with fcode[f] = rank | pos0<<3 | pos1<<6 | pos2<<9 | pos3<<12 | pos4<<15 (rank <= 5).  All blocks of a
 * program are evaluated by one launch (blocks of a program must not read each other's
 * destinations); registering a block invalidates that program's device image. */
int c3l_plan_add_block(c3l_plan* plan, int32_t program, int32_t deg, int64_t count, const int32_t* coord,
                       const int64_t* dest, int32_t n_entries, const double* coeff, const int32_t* fstart,
                       int32_t n_factors, const int64_t* foffset, const int32_t* fcode);

/* A dense output block: every index tuple of a padded rank-deg tensor (n^deg coordinates, flat
 * index e = sum_s i_s n^s, first index fastest), coordinate e written at position base + e (the
 * creation tensors of a workspace program).  Entries and factors as in c3l_plan_add_block; no
 * coordinate or destination arrays are stored. */
int c3l_plan_add_dense_block(c3l_plan* plan, int32_t program, int32_t deg, int64_t base, int32_t n_entries,
                             const double* coeff, const int32_t* fstart, int32_t n_factors, const int64_t* foffset,
                             const int32_t* fcode);

/* Scatter the complex device vector v into the base blocks of `kind` and run that
 * kind's contraction levels (atoms of that kind are then current). */
int c3l_plan_load(c3l_plan* plan, int32_t kind, const double* v, void* stream);

/* out[dest] = diag[dest]*x[dest] + sum of the program's entries, for every block of
 * `program` (out positions not covered by a block are left untouched).  diag and x
 * may be NULL (no diagonal term).  out == NULL runs a workspace program (creation
 * tensors): destinations are workspace offsets, diag and x must be NULL. */
int c3l_plan_apply(c3l_plan* plan, int32_t program, const double* diag, const double* x, double* out,
                   void* stream);

/* Mixed precision (no counterpart in the reference, which is FP64 throughout): with
 * bits = 32, loads of atom kind `kind` build its atoms in a complex64 mirror of the
 * workspace and program `program` evaluates there with FP32 accumulation (its output,
 * including the diag*x term, stays complex128); FP64-produced atoms that the program
 * reads are narrowed into the mirror whenever they changed.  bits = 64 restores FP64.
 * Workspace (creation) programs always run in FP64. */
int c3l_plan_set_precision(c3l_plan* plan, int32_t kind, int32_t program, int32_t bits);

/* Matrix cores for FP32 programs (no counterpart in the reference): output blocks of rank 4
 * whose entries are products of two atoms splitting the four positions between them (or of a
 * rank-4 atom and a scalar atom) run as 16x4x4x4 output tiles on v_mfma_f32_4x4x1_16b_f32
 * (complex products by the 3M split) instead of the per-coordinate gather kernels.  mode 0:
 * never; 1 (the default): blocks with at least 128 entries; 2: every such block.  FP64
 * programs are not affected. */
int c3l_plan_set_matrix_cores(c3l_plan* plan, int32_t mode);

/* Matrix-core tiles of `program` at its last FP32 launch (0 before one, or when no block
 * qualifies). */
int c3l_plan_matrix_core_tiles(c3l_plan* plan, int32_t program, int32_t* tiles);

/* Launch-shape autotuner (no counterpart in the reference, whose XLA compiler picks its own
 * schedules, cc/component_jax.py:248-462): times `reps` launches of `program` on x into out
 * (diag and x as for c3l_plan_apply; out must not be NULL) for each candidate image shape, by
 * coordinate descent from the default heuristics over the matrix-core threshold of an FP32
 * program (128 / none / 32 / 1 entries), lanes per coordinate (4 / 2 / 1) and the workgroup
 * target of that split (512 / 256 / 1024 / 2048 per block), coordinates per thread (2 / 1) and
 * its workgroup target, and the XCD slicing of each block's tiles (on / off), and keeps the
 * fastest for every plan with this plan's digest, program and precision (images are rebuilt at
 * their next launch).  The matrix-core threshold and the lane split change the summation order
 * of a coordinate's entries and are taken only at a 10 % gain, the other knobs (results bit for
 * bit unchanged) at 2 %.  ms_default / ms_best: time per launch of the default and the chosen
 * shape; shape_out (6 ints, nullable): lanes_max (log2), lane-split workgroup target,
 * coordinates per thread, its workgroup target, XCD slicing, matrix-core threshold (INT32_MAX:
 * none).  Synchronises the device. */
int c3l_plan_autotune(c3l_plan* plan, int32_t program, const double* diag, const double* x, double* out,
                      int32_t reps, void* stream, double* ms_default, double* ms_best, int32_t* shape_out);

/* Measurement hooks (no counterpart in the reference, which times with perf_counter,
 * cc/component_jax.py:53-65): when enabled, every program launch of c3l_plan_apply is
 * bracketed by HIP events on its stream; c3l_plan_program_time waits for them and
 * returns the accumulated kernel time and launch count of one program (a launch that
 * evaluates k twin plans at once, inside c3l_arnoldi_batch, counts k).
 * c3l_plan_block_info returns the metadata of one registered output block. */
int c3l_plan_set_timing(c3l_plan* plan, int32_t enable);
int c3l_plan_program_time(c3l_plan* plan, int32_t program, double* ms_total, int64_t* launches);
int32_t c3l_plan_block_count(c3l_plan* plan);
int c3l_plan_block_info(c3l_plan* plan, int32_t index, int32_t* program, int32_t* deg, int64_t* count,
                        int32_t* n_entries, int32_t* n_factors);

/* ---- tilde real coordinates -------------------------------------------------------- */
/* t[i] = ab[4i..4i+1] * x[k0[i]] + ab[4i+2..4i+3] * x[k1[i]]  (complex * real) */
int c3l_tilde_unpack(const double* x, double* t, const int64_t* k0, const int64_t* k1, const double* ab,
                     int64_t dim, void* stream);
/* x[c] = Re(w[c] * t[src[c]]) */
int c3l_tilde_pack(const double* t, double* x, const int64_t* src, const double* w, int64_t real_dim,
                   void* stream);

/* x[c] = Re(w[c] * t[src[c]] * inv_diag[src[c]]): the preconditioned pack of a tilde-symmetric
 * complex vector in one pass (TildeMap.pack after _preconditioner, symmetry.py:94-107,
 * newton_krylov.py:158-192). */
int c3l_tilde_pack_precondition(const double* t, double* x, const int64_t* src, const double* w, const double* inv_diag,
                                int64_t real_dim, void* stream);

/* ---- Krylov building blocks --------------------------------------------------------- */
/* Real: out[k] = sum_i V[k*ld + i] * w[i], k < ncols (partial sums in scratch of
 * at least ncols*c3l_reduction_chunks(dim) doubles). */
int c3l_rmultidot(const double* V, int64_t ld, int32_t ncols, const double* w, int64_t dim, double* scratch,
                  double* out, void* stream);
/* Real: w[i] += scale * sum_k alpha[k] * V[k*ld + i] */
int c3l_rmultiaxpy(const double* V, int64_t ld, int32_t ncols, const double* alpha, double scale, double* w,
                   int64_t dim, void* stream);
/* Complex: out[k] = sum_i conj?(V[k*ld+i]) * w[i] (conj when conjugate != 0). */
int c3l_zmultidot(const double* V, int64_t ld, int32_t ncols, const double* w, int64_t dim, int32_t conjugate,
                  double* scratch, double* out, void* stream);
/* c3l_zmultidot without the scratch argument and without any allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Its partial sums live in a
 * library-owned device array of 2^20 doubles (8 MiB) per GPU, held until the process ends:
 * calls with ncols * ceil(dim / rows) > 524288 complex partials (rows = 512 for ncols <= 64,
 * 1024 for ncols <= 256, else 2048) are rejected with C3L_EINVAL.  Every call on one GPU shares
 * that array, so calls that overlap on one GPU (concurrent streams or threads) race. */
int c3l_zmultidot_noalloc(const double* V, int64_t ld, int32_t ncols, const double* w, int64_t dim,
                          int32_t conjugate, double* out, void* stream);
/* Complex: w[i] += scale * sum_k alpha[k] * V[k*ld + i] */
int c3l_zmultiaxpy(const double* V, int64_t ld, int32_t ncols, const double* alpha, double scale, double* w,
                   int64_t dim, void* stream);
/* Rapidity preconditioner in tilde coordinates: y = pack(unpack(x) * inv_diag), fused
 * (rapidity_diagonal/_preconditioner, src/c3l/cc/newton_krylov.py:137-192). x != y. */
int c3l_tilde_precondition(const double* x, double* y, const int64_t* k0, const int64_t* k1, const double* ab,
                           const int64_t* src, const double* w, const double* inv_diag, int64_t real_dim,
                           void* stream);
/* y = a * x (x may equal y) */
int c3l_rscale(const double* x, double* y, double a, int64_t dim, void* stream);
/* y = a * x + b * y */
int c3l_raxpby(double a, const double* x, double b, double* y, int64_t dim, void* stream);
/* Number of row chunks used by the multidot reductions for a given dimension. */
int64_t c3l_reduction_chunks(int64_t dim);

/* For every shift z_s: y = (z_s I - T)^{-1} r with T upper triangular (m x m,
 * column-major complex), green[s] = sum_k w[k] y[k], last[s] = sum_k q[k] y[k]. */
int c3l_shifted_triangular_green(const double* T, int32_t m, const double* r, const double* w, const double* q,
                                 const double* shifts, int32_t n_shifts, double* green, double* last,
                                 void* stream);

/* ---- native Krylov drivers ----------------------------------------------------------- */

/* Tilde coordinate maps and rapidity preconditioner of one static problem (device). */
typedef struct c3l_tilde_maps {
  int64_t dim;             /* complex amplitudes */
  int64_t real_dim;        /* independent real coordinates */
  const int64_t* k0;       /* unpack gathers [dim] */
  const int64_t* k1;
  const double* ab;        /* [dim][4] alpha re, im, beta re, im */
  const int64_t* src;      /* pack gather [real_dim] */
  const double* w;         /* pack weights, complex [real_dim] */
  const double* inv_diag;  /* 1 / (-i Lambda), complex [dim] */
} c3l_tilde_maps;

/* Newton correction J x = b in real tilde coordinates (replaces the scipy gmres call of
 * solve_component_ccsd, src/c3l/cc/component_solver.py:181-191, and the JAX cycle
 * get_cycle, cc/_static_gmres.py:12-66).  Left-preconditioned GMRES(restart), at most
 * `cycles` restarts, target max(atol, rtol ||b||) on the true residual, checked after every
 * restart; the inner test on the preconditioned residual follows scipy.sparse.linalg.gmres
 * (initial ||M b|| min(1, target/||b||), rescaled after each restart).  info = 0 when the
 * target is met, else `cycles`.  On return w holds the true residual b - J x.
 * The plan's T atoms must be loaded (linearisation point).  Workspace: t_work, out_work
 * complex [dim]; V [(restart+1) x ld] with ld = c3l_krylov_basis_ld(real_dim) (each column
 * starts on a 256-byte boundary; entries past real_dim are never read) of doubles
 * (basis_bits = 64) or floats (basis_bits = 32: basis vectors stored in FP32, every dot,
 * coefficient and update computed in FP64; meant for FP32 Jacobian products); w, tmp
 * [real_dim]; scratch and coef at least
 * the sizes c3l_krylov_workspace(0, real_dim, restart, 0, ...) returns.  Orthogonalisation:
 * classical Gram-Schmidt with delayed reorthogonalisation (one reduction and one update pass
 * per iteration); each iteration applies J to the once-projected basis vector, so a cycle
 * that stops on the inner test has applied J once more than it has iterations.
 * x is overwritten with the step. */
/* Leading dimension (entries) of the c3l_gmres basis for `rows` real coordinates: rows rounded up
 * to a multiple of 64, so that every column starts on a cache-line boundary in FP32 and FP64
 * (no counterpart in the reference, whose basis is a list of NumPy vectors). */
int64_t c3l_krylov_basis_ld(int64_t rows);

int c3l_gmres(c3l_plan* plan, const c3l_tilde_maps* tm, const double* jvp_diag, double* t_work, double* out_work,
              const double* b, double* x, void* V, int32_t basis_bits, double* w, double* tmp, double* scratch,
              double* coef, double rtol, double atol, int32_t restart, int32_t cycles, int32_t* info,
              int32_t* iterations, double* true_rel, int32_t* jvp_calls, void* stream);

/* c3l_gmres for `count` independent systems of equal real dimension at once (the Newton
 * corrections of a group of problem instances, each solved by the reference's
 * solve_component_ccsd, src/c3l/cc/component_solver.py:181-191, one after the other).  `plan` is
 * the union plan of the systems (their plans side by side, one launch per program for all) and
 * `tm` its union tilde maps (real_dim = count * rd).  b, x, w, tmp are the systems' vectors side
 * by side; V holds (restart+1) columns of count*rd entries, column j being every system's basis
 * column j; scratch and coef hold `count` consecutive blocks of the c3l_gmres sizes for rd.
 * Host arrays of `count`: rtol, atol (per system), active (0 = skip the system: its x is set to
 * zero and nothing else is reported), and the results info, iterations, true_rel, jvp_calls.
 * Each system follows c3l_gmres exactly (same tolerance control and stopping rules); the systems
 * advance in lockstep: one union Jacobian product per step serves all, the orthogonalisation
 * runs over the systems still iterating, and restart cycles end together. */
int c3l_gmres_group(c3l_plan* plan, const c3l_tilde_maps* tm, int32_t count, const double* jvp_diag, double* t_work,
                    double* out_work, const double* b, double* x, void* V, int32_t basis_bits, double* w, double* tmp,
                    double* scratch, double* coef, const double* rtol, const double* atol, const int32_t* active,
                    int32_t restart, int32_t cycles, int32_t* info, int32_t* iterations, double* true_rel,
                    int32_t* jvp_calls, void* stream);

/* Shift-invariant Krylov resolvent on the sigma program (replaces the per-frequency GMRES
 * loop of component_retarded_green, src/c3l/cc/component_response.py:180-245, and the Arnoldi
 * resolvent arnoldi_resolvent, src/c3l/green/krylov.py:284-417).  One Arnoldi basis (classical
 * Gram-Schmidt with delayed reorthogonalisation) serves every shift; each shift runs GMRES on
 * it with its sink projection
 * green[s] = sink . x_m(z_s) updated incrementally.  V [(m_max+1) x c3l_krylov_basis_ld(dim)]
 * complex (columns padded to a cache-line multiple, padding never read) with
 * V[0] = b/beta on entry; H complex, column j at H + j*(m_max+1); rot_c [n_shifts*m_max]
 * doubles, rot_s and pi [n_shifts*m_max] complex, gamma and green [n_shifts] complex, resid
 * [n_shifts] doubles (relative GMRES residual); scratch and coef at least the sizes
 * c3l_krylov_workspace(1, dim, m_max, n_shifts, ...) returns; status >= 4 doubles.  Stops at
 * the first m >= min_dim with every shift's residual <= tolerance, at a breakdown, or at m_max.
 * The shifts absorb completed columns eight at a time; the stop test after each batch finds
 * the exact first column that meets the criterion and freezes the shift state there.  The
 * host reads the stop flag asynchronously (at most one batch late), so up to 16 further basis
 * vectors may be computed (and ignored) after the stop.  criterion = 0 tests the GMRES
 * residual (the per-frequency GMRES of component_retarded_green); criterion = 1 tests the
 * Galerkin residual |h_{m+1,m} e_m^T (z - H_m)^{-1} beta e_1| / beta (the stop rule of
 * arnoldi_resolvent, green/krylov.py:257-281,380-390), and resid then holds it. */
int c3l_arnoldi(c3l_plan* plan, const double* diag, double* V, double* H, int64_t dim, int32_t m_max,
                const double* shifts, int32_t n_shifts, const double* sink, double beta, double* rot_c,
                double* rot_s, double* pi, double* gamma, double* green, double* resid, double* scratch,
                double* coef, double* status, double tolerance, int32_t min_dim, double breakdown_tol,
                int32_t criterion, int32_t* m_out, int32_t* breakdown_out, double* worst_out, void* stream);

/* c3l_arnoldi for `count` (1..4) right-hand sides at once, in lockstep: basis i runs on
 * plans[i], and every sigma and orthogonalisation kernel of a step is one launch over all
 * bases that have not stopped (the addition and removal resolvents of a charge block,
 * src/c3l/cc/component_response.py:180-245, each run as its own c3l_arnoldi in the
 * reference's per-channel loop).  The plans must be twins: uploaded from the same lowered
 * plan (equal structure, checked), each with its own workspace and the same T atoms loaded;
 * diag, shifts and the scalars are shared.  Every device buffer of c3l_arnoldi holds `count`
 * consecutive per-basis blocks of the single-basis size (V, H, sink, rot_c, rot_s, pi, gamma,
 * green, resid, scratch, coef, status); beta, m_out, breakdown_out and worst_out are host
 * arrays of `count`.  Each basis stops on its own criterion and is dropped from the launches
 * after the check that sees its stop. */
int c3l_arnoldi_batch(c3l_plan* const* plans, int32_t count, const double* diag, double* V, double* H, int64_t dim,
                      int32_t m_max, const double* shifts, int32_t n_shifts, const double* sink, const double* beta,
                      double* rot_c, double* rot_s, double* pi, double* gamma, double* green, double* resid,
                      double* scratch, double* coef, double* status, double tolerance, int32_t min_dim,
                      double breakdown_tol, int32_t* m_out, int32_t* breakdown_out, double* worst_out, void* stream);

/* c3l_arnoldi_batch in three calls, so that one host thread can drive the resolvents of
 * several charge blocks, each on its own stream (the reference runs them one after the
 * other, component_response.py:180-245).  c3l_arnoldi_begin validates, initialises the shift
 * state on `stream` and returns the run; c3l_arnoldi_drive advances every given run one
 * Krylov step at a time in turn until each has stopped; c3l_arnoldi_end waits for the run's
 * stream, writes m_out / breakdown_out / worst_out (count entries) and frees the run (also
 * after an error).  Stop flags are read asynchronously, at most one shift batch late, so a
 * stopped basis may compute up to 16 further (ignored) basis vectors.  The buffers must
 * stay valid until c3l_arnoldi_end. */
typedef struct c3l_arnoldi_run c3l_arnoldi_run;
int c3l_arnoldi_begin(c3l_plan* const* plans, int32_t count, const double* diag, double* V, double* H, int64_t dim,
                      int32_t m_max, const double* shifts, int32_t n_shifts, const double* sink, const double* beta,
                      double* rot_c, double* rot_s, double* pi, double* gamma, double* green, double* resid,
                      double* scratch, double* coef, double* status, double tolerance, int32_t min_dim,
                      double breakdown_tol, void* stream, c3l_arnoldi_run** out);
int c3l_arnoldi_drive(c3l_arnoldi_run* const* runs, int32_t n_runs);
/* c3l_arnoldi_begin for a group of problem instances of equal dimension whose sigma plans are
 * side by side in one union plan (the resolvents of component_retarded_green,
 * src/c3l/cc/component_response.py:180-245, of every instance in the group).  n_slots (1..4)
 * twin copies of the union plan each serve per_slot bases (n_slots * per_slot <= 64): basis
 * i = s * per_slot + p is instance p's resolvent on slot s (e.g. slot 0 addition, slot 1
 * removal).  V holds n_slots blocks of (m_max+1) columns of per_slot*dim complex entries (column
 * j of slot s: every instance's basis column j side by side, V[0] of basis i = its b/beta); W is
 * the sigma output, n_slots blocks of per_slot*dim; diag the union diagonal (per_slot*dim).  H,
 * sink (dim each), beta and every per-basis workspace of c3l_arnoldi_batch hold n_slots*per_slot
 * consecutive blocks in basis order.  One sigma launch per step serves every slot that still
 * has a live basis; each basis stops on its own criterion.  Driven and ended like
 * c3l_arnoldi_begin's runs (c3l_arnoldi_drive, c3l_arnoldi_end with count = n_slots*per_slot). */
int c3l_arnoldi_group_begin(c3l_plan* const* slot_plans, int32_t n_slots, int32_t per_slot, const double* diag,
                            double* V, double* W, double* H, int64_t dim, int32_t m_max, const double* shifts,
                            int32_t n_shifts, const double* sink, const double* beta, double* rot_c, double* rot_s,
                            double* pi, double* gamma, double* green, double* resid, double* scratch, double* coef,
                            double* status, double tolerance, int32_t min_dim, double breakdown_tol, void* stream,
                            c3l_arnoldi_run** out);
int c3l_arnoldi_end(c3l_arnoldi_run* run, int32_t* m_out, int32_t* breakdown_out, double* worst_out);

/* Workspace sizes (in doubles) of c3l_arnoldi (complex_basis = 1, dim complex rows, m_max =
 * max_columns, n_shifts shifts) and c3l_gmres (complex_basis = 0, real_dim rows, restart =
 * max_columns; n_shifts ignored).  No counterpart in the reference. */
int c3l_krylov_workspace(int32_t complex_basis, int64_t dim, int32_t max_columns, int32_t n_shifts,
                         int64_t* scratch_doubles, int64_t* coef_doubles);

/* Measurement hooks of the drivers (no counterpart in the reference): when enabled, the two
 * bandwidth passes of every orthogonalisation step inside c3l_arnoldi / c3l_gmres (the
 * reduction kernel and the update kernel) are bracketed by HIP events on their stream.
 * c3l_krylov_timing waits for them and returns, for one family (0 Arnoldi reduction,
 * 1 Arnoldi update, 2 GMRES reduction, 3 GMRES update), the accumulated milliseconds, the
 * number of launches and their algorithmic bytes (basis columns and vectors read and
 * written). */
int c3l_krylov_set_timing(int32_t enable);
int c3l_krylov_timing(int32_t family, double* ms_total, int64_t* launches, double* bytes);

/* ---- learned warm start (MLP) ---------------------------------------------------------
 * No counterpart in the reference: a small MLP (bf16 weights, FP32 accumulation on the
 * matrix cores) maps model parameters to compressed initial amplitudes.  Matrices are
 * strided views: A(m, k) = A[m*a_rs + k*a_cs], B(k, n) = B[k*b_rs + n*b_cs]; C is
 * row-major with leading dimension ldc.  bf16 values are passed as uint16_t bit patterns. */
/* C = alpha * A B + beta * C (C is FP32) */
int c3l_bf16_gemm(const uint16_t* A, int64_t a_rs, int64_t a_cs, const uint16_t* B, int64_t b_rs, int64_t b_cs,
                  float* C, int64_t ldc, int32_t M, int32_t N, int32_t K, float alpha, float beta, void* stream);
/* z[m, n] += bias[n]; h = bf16(act ? GELU(z) : z) when h != NULL (GELU in its tanh form) */
int c3l_mlp_bias_act(float* z, const float* bias, uint16_t* h, int32_t M, int32_t N, int32_t act, void* stream);
/* dz = dh * GELU'(z), also as bf16 */
int c3l_mlp_gelu_backward(const float* z, const float* dh, float* dz, uint16_t* dz16, int64_t total, void* stream);
/* db[n] = sum_m dz[m, n] */
int c3l_mlp_column_sum(const float* dz, float* db, int32_t M, int32_t N, void* stream);
/* dy = scale * (y - target) (FP32 and bf16); loss_parts[b] = partial sums of (y - target)^2
 * over n_parts workgroups */
int c3l_mlp_mse(const float* y, const float* target, float* dy, uint16_t* dy16, float* loss_parts, int64_t total,
                float scale, int32_t n_parts, void* stream);
/* Adam step `step` (>= 1) on FP32 master weights w with moments m, v; refreshes the bf16
 * copy w16 when non-NULL */
int c3l_mlp_adam(float* w, float* m, float* v, const float* g, uint16_t* w16, int64_t n, float lr, float b1, float b2,
                 float eps, int32_t step, void* stream);
/* b = bf16(a), round to nearest even */
int c3l_to_bf16(const float* a, uint16_t* b, int64_t n, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* C3L_HIP_H */
