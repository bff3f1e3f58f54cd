/*
 * jja.h -- C ABI of libjja, the MI355X time-evolution engine for Josephson
 * junction arrays.
 *
 * The library replaces the reference's hot loop
 *     time_evolution_core(problem, th_store_mask, I_store_mask) -> (th_out, I_out)
 * (time_evolution.py:461-582) and the native routines it leans on: the per-step
 * numpy passes (Ic sin, stencil terms, noise), scipy's sparse A / A^T products,
 * and SuperLU's factorize-once / solve-every-step (time_evolution.py:504-506,
 * 560-569; scipy/sparse/linalg/_dsolve/linsolve.py:596). The Python package
 * pyjjasim_amd calls it through ctypes from its TimeEvolutionProblem.compute()
 * (the reference's time_evolution(), time_evolution.py:422-458, stays in Python).
 *
 * Conventions
 *   - every function returns 0 (JJA_OK) or a negative JJA_E* code; the message
 *     is available from jja_last_error() (thread-local);
 *   - plan creation copies host arrays; run-time pointers (state, inputs,
 *     outputs) are DEVICE pointers owned by the caller;
 *   - no function allocates device memory after jja_plan_create, except that
 *     jja_evolve grows one plan-owned descriptor buffer when it first sees more
 *     problems or columns;
 *   - a plan may be used from one host thread at a time.
 */
#ifndef JJA_H
#define JJA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JJA_OK 0
#define JJA_EINVAL (-1)       /* invalid argument or shape (Python: ValueError)   */
#define JJA_ENOMEM (-2)       /* host or device allocation failed                */
#define JJA_EFACTOR (-3)      /* face matrix not positive definite               */
#define JJA_EHIP (-4)         /* HIP runtime error; message has hipGetErrorString */
#define JJA_EUNSUPPORTED (-5) /* problem too large for this kernel variant       */

typedef struct jja_factor jja_factor;
typedef struct jja_plan jja_plan;

/* ------------------------------------------------------------------ factor */

/*
 * Cholesky K = L L^T of a symmetric positive definite matrix given in full
 * compressed-column form (both triangles), rows already in the desired order.
 * Replaces scipy.sparse.linalg.factorized(K) (time_evolution.py:505-506).
 * On success *nnz_out is the number of stored entries of L.
 */
int jja_cholesky(int32_t n, const int64_t* colptr, const int32_t* rowidx, const double* val,
                 jja_factor** out, int64_t* nnz_out);
/* Copy L out: colptr[n+1], rowidx[nnz], val[nnz]; rows ascending, diagonal first. */
int jja_factor_get(const jja_factor* f, int64_t* colptr, int32_t* rowidx, double* val);
void jja_factor_free(jja_factor* f);
const char* jja_factor_last_error(void);

/* ------------------------------------------------------------------- plan */

/*
 * One circuit: topology, per-junction coefficients and the stage schedule of
 * the triangular solves. All pointers are HOST arrays, copied by plan creation.
 *
 * Face positions: the solve works on faces renumbered by face_pos. Rows of the
 * factor are grouped into dense blocks; blocks of one stage are independent.
 * Forward solve of stage s, for each row r of the stage:
 *     t_r = b_r - sum_e lo_val[e] * z[lo_col[e]]           (e in lo_ptr[r]..)
 *     z_r = sum_{c <= a} inv(a, c) * t_{block+c}             (a = r - block)
 * Backward solve, stages in reverse:
 *     t_r = z_r - sum_e up_val[e] * x[up_col[e]]
 *     x_r = sum_{c >= a} inv(c, a) * t_{block+c}
 * where inv is the inverse of the block's diagonal part of L. Packed storage of
 * a block of size B at offset o:  inv_cm[o + c*B - c*(c-1)/2 + (a - c)] = inv(a, c)
 * (column-major lower), inv_rm[o + c*(c+1)/2 + a] = inv(c, a) (row-major lower).
 */
typedef struct jja_problem_desc {
  int32_t n_junctions;
  int32_t n_faces;
  int32_t stencil_width; /* 3, 4 or 5 (time_evolution.py:401-408) */
  int32_t n_stages;
  /* cycle matrix A by face, ascending junction: code = 2*j + (A[f,j] < 0) */
  const int32_t* face_ptr;  /* [n_faces+1] */
  const int32_t* face_code; /* [face_ptr[n_faces]] */
  /* A^T by junction, two slots, ascending face: code = 2*f + (A[f,j] < 0), -1 empty */
  const int32_t* junc_code; /* [2*n_junctions] */
  const double* Ic;         /* [Nj] critical current */
  const double* Rv;         /* [Nj] 1/(dt R) */
  const double* coef;       /* [stencil_width][Nj] c_k = S1[k] Rv + S2[k] Cv (time_evolution.py:477-499) */
  const int32_t* face_pos;  /* [Nf] */
  const int32_t* stage_ptr; /* [n_stages+1] */
  const int64_t* lo_ptr;    /* [Nf+1] */
  const int32_t* lo_col;
  const double* lo_val;
  const int64_t* up_ptr;    /* [Nf+1] */
  const int32_t* up_col;
  const double* up_val;
  const int32_t* row_block; /* [Nf] first row of the row's block */
  const int32_t* row_bsize; /* [Nf] block size */
  const int64_t* row_boff;  /* [Nf] block offset into inv_cm / inv_rm */
  int64_t n_inv;
  const double* inv_cm;
  const double* inv_rm;
} jja_problem_desc;

/* Create a plan for n_problems circuits on the current HIP device. */
int jja_plan_create(int32_t n_problems, const jja_problem_desc* problems, jja_plan** out);
int jja_plan_destroy(jja_plan* plan);
/* Dynamic LDS bytes one column of problem p needs (0 if it cannot run). */
int64_t jja_plan_lds_bytes(const jja_plan* plan, int32_t p);

/* ---------------------------------------------------------------- evolve */

/*
 * An input term (time_evolution.py:119-130, 342-356), evaluated at element i
 * (junction or face), column w, absolute step t:
 *     value = base[i*s_i + w*s_w + (t - t0)*s_t] * (amp ? amp[w*a_w + (t - t0)*a_t] : 1)
 * zero != 0 marks the reference's "skip this term" flag (np.allclose(x, 0) for an
 * input that is not time dependent, time_evolution.py:509-519); a zero current
 * source is still added to the stored current, as the reference does.
 */
typedef struct jja_source {
  const double* base;
  int64_t s_i, s_w, s_t;
  const double* amp;
  int64_t a_w, a_t;
  int64_t t0;
  int32_t zero;
  int32_t pad_;
} jja_source;

/* Run-time state of one problem (all DEVICE pointers, caller-owned). */
typedef struct jja_problem_run {
  int32_t W;             /* number of columns (bias points) */
  int32_t column_offset; /* noise counter of column 0 */
  double dt;
  double* theta;         /* [stencil_width-1][W][Nj]; theta at step m in slot m mod (sw-1) */
  double* theta_s;       /* [W][Nj] voltage-source phase; NULL when Vs.zero */
  jja_source Is, Vs, f, T;
  const double* noise;   /* optional [n_steps][W][Nj] standard normals (index t - noise_t0) */
  int64_t noise_t0;
  uint64_t seed;         /* Philox key when noise == NULL */
  const int32_t* th_slot; /* [*] output slot of theta per absolute step - slot_t0, or -1 */
  const int32_t* I_slot;
  int64_t slot_t0;
  double* th_out;        /* [slots][W][Nj] */
  double* I_out;         /* [slots][W][Nj] */
} jja_problem_run;

/*
 * Advance every column of every problem by n_steps steps, starting at absolute
 * step step0, on `stream` (a hipStream_t; NULL = default stream). One
 * workgroup integrates one column for all n_steps steps (persistent launch).
 * block_size: threads per workgroup (64..1024, multiple of 64; 0 = default).
 */
int jja_evolve(jja_plan* plan, const jja_problem_run* runs, int32_t n_steps, int64_t step0,
               int32_t block_size, void* stream);

/*
 * Vortex configuration n[f] = -sum_j A[f,j] rint(theta_j / (2 pi))
 * (time_evolution.py:734-755) for `cols` phase vectors of problem p:
 * theta [cols][Nj] -> n [cols][Nf] (int32).
 */
int jja_vortex(jja_plan* plan, int32_t p, const double* theta, int32_t cols, int32_t* n_out,
               void* stream);

const char* jja_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* JJA_H */
