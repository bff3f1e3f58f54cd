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
 *   - no function allocates device or host memory after jja_plan_create: the
 *     plan's launch descriptors, work lists and barrier words are sized there
 *     (jja_problem_desc.max_columns);
 *   - a plan may be used from one host thread at a time, and its launches run
 *     on one stream at a time: jja_evolve or jja_adjoint on a stream other than
 *     the one of the plan's previous launch, while that launch has not
 *     finished, is refused with JJA_EBUSY (the plan's launch descriptors are
 *     shared by its launches). Different plans run concurrently on different
 *     streams.
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
#define JJA_EBUSY (-6)        /* plan still running a launch on another stream   */

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
 * One circuit: per-junction coefficients, cycle-matrix layouts and the stage
 * schedule of the triangular solves, laid out for workgroups of `block_size`
 * threads. All pointers are HOST arrays, copied by plan creation.
 *
 * Faces are renumbered to solve positions 0..Nf-1 (face_of_pos). Work that a
 * workgroup spreads over its threads is a list of slices (jja_slice):
 *   - G = flags & 0xff consecutive threads share one row (G a power of two
 *     <= 64); thread tid works on row slice_rows[rowbase + tid/G] (-1: none) and
 *     takes that row's entries tid%G, tid%G + G, ...;
 *   - every address is known without reading an index first.
 * Face slices (workgroup-wide): code `it` (it < n_iter <= 8) of thread tid at
 *     base + it*block_size + tid; the row of tid at slice_rows[rowbase + tid].
 *     Face codes pad with -1.
 * Solve records (one wavefront of 64 lanes each): every record holds 8 entries
 *     per lane, zero-padded. Value `it` of lane l at base + (it/2)*128 + 2*l +
 *     it%2 (16-byte loads); index `it` of lane l at ibase + 8*l + it, except
 *     in a block-phase record (flag BLOCK) with 4-byte indices, whose lane
 *     reads consecutive scratch rows: lane l has one word at ent_off slot
 *     ibase + l, base | len << 28, and entry `it` is base + it*G for it < len,
 *     else 0; the row
 *     of lane l at slice_rows[rowbase + l/G]. An index addresses the column's
 *     face store in doubles: z at its position, the stage scratch t at
 *     n_faces + stage row. Wave w of the workgroup runs records solve_wave_ptr[w] ..
 *     solve_wave_ptr[w+1]-5: a multiple of three (a record with G 1 and no rows
 *     is a no-op), followed by 4 records that it loads ahead but never executes.
 *     Every wave's program has the same number of BARRIER records.
 * A row may span several consecutive slices; only the last one (flag FINISH)
 * reduces over the G lanes and writes the row. Row slots are filled from the
 * front: flags >> 16 is the number of slots holding a row, and threads of the
 * other slots load nothing.
 *
 * Face phase (G = 1, entries = junction codes 2*j + (A[f,j] < 0), ascending j):
 *     b_p = sum over codes of +-u_j,  u = x/c0 - theta_s.
 * The rows of the factor L of the reordered face matrix form dense blocks;
 * blocks of one stage are independent. The solve program lists the slices of
 * the forward then the backward solve in execution order; for each stage:
 *     off-block (flag BLOCK clear): t[row - r0] = z[row] - sum val * z[idx]
 *     block     (flag BLOCK set):   z[row]      = sum val * t[idx]
 * where the block entries are rows (forward) or columns (backward) of the
 * inverse of the row's dense diagonal block and idx is a stage-local row.
 * Flag BARRIER ends a phase: the workgroup synchronises after the slice (all
 * parts of a split column: a column barrier). Flag WG_BARRIER ends a phase of
 * a split column whose rows are read and written by one workgroup only: that
 * workgroup synchronises. Every wave meets the same barriers in the same order.
 */
#define JJA_SLICE_BLOCK 0x100
#define JJA_SLICE_FINISH 0x200
#define JJA_SLICE_BARRIER 0x400
#define JJA_SLICE_WG_BARRIER 0x800
/* an off-block record of a global store (4-byte indices) whose lanes each span
 * fewer than 2^16 positions: slot ibase + l holds lane l's base, slots ibase +
 * 64 + 4l + q hold its entries 2q and 2q+1 as 16-bit offsets from the base */
#define JJA_SLICE_NARROW 0x1000

/* most sine harmonics of a current-phase relation on the device */
#define JJA_CPR_MAX_TERMS 8
/* sine modes one wave of the lattice solve accumulates at once (even) */
#define JJA_LATTICE_CHUNK 14
/* the lattice solve cuts each mode's tridiagonal recurrence into
 * min(R, JJA_LATTICE_TRI_THREADS * parts / C) blocks */
#define JJA_LATTICE_TRI_THREADS 1024

typedef struct jja_slice {
  int64_t base;    /* first value (face slices: first code) */
  int64_t ibase;   /* first index (solve slices) */
  int32_t rowbase; /* first row slot in slice_rows */
  int32_t r0;      /* first position of the slice's stage */
  int32_t n_iter;  /* entries per thread, 0..8 */
  int32_t flags;   /* G | JJA_SLICE_* | rows << 16 */
} jja_slice;

typedef struct jja_problem_desc {
  int32_t n_junctions;
  int32_t n_faces;
  int32_t stencil_width;    /* 3, 4 or 5 (time_evolution.py:401-408) */
  int32_t block_size;       /* threads per column the layouts are built for (parts workgroups) */
  const double* Ic;         /* [Nj] critical current */
  const double* Rv;         /* [Nj] 1/(dt R) */
  const double* coef;       /* [stencil_width][Nj] c_k = S1[k] Rv + S2[k] Cv (time_evolution.py:477-499) */
  /* A^T by junction: two slots, ascending face index, code = 2*position + (A[f,j] < 0), -1 empty */
  const int32_t* junc_code; /* [2*Nj] */
  const int32_t* face_of_pos; /* [Nf] */
  int32_t n_face_slices;
  int32_t n_solve_slices;        /* all records of all waves, lookahead records included */
  const jja_slice* face_slices;  /* [n_face_slices] */
  const jja_slice* solve_slices; /* [n_solve_slices], per-wave programs in execution order */
  int64_t n_face_entries;
  const int32_t* face_code;      /* [n_face_entries] */
  int64_t n_slice_rows;
  const int32_t* slice_rows;     /* [n_slice_rows] */
  int64_t n_entries;
  const double* ent_val;         /* [n_entries] */
  int64_t n_index;
  const void* ent_off;           /* [n_index] face-store indices (doubles), index_bytes wide */
  int32_t max_stage_rows;        /* size of the stage scratch t */
  /* the first stage's rows have no off-block entries, so its scratch equals b:
   * the face phase writes b[p] also to t[p] for positions p < face_copy_rows,
   * and the program has no records for that phase */
  int32_t face_copy_rows;
  const int32_t* solve_wave_ptr; /* [block_size/64 + 1] first record of each wave's program */
  /* 2: the face vector and stage scratch live in the workgroup's LDS (indices are
   *    uint16 absolute LDS addresses in doubles);
   * 4: they live in a per-column global array (jja_problem_run.zbuf, uint32
   *    indices; each lane's 8 indices are two 16-byte loads), for circuits whose
   *    face vector does not fit in LDS. */
  int32_t index_bytes;
  /* workgroups per column: 1, or 2..32 (a power of two) with index_bytes 4. A
   * column of a split problem is worked by `parts` workgroups of
   * block_size/parts threads that meet at column barriers in global memory.
   * Every problem of a plan has the same block_size/parts; part counts may
   * differ, but once any problem is split, every problem of the plan needs
   * index_bytes 4, and jja_evolve / jja_adjoint launch cooperatively (all
   * workgroups of the launch resident at once, or JJA_EUNSUPPORTED). */
  int32_t parts;
  /* 0, or 1: split columns are worked two at a time (columns 2q and 2q+1 of a
   * problem share one set of parts workgroups, which stream the solve records
   * once for both and meet at one column barrier; the last column is alone when
   * W is odd). Needs parts >= 2 and block_size/parts <= 512; the same value for
   * every problem of a plan. Results are those of unpaired columns, bit for bit. */
  int32_t column_pairs;
  /* the most columns (W) a run of this problem will have, 1 .. 2^20: plan
   * creation sizes the launch descriptors, work lists and barrier words for it,
   * so that jja_evolve and jja_adjoint allocate nothing; a run with more
   * columns is refused (JJA_EINVAL) */
  int32_t max_columns;
  /* current-phase relation: 0 terms for the default Ic sin(theta)
   * (static_problem.py:77-85); n terms for Ic * sum_{k<n} cpr_coef[k] sin((k+1) theta),
   * the device form of a sine-series relation (jja_adjoint: default only) */
  int32_t cpr_terms;
  int32_t cpr_pad_;
  double cpr_coef[JJA_CPR_MAX_TERMS];
  /* Lattice solve: 0 rows, or the face system of a rectangular lattice of
   * R x C faces at positions r*C + c whose face matrix is the Kronecker sum
   * K = g (T_R (x) I_C) + g_c (I_R (x) T_C), T_n = tridiag(-1, 2, -1). The
   * step then solves K by a sine transform (DST-I) of every row, one
   * tridiagonal solve along r per sine mode and the inverse transform, in
   * place of the staged factor program (which is then empty: n_solve_slices 0,
   * face_copy_rows 0). The reference factorises K with SuperLU for every
   * circuit (time_evolution.py:505-506); this is an exact direct solve of the
   * same K. Needs column_pairs 0, R >= 2, C >= 2, R*C = n_faces and
   * max_stage_rows >= (n_faces & 1) + max(2*R*ES, 4*B*C) with H = (C+1)/2,
   * ES = H + (H & 1) and B = max(1, min(R, JJA_LATTICE_TRI_THREADS * parts / C))
   * blocks per tridiagonal solve. */
  int32_t lattice_rows;        /* R (0: no lattice solve) */
  int32_t lattice_cols;        /* C */
  /* [ES][CP], CP = C rounded up to a multiple of JJA_LATTICE_CHUNK:
   * sqrt(2/(C+1)) sin(pi (i+1)(j+1)/(C+1)) for i < H and j < C, else 0 */
  const double* lattice_sine;
  /* [R][CP]: 1 / pivot of sine mode k's tridiagonal factor g T_R + g_c mu_k I
   * at row r; the solve is d_r = (y_r + g d_{r-1}) inv[r][k], then
   * x_r = d_r - ((-g) inv[r][k]) x_{r+1} */
  const double* lattice_inv;
  double lattice_g;            /* g */
} jja_problem_desc;

/*
 * Check every size, index and offset of n_problems descriptors without touching
 * a device (jja_plan_create runs the same checks before it uploads anything).
 */
int jja_plan_validate(int32_t n_problems, const jja_problem_desc* problems);
/* Create a plan for n_problems circuits on the current HIP device. */
int jja_plan_create(int32_t n_problems, const jja_problem_desc* problems, jja_plan** out);
int jja_plan_destroy(jja_plan* plan);
/*
 * Status of the plan's split launches so far (waits for the device): JJA_OK, or
 * JJA_EHIP if a column barrier of a split launch gave up waiting (its results
 * are invalid). No counterpart in the reference: a split column is this
 * library's own way to spread one large circuit over many compute units.
 */
int jja_plan_status(jja_plan* plan);
/*
 * Polls a column barrier of the plan's split launches makes before it gives up
 * (default 2^21, each poll sleeping briefly: far longer than any phase of a
 * step). A small value forces the give-up path, which jja_plan_status then
 * reports; every workgroup still runs to the end of its launch. No
 * counterpart in the reference.
 */
int jja_plan_set_wait_limit(jja_plan* plan, uint32_t polls);
/*
 * Workgroups of the plan's split launches (since the previous call, which
 * resets the count; waits for the device) whose column had every part on one
 * XCD: their column barriers skip the L2 write-back. A
 * diagnostic of the launch layout; returns the count or a negative JJA_E*
 * code. No counterpart in the reference.
 */
int64_t jja_plan_one_xcd_parts(jja_plan* plan);
/* Dynamic LDS bytes one column of problem p needs (0 if it cannot run). */
int64_t jja_plan_lds_bytes(const jja_plan* plan, int32_t p);
/* Doubles per column of the global face store problem p needs (0: it uses LDS). */
int64_t jja_plan_zbuf_doubles(const jja_plan* plan, int32_t p);

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
  double* u;             /* [W][Nj] scratch: x/c0 - theta_s of the step being solved, for a
                            problem with the LDS face store (index_bytes 2) in a plan without
                            split columns; may be NULL otherwise (a global store's face phase
                            forms x/c0 - theta_s from the history slot) */
  jja_source Is, Vs, f, T;
  /* noise amplitude sqrt(2 T Rv) when T is constant in time (sigma.base != NULL);
   * otherwise the kernel forms it from T */
  jja_source sigma;
  const double* noise;   /* optional [n_steps][W][Nj] standard normals (index t - noise_t0) */
  int64_t noise_t0;
  uint64_t seed;         /* Philox key when noise == NULL */
  const int32_t* th_slot; /* [*] output slot of theta per absolute step - slot_t0, or -1 */
  const int32_t* I_slot;
  int64_t slot_t0;
  double* th_out;        /* [slots][W][Nj] */
  double* I_out;         /* [slots][W][Nj] */
  const double* Ic;      /* [Nj] critical currents of this run, or NULL: the plan's */
  double* zbuf;          /* [W][n_faces + max_stage_rows] face store when the problem was
                            planned with index_bytes 4 (see jja_plan_zbuf_doubles), else NULL */
} jja_problem_run;

/*
 * Advance every column of every problem by n_steps steps, starting at absolute
 * step step0, on `stream` (a hipStream_t; NULL = default stream). One
 * workgroup of block_size/parts threads (parts workgroups, for a split
 * problem) integrates one column for all n_steps steps, largest problems
 * first. A plan with any split problem launches cooperatively; check
 * jja_plan_status after its runs. Steps step0 .. step0 + n_steps - 1 must lie
 * within what every run's sources, noise and slot maps cover: those arrays
 * carry no length here, so the caller keeps launches inside the horizon.
 */
int jja_evolve(jja_plan* plan, const jja_problem_run* runs, int32_t n_steps, int64_t step0,
               void* stream);

/* --------------------------------------------------------------- adjoint */

/*
 * Reverse-mode sweep of the stencil-3 step (DefaultCPR) for the gradients of a
 * loss of the stored phases. The step is theta_k = M x_k + (terms in f, theta_s)
 * with M = D A^T K^-1 A D - D, D = diag(1/c0), symmetric; x_k depends on
 * theta_{k-1}, theta_{k-2} through a_k = 2 Ic cos(arg_k) + c1 and
 * b_k = c2 - Ic cos(arg_k), arg_k = 2 theta_{k-1} - theta_{k-2}. With lambda_k
 * the adjoint of theta_k and mu_k = M lambda_k (the same face phase, solve and
 * junction gather as the forward step, applied to lambda):
 *     lambda_{k-1} += a_k mu_k,   lambda_{k-2} += b_k mu_k,
 *     dL/dIc += sin(arg_k) mu_k,  dL/dIs_k = -mu_k,
 *     dL/df_k = -2 pi K^-1 A D lambda_k  (the solve's face vector).
 * All pointers are DEVICE pointers; arrays are [W][N] per problem.
 */
typedef struct jja_adjoint_run {
  int32_t W;
  int32_t prime;          /* 1: first call of a sweep (see jja_adjoint) */
  const double* Ic;       /* [Nj] or NULL: the plan's */
  const double* hist;     /* theta_m at hist[(m - hist_m0) * W * Nj + w * Nj + j] for
                             m in [k_lo - 2, k_hi - 1] */
  int64_t hist_m0;
  double* lam;            /* [2][W][Nj] adjoint of step m in slot m mod 2 */
  double* u;              /* [W][Nj] scratch: lambda_k / c0 of the step being solved */
  const double* g;        /* [slots][W][Nj] upstream gradient of stored theta */
  const int32_t* g_slot;  /* slot of step m - slot_t0 in g, or -1 */
  int64_t slot_t0;
  int64_t g_lo;           /* g_slot covers steps [g_lo, g_hi) */
  int64_t g_hi;
  double* grad_Ic;        /* [W][Nj] += sum_k sin(arg_k) mu_k */
  double* grad_Is;        /* [W][Nj] += sum_k -mu_k */
  double* grad_f;         /* [W][Nf] by solve position: += sum_k -2 pi K^-1 A D lambda_k */
  double* zbuf;           /* face store as in jja_problem_run */
} jja_adjoint_run;

/*
 * Process steps k_hi down to k_lo (k_hi >= k_lo >= 0). On entry lam[k_hi mod 2]
 * holds lambda_{k_hi} complete and u = lambda_{k_hi}/c0, and lam[(k_hi - 1) mod 2]
 * the contributions to lambda_{k_hi - 1} from later steps; with prime = 1 the
 * kernel first adds g_{k_hi} to lam[k_hi mod 2] and forms u itself. On exit the
 * same holds for k_lo - 1. The stencil width must be 3.
 */
int jja_adjoint(jja_plan* plan, const jja_adjoint_run* runs, int64_t k_hi, int64_t k_lo, void* stream);

/*
 * Solve K x = b for `cols` right-hand sides with the factor and solve program
 * of problem i (K: the matrix its descriptor was built from; b and x in face
 * order, [cols][n] with leading dimensions ldb, ldx >= n; DEVICE pointers).
 * zbuf: [cols][n + max_stage_rows] doubles of device scratch when the problem
 * has the global face store (jja_plan_zbuf_doubles > 0), else NULL. Replaces
 * the factorize-once / solve-many pattern of scipy.sparse.linalg.factorized as
 * the reference uses it for K (time_evolution.py:505-506) and for the cycle and
 * cut-space systems A A^T and Mr Mr^T (josephson_circuit.py:93-128). A plan
 * made only for solving has no junctions (n_junctions 0; jja_evolve refuses
 * it). Not for plans whose columns are split over workgroups.
 */
int jja_solve(jja_plan* plan, int32_t i, const double* b, int64_t ldb, double* x, int64_t ldx, int32_t cols,
              double* zbuf, void* stream);

/*
 * jja_solve for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc); no counterpart in the
 * reference. It calls no hipMalloc, hipFree or allocating library routine: the
 * global face store of its columns (cols * jja_plan_zbuf_doubles(plan, i)
 * doubles) is a library array of 2^21 doubles (16 MiB) per GPU, held until the
 * process ends; a call needing more is refused (JJA_EINVAL). Every call on one
 * GPU uses that GPU's array, so calls that overlap on one GPU (other streams,
 * other threads) race. The plan's device must be the current device. Results
 * equal jja_solve's, bit for bit.
 */
int jja_solve_noalloc(jja_plan* plan, int32_t i, const double* b, int64_t ldb, double* x, int64_t ldx,
                      int32_t cols, void* stream);

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
