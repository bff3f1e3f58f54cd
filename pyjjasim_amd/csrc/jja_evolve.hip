// libjja: MI355X (gfx950) kernels and device plan for the semi-implicit RCSJ
// time evolution of Josephson junction arrays.
//
// One workgroup integrates one column (problem w) of one circuit for a whole
// chunk of steps: columns never exchange data, so the launch needs no grid-wide
// synchronisation. Per step (stencil 3 shown; time_evolution.py:523-580):
//
//   junctions  x_j = eta - Is + Ic sin(2 th1 - th2) + c1 th1 + c2 th2
//   faces      b_f = sum_j A_fj (x_j / c0_j - ths_j) - 2 pi f_f      (LDS)
//   solve      J = K^-1 b : staged forward/backward triangular solves  (LDS)
//   junctions  y_j = sum_f A_fj J_f ; th_new = (y - x) / c0 ; I = y + Is
//
// The face vector lives in LDS for the whole step. The oldest phase slot of the
// history holds x between the first and the last junction phase, then the new
// phase, so the state needs stencil_width-1 vectors per column and nothing else.
// Arithmetic follows the reference's operation order term by term (no FMA
// contraction in the step expressions), so results agree with the CPU scheme to
// rounding of sin and of the linear solve.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/jja.h"

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

#define JJA_HIP(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(JJA_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
  } while (0)

constexpr double kTwoPi = 6.283185307179586;  // 2 * np.pi

// Device-resident description of one circuit (all pointers device memory).
struct DevProblem {
  int32_t Nj, Nf, sw, n_stages;
  int32_t max_stage_rows, pad_;
  const int32_t* face_ptr;
  const int32_t* face_code;
  const int32_t* junc_code;
  const double* Ic;
  const double* Rv;
  const double* coef;
  const int32_t* face_pos;
  const int32_t* stage_ptr;
  const int64_t* lo_ptr;
  const int32_t* lo_col;
  const double* lo_val;
  const int64_t* up_ptr;
  const int32_t* up_col;
  const double* up_val;
  const int32_t* row_block;
  const int32_t* row_bsize;
  const int64_t* row_boff;
  const double* inv_cm;
  const double* inv_rm;
};

struct WorkItem {
  int32_t problem;
  int32_t column;
};

__device__ __forceinline__ double src_at(const jja_source& s, int64_t i, int64_t w, int64_t t) {
  const int64_t tt = t - s.t0;
  double v = s.base[i * s.s_i + w * s.s_w + tt * s.s_t];
  if (s.amp) v = v * s.amp[w * s.a_w + tt * s.a_t];
  return v;
}

// Philox4x32-10 (Salmon et al., SC'11) + Box-Muller; counter (j>>1, w, step, 0).
__device__ __forceinline__ double philox_normal(uint64_t seed, uint32_t j, uint32_t w,
                                                uint64_t step) {
  uint32_t c0 = j >> 1, c1 = w, c2 = static_cast<uint32_t>(step), c3 = 0u;
  uint32_t k0 = static_cast<uint32_t>(seed), k1 = static_cast<uint32_t>(seed >> 32);
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
    const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
    const uint32_t n0 = static_cast<uint32_t>(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = static_cast<uint32_t>(p0 >> 32) ^ c3 ^ k1;
    c1 = static_cast<uint32_t>(p1);
    c3 = static_cast<uint32_t>(p0);
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  const uint64_t a = ((static_cast<uint64_t>(c1) << 32) | c0) >> 11;
  const uint64_t b = ((static_cast<uint64_t>(c3) << 32) | c2) >> 11;
  const double u1 = static_cast<double>(a + 1) * 0x1.0p-53;
  const double u2 = static_cast<double>(b) * 0x1.0p-53;
  const double rho = sqrt(-2.0 * log(u1));
  const double ang = kTwoPi * u2;
  return (j & 1u) ? rho * sin(ang) : rho * cos(ang);
}

__device__ __forceinline__ int64_t col_start(int32_t c, int32_t B) {
  return static_cast<int64_t>(c) * B - (static_cast<int64_t>(c) * (c - 1)) / 2;
}

__device__ __forceinline__ int64_t row_start(int32_t c) {
  return (static_cast<int64_t>(c) * (c + 1)) / 2;
}

template <int BS>
__global__ __launch_bounds__(BS) void evolve_kernel(const DevProblem* __restrict__ probs,
                                                    const jja_problem_run* __restrict__ runs,
                                                    const WorkItem* __restrict__ work,
                                                    int32_t n_steps, int64_t step0) {
  extern __shared__ double lds[];
  const WorkItem wi = work[blockIdx.x];
  const DevProblem& P = probs[wi.problem];
  const jja_problem_run& R = runs[wi.problem];
  const int tid = threadIdx.x;
  const int32_t Nj = P.Nj, Nf = P.Nf, sw = P.sw, S = sw - 1;
  const int64_t w = wi.column;
  const int64_t W = R.W;
  double* __restrict__ zs = lds;        // face vector, by solve position
  double* __restrict__ ts = lds + Nf;   // per-stage scratch
  const size_t slot_stride = static_cast<size_t>(W) * Nj;
  double* __restrict__ th = R.theta + static_cast<size_t>(w) * Nj;
  double* __restrict__ ths = R.theta_s ? R.theta_s + static_cast<size_t>(w) * Nj : nullptr;
  const double* __restrict__ c0 = P.coef;
  const double* __restrict__ c1 = P.coef + Nj;
  const double* __restrict__ c2 = P.coef + 2 * static_cast<size_t>(Nj);
  const double* __restrict__ c3 = P.coef + 3 * static_cast<size_t>(Nj);
  const double* __restrict__ c4 = P.coef + 4 * static_cast<size_t>(Nj);
  const uint32_t noise_col = static_cast<uint32_t>(R.column_offset + w);

  for (int32_t n = 0; n < n_steps; ++n) {
    const int64_t t = step0 + n;
    // history slot of phase at step m is m mod S (m >= -S)
    auto slot = [&](int64_t m) -> size_t {
      int64_t r = m % S;
      return static_cast<size_t>(r < 0 ? r + S : r) * slot_stride;
    };
    const size_t s_new = slot(t), s1 = slot(t - 1), s2 = slot(t - 2);
    const size_t s3 = sw >= 4 ? slot(t - 3) : 0, s4 = sw >= 5 ? slot(t - 4) : 0;

    // ---- junction phase: x_j into the oldest slot (time_evolution.py:533-558)
    for (int32_t j = tid; j < Nj; j += BS) {
      const double t1 = th[s1 + j];
      const double t2 = th[s2 + j];
      double X;
      if (sw == 3) {
        X = P.Ic[j] * sin(2.0 * t1 - t2) + c1[j] * t1 + c2[j] * t2;
      } else if (sw == 4) {
        const double t3 = th[s3 + j];
        X = P.Ic[j] * sin(3.0 * t1 - 3.0 * t2 + t3) + c1[j] * t1 + c2[j] * t2 + c3[j] * t3;
      } else {
        const double t3 = th[s3 + j];
        const double t4 = th[s4 + j];
        X = P.Ic[j] * sin(4.0 * t1 - 6.0 * t2 + 4.0 * t3 - t4) + c1[j] * t1 + c2[j] * t2 +
            c3[j] * t3 + c4[j] * t4;
      }
      double eta = 0.0;
      if (!R.T.zero) {
        const double z = R.noise ? R.noise[static_cast<size_t>(t - R.noise_t0) * slot_stride +
                                           static_cast<size_t>(w) * Nj + j]
                                 : philox_normal(R.seed, static_cast<uint32_t>(j), noise_col,
                                                 static_cast<uint64_t>(t));
        eta = sqrt(2.0 * src_at(R.T, j, w, t) * P.Rv[j]) * z;
      }
      const double x = R.Is.zero ? eta + X : (eta - src_at(R.Is, j, w, t)) + X;
      th[s_new + j] = x;
    }
    __syncthreads();

    // ---- face phase: b = A (x/c0 - ths) - 2 pi f (time_evolution.py:560-569)
    for (int32_t f = tid; f < Nf; f += BS) {
      double acc = 0.0;
      for (int32_t e = P.face_ptr[f]; e < P.face_ptr[f + 1]; ++e) {
        const int32_t code = P.face_code[e];
        const int32_t j = code >> 1;
        double u = th[s_new + j] / c0[j];
        if (ths) u = u - ths[j];
        acc = (code & 1) ? acc - u : acc + u;
      }
      if (!R.f.zero) acc = acc - kTwoPi * src_at(R.f, f, w, t);
      zs[P.face_pos[f]] = acc;
    }
    __syncthreads();

    // ---- forward solve L z = b, stage by stage
    for (int32_t s = 0; s < P.n_stages; ++s) {
      const int32_t r0 = P.stage_ptr[s], r1 = P.stage_ptr[s + 1];
      for (int32_t r = r0 + tid; r < r1; r += BS) {
        double acc = zs[r];
        for (int64_t e = P.lo_ptr[r]; e < P.lo_ptr[r + 1]; ++e)
          acc = __builtin_fma(-P.lo_val[e], zs[P.lo_col[e]], acc);
        ts[r - r0] = acc;
      }
      __syncthreads();
      for (int32_t r = r0 + tid; r < r1; r += BS) {
        const int32_t b0 = P.row_block[r], B = P.row_bsize[r], a = r - b0;
        const double* __restrict__ inv = P.inv_cm + P.row_boff[r];
        double z = 0.0;
        for (int32_t c = 0; c <= a; ++c)
          z = __builtin_fma(inv[col_start(c, B) + (a - c)], ts[b0 + c - r0], z);
        (void)B;
        zs[r] = z;
      }
      __syncthreads();
    }
    // ---- backward solve L^T J = z
    for (int32_t s = P.n_stages - 1; s >= 0; --s) {
      const int32_t r0 = P.stage_ptr[s], r1 = P.stage_ptr[s + 1];
      for (int32_t r = r0 + tid; r < r1; r += BS) {
        double acc = zs[r];
        for (int64_t e = P.up_ptr[r]; e < P.up_ptr[r + 1]; ++e)
          acc = __builtin_fma(-P.up_val[e], zs[P.up_col[e]], acc);
        ts[r - r0] = acc;
      }
      __syncthreads();
      for (int32_t r = r0 + tid; r < r1; r += BS) {
        const int32_t b0 = P.row_block[r], B = P.row_bsize[r], a = r - b0;
        const double* __restrict__ inv = P.inv_rm + P.row_boff[r];
        double z = 0.0;
        for (int32_t c = a; c < B; ++c)
          z = __builtin_fma(inv[row_start(c) + a], ts[b0 + c - r0], z);
        zs[r] = z;
      }
      __syncthreads();
    }

    // ---- junction update (time_evolution.py:570-580)
    const int32_t tho = R.th_slot ? R.th_slot[t - R.slot_t0] : -1;
    const int32_t Io = R.I_slot ? R.I_slot[t - R.slot_t0] : -1;
    for (int32_t j = tid; j < Nj; j += BS) {
      double y = 0.0;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int32_t code = P.junc_code[2 * static_cast<size_t>(j) + k];
        if (code >= 0) {
          const double J = zs[P.face_pos[code >> 1]];
          y = (code & 1) ? y - J : y + J;
        }
      }
      const double x = th[s_new + j];
      const double tn = (y - x) / c0[j];
      th[s_new + j] = tn;
      if (tho >= 0) R.th_out[static_cast<size_t>(tho) * slot_stride + static_cast<size_t>(w) * Nj + j] = tn;
      if (Io >= 0)
        R.I_out[static_cast<size_t>(Io) * slot_stride + static_cast<size_t>(w) * Nj + j] =
            y + src_at(R.Is, j, w, t);
      if (ths) ths[j] += src_at(R.Vs, j, w, t) * R.dt;
    }
    __syncthreads();
  }
}

__global__ void vortex_kernel(const DevProblem* __restrict__ probs, int32_t p,
                              const double* __restrict__ theta, int32_t cols,
                              int32_t* __restrict__ n_out) {
  const DevProblem& P = probs[p];
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<int64_t>(P.Nf) * cols) return;
  const int64_t col = idx / P.Nf;
  const int32_t f = static_cast<int32_t>(idx - col * P.Nf);
  const double* th = theta + col * P.Nj;
  int64_t acc = 0;
  for (int32_t e = P.face_ptr[f]; e < P.face_ptr[f + 1]; ++e) {
    const int32_t code = P.face_code[e];
    const int64_t k = static_cast<int64_t>(rint(th[code >> 1] / kTwoPi));
    acc += (code & 1) ? k : -k;
  }
  n_out[idx] = static_cast<int32_t>(acc);
}

struct Plan {
  int device = 0;
  std::vector<void*> allocations;
  std::vector<DevProblem> problems;
  std::vector<int64_t> lds_bytes;
  DevProblem* d_problems = nullptr;
  jja_problem_run* d_runs = nullptr;
  WorkItem* d_work = nullptr;
  size_t work_cap = 0;
  int64_t lds_limit = 0;
};

template <typename T>
int upload(Plan* plan, const T* host, size_t n, const T** out) {
  *out = nullptr;
  if (n == 0) return JJA_OK;
  if (!host) return fail(JJA_EINVAL, "jja_plan_create: missing array");
  void* d = nullptr;
  JJA_HIP(hipMalloc(&d, n * sizeof(T)));
  plan->allocations.push_back(d);
  JJA_HIP(hipMemcpy(d, host, n * sizeof(T), hipMemcpyHostToDevice));
  *out = static_cast<const T*>(d);
  return JJA_OK;
}

void free_plan(Plan* plan) {
  for (void* p : plan->allocations) (void)hipFree(p);
  delete plan;
}

int validate(const jja_problem_desc& d, int64_t* max_stage_rows) {
  if (d.n_junctions <= 0 || d.n_faces < 0 || d.stencil_width < 3 || d.stencil_width > 5 ||
      d.n_stages < 0)
    return fail(JJA_EINVAL, "jja_plan_create: bad sizes");
  if (!d.face_ptr || !d.junc_code || !d.Ic || !d.Rv || !d.coef)
    return fail(JJA_EINVAL, "jja_plan_create: missing topology or coefficients");
  const int32_t Nf = d.n_faces, Nj = d.n_junctions;
  for (int32_t f = 0; f < Nf; ++f)
    if (d.face_ptr[f + 1] < d.face_ptr[f]) return fail(JJA_EINVAL, "face_ptr not monotone");
  for (int32_t e = 0; e < d.face_ptr[Nf]; ++e)
    if (d.face_code[e] < 0 || (d.face_code[e] >> 1) >= Nj)
      return fail(JJA_EINVAL, "face_code out of range");
  for (int64_t e = 0; e < 2 * static_cast<int64_t>(Nj); ++e)
    if (d.junc_code[e] < -1 || (d.junc_code[e] >> 1) >= Nf)
      return fail(JJA_EINVAL, "junc_code out of range");
  if (Nf > 0) {
    if (!d.face_pos || !d.stage_ptr || !d.lo_ptr || !d.up_ptr || !d.row_block ||
        !d.row_bsize || !d.row_boff)
      return fail(JJA_EINVAL, "jja_plan_create: missing solve schedule");
    if (d.stage_ptr[0] != 0 || d.stage_ptr[d.n_stages] != Nf)
      return fail(JJA_EINVAL, "stage_ptr must span all faces");
    std::vector<char> seen(Nf, 0);
    for (int32_t f = 0; f < Nf; ++f) {
      const int32_t q = d.face_pos[f];
      if (q < 0 || q >= Nf || seen[q]) return fail(JJA_EINVAL, "face_pos is not a permutation");
      seen[q] = 1;
    }
    int64_t ms = 0;
    for (int32_t s = 0; s < d.n_stages; ++s) {
      const int32_t r0 = d.stage_ptr[s], r1 = d.stage_ptr[s + 1];
      if (r1 < r0) return fail(JJA_EINVAL, "stage_ptr not monotone");
      ms = std::max<int64_t>(ms, r1 - r0);
      for (int32_t r = r0; r < r1; ++r) {
        const int32_t b0 = d.row_block[r], B = d.row_bsize[r];
        if (b0 < r0 || B <= 0 || r < b0 || r >= b0 + B || b0 + B > r1)
          return fail(JJA_EINVAL, "block rows must stay inside their stage");
        const int64_t need = d.row_boff[r] + static_cast<int64_t>(B) * (B + 1) / 2;
        if (d.row_boff[r] < 0 || need > d.n_inv) return fail(JJA_EINVAL, "row_boff out of range");
        for (int64_t e = d.lo_ptr[r]; e < d.lo_ptr[r + 1]; ++e)
          if (d.lo_col[e] < 0 || d.lo_col[e] >= r0) return fail(JJA_EINVAL, "lo_col must point to earlier stages");
        for (int64_t e = d.up_ptr[r]; e < d.up_ptr[r + 1]; ++e)
          if (d.up_col[e] < r1 || d.up_col[e] >= Nf) return fail(JJA_EINVAL, "up_col must point to later stages");
      }
    }
    *max_stage_rows = ms;
  } else {
    *max_stage_rows = 0;
  }
  return JJA_OK;
}

template <int BS>
int launch_evolve(Plan* plan, int32_t n_work, int64_t lds, int32_t n_steps, int64_t step0,
                  hipStream_t stream) {
  auto* kern = evolve_kernel<BS>;
  JJA_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
  hipLaunchKernelGGL(kern, dim3(n_work), dim3(BS), static_cast<size_t>(lds), stream,
                     plan->d_problems, plan->d_runs, plan->d_work, n_steps, step0);
  JJA_HIP(hipGetLastError());
  return JJA_OK;
}

}  // namespace

extern "C" {

const char* jja_last_error(void) { return g_error.c_str(); }

int jja_plan_create(int32_t n_problems, const jja_problem_desc* descs, jja_plan** out) {
  if (!out || n_problems <= 0 || !descs) return fail(JJA_EINVAL, "jja_plan_create: invalid arguments");
  *out = nullptr;
  Plan* plan = new Plan();
  int rc = JJA_OK;
  hipDeviceProp_t prop;
  if (hipGetDevice(&plan->device) != hipSuccess ||
      hipGetDeviceProperties(&prop, plan->device) != hipSuccess) {
    delete plan;
    return fail(JJA_EHIP, "jja_plan_create: no HIP device");
  }
  plan->lds_limit = static_cast<int64_t>(prop.sharedMemPerBlock);
  for (int32_t p = 0; p < n_problems && rc == JJA_OK; ++p) {
    const jja_problem_desc& d = descs[p];
    int64_t ms = 0;
    rc = validate(d, &ms);
    if (rc) break;
    DevProblem dp{};
    dp.Nj = d.n_junctions;
    dp.Nf = d.n_faces;
    dp.sw = d.stencil_width;
    dp.n_stages = d.n_stages;
    dp.max_stage_rows = static_cast<int32_t>(ms);
    const size_t Nj = d.n_junctions, Nf = d.n_faces;
    const size_t nnzA = static_cast<size_t>(d.face_ptr[Nf]);
    const size_t nlo = Nf ? static_cast<size_t>(d.lo_ptr[Nf]) : 0;
    const size_t nup = Nf ? static_cast<size_t>(d.up_ptr[Nf]) : 0;
    if ((rc = upload(plan, d.face_ptr, Nf + 1, &dp.face_ptr))) break;
    if ((rc = upload(plan, d.face_code, nnzA, &dp.face_code))) break;
    if ((rc = upload(plan, d.junc_code, 2 * Nj, &dp.junc_code))) break;
    if ((rc = upload(plan, d.Ic, Nj, &dp.Ic))) break;
    if ((rc = upload(plan, d.Rv, Nj, &dp.Rv))) break;
    if ((rc = upload(plan, d.coef, Nj * d.stencil_width, &dp.coef))) break;
    if (Nf) {
      if ((rc = upload(plan, d.face_pos, Nf, &dp.face_pos))) break;
      if ((rc = upload(plan, d.stage_ptr, static_cast<size_t>(d.n_stages) + 1, &dp.stage_ptr))) break;
      if ((rc = upload(plan, d.lo_ptr, Nf + 1, &dp.lo_ptr))) break;
      if ((rc = upload(plan, d.lo_col, nlo, &dp.lo_col))) break;
      if ((rc = upload(plan, d.lo_val, nlo, &dp.lo_val))) break;
      if ((rc = upload(plan, d.up_ptr, Nf + 1, &dp.up_ptr))) break;
      if ((rc = upload(plan, d.up_col, nup, &dp.up_col))) break;
      if ((rc = upload(plan, d.up_val, nup, &dp.up_val))) break;
      if ((rc = upload(plan, d.row_block, Nf, &dp.row_block))) break;
      if ((rc = upload(plan, d.row_bsize, Nf, &dp.row_bsize))) break;
      if ((rc = upload(plan, d.row_boff, Nf, &dp.row_boff))) break;
      if ((rc = upload(plan, d.inv_cm, static_cast<size_t>(d.n_inv), &dp.inv_cm))) break;
      if ((rc = upload(plan, d.inv_rm, static_cast<size_t>(d.n_inv), &dp.inv_rm))) break;
    }
    plan->problems.push_back(dp);
    const int64_t bytes = (static_cast<int64_t>(d.n_faces) + ms) * static_cast<int64_t>(sizeof(double));
    plan->lds_bytes.push_back(bytes);
  }
  if (rc == JJA_OK) {
    const DevProblem* dprob = nullptr;
    rc = upload(plan, plan->problems.data(), plan->problems.size(), &dprob);
    plan->d_problems = const_cast<DevProblem*>(dprob);
  }
  if (rc == JJA_OK) {
    void* d = nullptr;
    if (hipMalloc(&d, sizeof(jja_problem_run) * n_problems) != hipSuccess) {
      rc = fail(JJA_ENOMEM, "jja_plan_create: device allocation failed");
    } else {
      plan->allocations.push_back(d);
      plan->d_runs = static_cast<jja_problem_run*>(d);
    }
  }
  if (rc != JJA_OK) {
    free_plan(plan);
    return rc;
  }
  *out = reinterpret_cast<jja_plan*>(plan);
  return JJA_OK;
}

int jja_plan_destroy(jja_plan* p) {
  if (!p) return JJA_OK;
  free_plan(reinterpret_cast<Plan*>(p));
  return JJA_OK;
}

int64_t jja_plan_lds_bytes(const jja_plan* p, int32_t i) {
  const Plan* plan = reinterpret_cast<const Plan*>(p);
  if (!plan || i < 0 || i >= static_cast<int32_t>(plan->lds_bytes.size())) return 0;
  return plan->lds_bytes[i] <= plan->lds_limit ? plan->lds_bytes[i] : 0;
}

int jja_evolve(jja_plan* p, const jja_problem_run* runs, int32_t n_steps, int64_t step0,
               int32_t block_size, void* stream_ptr) {
  Plan* plan = reinterpret_cast<Plan*>(p);
  if (!plan || !runs || n_steps < 0 || step0 < 0) return fail(JJA_EINVAL, "jja_evolve: invalid arguments");
  if (n_steps == 0) return JJA_OK;
  const int32_t np = static_cast<int32_t>(plan->problems.size());
  std::vector<WorkItem> work;
  int64_t lds = 0;
  for (int32_t i = 0; i < np; ++i) {
    const jja_problem_run& r = runs[i];
    const DevProblem& d = plan->problems[i];
    if (r.W <= 0 || !r.theta) return fail(JJA_EINVAL, "jja_evolve: run needs W > 0 and theta");
    if (r.dt <= 0) return fail(JJA_EINVAL, "jja_evolve: dt must be positive");
    if (!r.Is.base || !r.Vs.base || !r.f.base || !r.T.base)
      return fail(JJA_EINVAL, "jja_evolve: every source needs a base array");
    if (!r.Vs.zero && !r.theta_s) return fail(JJA_EINVAL, "jja_evolve: non-zero Vs needs theta_s");
    if (plan->lds_bytes[i] > plan->lds_limit)
      return fail(JJA_EUNSUPPORTED, "jja_evolve: face vector of problem " + std::to_string(i) +
                                        " needs " + std::to_string(plan->lds_bytes[i]) +
                                        " bytes of LDS; limit " + std::to_string(plan->lds_limit));
    (void)d;
    lds = std::max(lds, plan->lds_bytes[i]);
    for (int32_t w = 0; w < r.W; ++w) work.push_back({i, w});
  }
  if (work.empty()) return JJA_OK;
  if (work.size() > plan->work_cap) {
    if (plan->d_work) {
      auto it = std::find(plan->allocations.begin(), plan->allocations.end(), static_cast<void*>(plan->d_work));
      if (it != plan->allocations.end()) plan->allocations.erase(it);
      JJA_HIP(hipFree(plan->d_work));
      plan->d_work = nullptr;
    }
    void* d = nullptr;
    JJA_HIP(hipMalloc(&d, work.size() * sizeof(WorkItem)));
    plan->allocations.push_back(d);
    plan->d_work = static_cast<WorkItem*>(d);
    plan->work_cap = work.size();
  }
  hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
  // pageable source: the call returns once the data is staged, and the copy is
  // ordered after earlier work on the stream
  JJA_HIP(hipMemcpyAsync(plan->d_runs, runs, sizeof(jja_problem_run) * np, hipMemcpyHostToDevice, stream));
  JJA_HIP(hipMemcpyAsync(plan->d_work, work.data(), sizeof(WorkItem) * work.size(),
                         hipMemcpyHostToDevice, stream));
  const int32_t n_work = static_cast<int32_t>(work.size());
  if (lds < 8) lds = 8;
  switch (block_size) {
    case 64: return launch_evolve<64>(plan, n_work, lds, n_steps, step0, stream);
    case 128: return launch_evolve<128>(plan, n_work, lds, n_steps, step0, stream);
    case 256: return launch_evolve<256>(plan, n_work, lds, n_steps, step0, stream);
    case 512: return launch_evolve<512>(plan, n_work, lds, n_steps, step0, stream);
    case 0:
    case 1024: return launch_evolve<1024>(plan, n_work, lds, n_steps, step0, stream);
    default: return fail(JJA_EINVAL, "jja_evolve: block_size must be 64, 128, 256, 512 or 1024");
  }
}

int jja_vortex(jja_plan* p, int32_t i, const double* theta, int32_t cols, int32_t* n_out,
               void* stream_ptr) {
  Plan* plan = reinterpret_cast<Plan*>(p);
  if (!plan || i < 0 || i >= static_cast<int32_t>(plan->problems.size()) || !theta || !n_out || cols < 0)
    return fail(JJA_EINVAL, "jja_vortex: invalid arguments");
  const int64_t total = static_cast<int64_t>(plan->problems[i].Nf) * cols;
  if (total == 0) return JJA_OK;
  const int threads = 256;
  const int64_t blocks = (total + threads - 1) / threads;
  hipLaunchKernelGGL(vortex_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(threads), 0,
                     static_cast<hipStream_t>(stream_ptr), plan->d_problems, i, theta, cols, n_out);
  JJA_HIP(hipGetLastError());
  return JJA_OK;
}

}  // extern "C"
