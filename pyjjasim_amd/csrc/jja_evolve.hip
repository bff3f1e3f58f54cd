// libjja: MI355X (gfx950) kernels and device plan for the semi-implicit RCSJ
// time evolution of Josephson junction arrays.
//
// One workgroup integrates one column (problem w) of one circuit for a whole
// chunk of steps: columns never exchange data, so a launch needs no grid-wide
// synchronisation. Per step (stencil 3 shown; time_evolution.py:523-580):
//
//   junctions  x_j = eta - Is + Ic sin(2 th1 - th2) + c1 th1 + c2 th2
//              u_j = x_j / c0_j - ths_j                            (scratch)
//   faces      b_f = sum_j A_fj u_j - 2 pi f_f                      (LDS)
//   solve      J = K^-1 b : staged forward/backward triangular solves (LDS)
//   junctions  y_j = sum_f A_fj J_f ; th_new = (y - x) / c0 ; I = y + Is
//
// The face vector lives in LDS for the whole step. The oldest phase slot of the
// history holds x between the junction passes, then the new phase, so the
// state is stencil_width-1 vectors per column. The junction pass that finishes
// step t also prepares step t+1. Arithmetic of the step expressions follows the
// reference's operation order (no FMA contraction; a division by c0 is a
// product with the reciprocal 1/c0), so results agree with the CPU scheme to
// the rounding of sin, of those quotients and of the linear solve.
//
// The solve is a linear program of slices whose addresses do not depend on any
// data: each thread prefetches the entries of the next slice (also across the
// phase barriers, which only order LDS traffic) while it processes the current
// one, so global-memory latency overlaps the LDS gathers and reductions.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/jja.h"
#include "jja_math.h"

// Device scratch of jja_solve_noalloc (a global face store for its columns):
// 2^21 doubles (16 MiB) per GPU, held until the process ends. Calls on one
// GPU share it; calls that overlap on one GPU race.
constexpr int64_t kScratchDoubles = int64_t{1} << 21;
static __device__ __align__(256) double scratch[kScratchDoubles];

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
constexpr int kMaxIter = 8;                   // entries per lane and solve record
constexpr int kFaceIter = 4;                  // junction codes per thread and face record
constexpr int kFaceGroup = 4;                 // face records whose loads are issued together
constexpr int kWave = 64;                     // lanes of a wavefront (solve records are per wave)
// buffer resource word 3 for gfx9-family raw buffers (no format conversion),
// and a byte offset beyond any resource range (ranges stay below 2^31)
constexpr int kBufferWord3 = 0x00020000;
constexpr uint32_t kOutOfRange = 0x80000000u;

// Pointers read out of the descriptor structs are generic; casting them to the
// global (1) or constant (4) address space lets the compiler use global and
// scalar loads instead of flat ones, which would also stall on LDS traffic.
#define GAS __attribute__((address_space(1)))
#define CAS __attribute__((address_space(4)))
#define LAS __attribute__((address_space(3)))
template <typename T>
__device__ __forceinline__ const GAS T* gbl(const T* p) {
  return (const GAS T*)p;
}
template <typename T>
__device__ __forceinline__ GAS T* gblw(T* p) {
  return (GAS T*)p;
}

typedef double vdouble2 __attribute__((ext_vector_type(2)));
typedef double vdouble4 __attribute__((ext_vector_type(4)));
typedef uint32_t vuint2 __attribute__((ext_vector_type(2)));
typedef uint32_t vuint4 __attribute__((ext_vector_type(4)));
typedef int32_t vint2 __attribute__((ext_vector_type(2)));

struct RecWords {
  vuint4 a, b;
};
static_assert(sizeof(RecWords) == sizeof(jja_slice), "record layout");

// Uniform load of one program record through the scalar cache.
__device__ __forceinline__ jja_slice load_rec(const jja_slice* prog, int32_t k) {
  const CAS vuint4* q = (const CAS vuint4*)(prog + k);
  RecWords w;
  w.a = q[0];
  w.b = q[1];
  return __builtin_bit_cast(jja_slice, w);
}

// Device-resident description of one circuit (all pointers device memory).
struct DevProblem {
  int32_t Nj, Nf, sw, BS;
  int32_t n_face_slices, n_solve_slices, max_stage_rows, copy_rows;
  int32_t index_bytes;
  int32_t val_bytes;        // bytes of ent_val (below 2^31: the solve reads it through a buffer resource)
  const double* Ic;
  const double* Rv;
  const double* coef;
  const int32_t* junc_code;
  const int32_t* face_of_pos;
  const jja_slice* face_slices;
  const jja_slice* solve_slices;
  const int32_t* face_code;
  const int32_t* slice_rows;
  const double* ent_val;
  const void* ent_off;      // LDS indices: uint16 (face store in LDS) or uint32 (in global memory)
  const int32_t* wave_ptr;  // [BS/64 + 1]
  const double* zero;       // 8 zero doubles: absent input terms read here
  const int32_t* face_fid;  // [n_slice_rows]: face index of a face record's row slot, or -1
  int32_t parts;            // workgroups per column (1, or more with a global face store)
  int32_t coef_si;          // junction stride of coef: 1, or 0 when every row is one value
  const double* rc0;        // [Nj] 1 / c0 (stride coef_si)
  int32_t cpr_terms;        // 0: Ic sin(arg); else Ic * sum_k cpr_coef[k] sin((k+1) arg)
  double cpr_coef[JJA_CPR_MAX_TERMS];
  // lattice solve (lat_rows 0: the staged factor program solves K)
  int32_t lat_rows, lat_cols;  // R, C: faces at positions r*C + c
  int32_t lat_es, lat_cp;      // folded row length ES (even) and sine-table row stride CP
  int32_t lat_ef;              // folded rows stored per row: ES (LDS store), ES rounded up to 4 (global)
  int32_t lat_mp, lat_bp;      // split column: modes per part and blocks per mode of the part-local
                               // tridiagonal solve in LDS (lat_mp 0: the column-wide one)
  int32_t lat_pad_;
  int64_t lat_t0;              // store offset of the folded rows (even: 16-byte loads)
  const double* lat_sine;      // [ES][CP]
  const double* lat_sine_p;    // [ES][2][CP/2]: the same table, even and odd modes apart (matrix-core B operands)
  const double* lat_inv;       // [R][CP] inverse Thomas pivots
  double lat_g;
};

// The current-phase relation's sum at one argument: sin(arg) for the default
// relation, else sum_k a_k sin((k+1) arg) accumulated in ascending k (each
// (k+1) arg formed as one product, as numpy forms k * theta).
__device__ __forceinline__ double cpr_sum(const DevProblem& P, double arg) {
  if (P.cpr_terms == 0) return jja_sin(arg);
  double acc = P.cpr_coef[0] * jja_sin(arg);
  for (int k = 1; k < P.cpr_terms; ++k) acc = acc + P.cpr_coef[k] * jja_sin(static_cast<double>(k + 1) * arg);
  return acc;
}

// Stencil coefficient k of junction j (an identical value for every junction
// is read from one address).
__device__ __forceinline__ double coef_at(const DevProblem& P, int k, int32_t j) {
  return gbl(P.coef)[static_cast<int64_t>(k) * P.Nj + static_cast<int64_t>(j) * P.coef_si];
}
// 1 / c0 of junction j: the step divides by c0 as a product with its
// reciprocal (within one rounding of the quotient)
__device__ __forceinline__ double rc0_at(const DevProblem& P, int32_t j) {
  return gbl(P.rc0)[static_cast<int64_t>(j) * P.coef_si];
}

struct WorkItem {
  int32_t problem;
  int32_t column;  // column (times the problem's parts, plus the part, in a split launch)
  int32_t slot;    // the column's barrier counter in a split launch
};

// Workgroup barrier that orders LDS traffic only: global loads in flight (the
// prefetched solve entries) stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Philox4x32-10 (Salmon et al., SC'11) + Box-Muller for the junction pair
// (j, j+1), j even; counter (j>>1, w, step, 0), key = seed.
__device__ __forceinline__ void normal_pair(uint64_t seed, uint32_t j, uint32_t w, uint64_t step,
                                            double* z0, double* z1) {
  uint32_t c0 = j >> 1, c1 = w, c2 = static_cast<uint32_t>(step), c3 = 0u;
  uint32_t k0 = static_cast<uint32_t>(seed), k1 = static_cast<uint32_t>(seed >> 32);
  // an empty statement that the compiler must assume changes the key: the
  // round keys are then formed here (two scalar adds per round) instead of
  // being hoisted out of the junction loop as 20 live scalar registers, which
  // spilled to vector lanes and cost 22 v_readlane per pair
  __asm__ volatile("" : "+s"(k0), "+s"(k1));
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
  double sn, cs;
  jja_sincos(ang, &sn, &cs);  // one argument reduction for both
  *z0 = rho * cs;
  *z1 = rho * sn;
}

// One 64-bit DPP move (two 32-bit halves); rows outside ROW_MASK and lanes
// without a source read 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), CTRL, ROW_MASK, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), CTRL, ROW_MASK, 0xf, true);
  return __builtin_bit_cast(double, static_cast<long long>(static_cast<unsigned>(lo)) |
                                        (static_cast<long long>(hi) << 32));
}

// Sum over each aligned group of G lanes (G a power of two <= 64), left in the
// group's last lane. Fixed order (shift-add within 16-lane rows, then row
// broadcasts), so the bits do not depend on anything but the inputs.
__device__ __forceinline__ double group_sum_last(double v, int G) {
  if (G >= 2) v += dpp_f64<0x111, 0xf>(v);   // row_shr:1
  if (G >= 4) v += dpp_f64<0x112, 0xf>(v);   // row_shr:2
  if (G >= 8) v += dpp_f64<0x114, 0xf>(v);   // row_shr:4
  if (G >= 16) v += dpp_f64<0x118, 0xf>(v);  // row_shr:8
  if (G >= 32) v += dpp_f64<0x142, 0xa>(v);  // row_bcast:15 into rows 1, 3
  if (G >= 64) v += dpp_f64<0x143, 0xc>(v);  // row_bcast:31 into rows 2, 3
  return v;
}

// Where a column keeps its face vector z (by solve position) and the stage
// scratch t after it, and which threads work on the column. FaceStore<false>:
// the workgroup's LDS, addressed by the plan's 16-bit indices as absolute LDS
// addresses (the evolve and adjoint kernels have no static LDS, checked on the
// host before every launch). FaceStore<true>: a per-column array in global
// memory for circuits whose face vector does not fit in LDS, addressed by
// 32-bit indices. SplitStore: the global array of a column that several
// workgroups (its parts) share. The arithmetic on the values is the same in all.
//
// vt and vbs are the thread's index within the column and the column's thread
// count (threadIdx.x and the block size, or part * blockDim + threadIdx.x and
// parts * blockDim). sync() orders all memory between the column's threads;
// barrier() orders the face vector only (the solve's phase boundaries).
template <bool GLOBAL>
struct FaceStore;

template <>
struct FaceStore<false> {
  static constexpr bool kGlobal = false;
  double* p;
  int32_t vt, vbs;
  __device__ __forceinline__ double rd(int64_t i) const { return p[i]; }
  __device__ __forceinline__ double gather(uint32_t i) const {
    return *reinterpret_cast<const LAS double*>(static_cast<uintptr_t>(i) << 3);
  }
  __device__ __forceinline__ void wr(int64_t i, double v) const { p[i] = v; }
  // elements i and i+1 (i even) in one 16-byte load
  __device__ __forceinline__ vdouble2 rd2(int64_t i) const {
    return *reinterpret_cast<const LAS vdouble2*>(static_cast<uintptr_t>(i) << 3);
  }
  // orders LDS traffic only: global loads in flight stay in flight
  __device__ __forceinline__ void barrier() const { lds_barrier(); }
  __device__ __forceinline__ void wg_barrier() const { lds_barrier(); }
  __device__ __forceinline__ void sync() const { __syncthreads(); }
};

template <>
struct FaceStore<true> {
  static constexpr bool kGlobal = true;
  GAS double* p;
  int32_t vt, vbs;
  __device__ __forceinline__ double rd(int64_t i) const { return p[i]; }
  __device__ __forceinline__ double gather(uint32_t i) const { return p[i]; }
  __device__ __forceinline__ void wr(int64_t i, double v) const { p[i] = v; }
  __device__ __forceinline__ vdouble2 rd2(int64_t i) const {
    return *reinterpret_cast<const GAS vdouble2*>(p + i);
  }
  __device__ __forceinline__ void barrier() const { __syncthreads(); }
  __device__ __forceinline__ void wg_barrier() const { __syncthreads(); }
  __device__ __forceinline__ void sync() const { __syncthreads(); }
};

// Dynamic LDS a split workgroup may take for its part of a lattice solve (the
// workgroup's 160 KiB less room for its static variables)
constexpr int64_t kSplitLdsMax = 163840 - 256;

// Default polls of one column barrier before a workgroup gives up (each poll
// sleeps briefly; far longer than any phase of a step takes).
constexpr uint32_t kWaitLimit = 1u << 21;
// s_waitcnt immediate (gfx9 encoding) that waits for vmcnt 0 only: expcnt and
// lgkmcnt fields at their maxima
constexpr int kWaitVmem = 0x0F70;

// The XCD (accelerator complex die) the calling wave runs on: hardware
// register XCC_ID, bits 3:0.
__device__ __forceinline__ uint32_t xcd_id() {
  return __builtin_amdgcn_s_getreg((3 << 11) | 20) & 15u;
}

struct SplitStore {
  static constexpr bool kGlobal = true;
  GAS double* p;
  int32_t vt, vbs;
  uint32_t* ctr;         // the column's arrival counter, zero at launch
  uint32_t* xcds;        // bit x set by each part running on XCD x, zero at launch
  uint32_t parts;
  uint32_t wait_limit;   // polls before a wait gives up
  int32_t* err;          // [0] set when a wait gave up; [1] counts one-XCD workgroups
  int32_t* stopped;      // this workgroup's flag (LDS): it no longer waits
  mutable uint32_t gen;  // arrivals to wait for (thread 0)
  mutable bool one_xcd;  // every part of the column shares this XCD's L2 (thread 0)
  LAS double* lds;       // this workgroup's dynamic LDS (the part-local tridiagonal solve)
  __device__ __forceinline__ double rd(int64_t i) const { return p[i]; }
  __device__ __forceinline__ double gather(uint32_t i) const { return p[i]; }
  __device__ __forceinline__ void wr(int64_t i, double v) const { p[i] = v; }
  __device__ __forceinline__ vdouble2 rd2(int64_t i) const {
    return *reinterpret_cast<const GAS vdouble2*>(p + i);
  }
  // Every wave waits for its own memory operations (stores reach the L2,
  // which the L1 writes through to), the workgroup meets, and thread 0 alone
  // arrives on the column's counter, polls it with relaxed loads (an acquire
  // per poll would invalidate caches under the working workgroups) and
  // acquires once at device scope (the invalidation covers the compute unit's
  // other waves); then the workgroup meets again.
  //   Parts on several XCDs (separate L2s) release at device scope before the
  // arrival: an L2 write-back. Parts on one XCD skip it: their stores already
  // meet in that XCD's L2. The parts publish their XCDs at launch
  // (split_store); each workgroup reads the column's set after its first
  // barrier, which always releases.
  //   A wait that gives up sets err and stops this workgroup's waits for the
  // rest of the launch (its results are then invalid, and the host reports
  // the error).
  __device__ __forceinline__ void sync() const {
    if (parts == 1) {  // a column of one workgroup (in a launch with split columns)
      __syncthreads();
      return;
    }
    __builtin_amdgcn_s_waitcnt(kWaitVmem);
    __syncthreads();
    if (threadIdx.x == 0 && !*stopped) {
      gen += parts;
      if (one_xcd)
        __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      else
        __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      uint32_t spins = 0;
      while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < gen) {
        if (++spins > wait_limit) {
          *stopped = 1;
          __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      // after the first barrier every part has published its XCD
      if (gen == parts && !*stopped) {
        one_xcd = __builtin_popcount(__hip_atomic_load(xcds, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == 1;
        if (one_xcd) __hip_atomic_fetch_add(err + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    __syncthreads();
  }
  __device__ __forceinline__ void barrier() const { sync(); }
  // the waves of this workgroup only (a part-local phase: its rows' readers
  // and writers are all in this workgroup)
  __device__ __forceinline__ void wg_barrier() const { __syncthreads(); }
};

// Entries of one solve record for one lane, held in registers: 8 values and
// their indices (two 16-bit halves per word, or one 32-bit index per word).
template <bool GLOBAL>
struct SliceRegs {
  double v[kMaxIter];
  uint32_t idx[GLOBAL ? kMaxIter : kMaxIter / 2];
  int32_t row;
};

template <bool GLOBAL>
__device__ __forceinline__ uint32_t slice_index(const SliceRegs<GLOBAL>& r, int i) {
  if constexpr (GLOBAL)
    return r.idx[i];
  else
    return (r.idx[i / 2] >> (16 * (i % 2))) & 0xffffu;
}

// Face-store index of entry i of a record. In a global store a block-phase
// record's lane reads consecutive scratch rows: its word (idx[0]) is
// base | len << 28, and entry i is base + i*G for i < len, else 0 (a padding
// entry: its value is zero). LDS stores keep explicit indices: there the
// decode costs more than the index bytes it saves.
template <bool GLOBAL>
__device__ __forceinline__ uint32_t record_index(const SliceRegs<GLOBAL>& r, int32_t flags, int i) {
  if (GLOBAL && (flags & JJA_SLICE_BLOCK)) {
    const uint32_t w = r.idx[0];
    const uint32_t G = static_cast<uint32_t>(flags & 0xff);
    return static_cast<uint32_t>(i) < (w >> 28) ? (w & 0x0FFFFFFFu) + static_cast<uint32_t>(i) * G : 0u;
  }
  // a NARROW off-block record of a global store: lane base (idx[0]) plus a
  // 16-bit offset (two per word in idx[1..4])
  if (GLOBAL && (flags & JJA_SLICE_NARROW))
    return r.idx[0] + ((r.idx[1 + i / 2] >> (16 * (i % 2))) & 0xffffu);
  return slice_index(r, i);
}

// Issue the loads of one solve record: always the same vector loads, each a
// scalar base (from the record) plus a per-lane offset that does not change
// between records, so no address arithmetic runs per record.
// The record values go through a buffer resource over ent_val: a value pair at
// or past the record's iteration count is read at an offset beyond the
// resource's range, which returns zeros without a memory access, so a short
// record moves only the bytes it uses and every record issues the same loads.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t value_resource(const DevProblem& P) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(P.ent_val), 0, P.val_bytes, kBufferWord3);
}

template <bool GLOBAL>
__device__ __forceinline__ void load_slice(const DevProblem& P, __amdgpu_buffer_rsrc_t vres,
                                           const jja_slice& m, int lane, SliceRegs<GLOBAL>& r) {
  const int lg = __builtin_ctz(static_cast<unsigned>(m.flags & 0xff));
  const uint32_t voff = static_cast<uint32_t>(m.base) * 8u + static_cast<uint32_t>(lane) * 16u;
  // the raw row; whether this lane's slot holds one is decided when the record
  // is processed (a select here would wait for the load)
  r.row = gbl(P.slice_rows)[m.rowbase + (lane >> lg)];
#pragma unroll
  for (int q = 0; q < kMaxIter / 2; ++q) {
    const uint32_t off = 2 * q < m.n_iter ? voff + q * 16u * kWave : kOutOfRange;
    const vdouble2 d = __builtin_bit_cast(
        vdouble2, __builtin_amdgcn_raw_buffer_load_b128(vres, static_cast<int>(off), 0, 0));
    r.v[2 * q] = d.x;
    r.v[2 * q + 1] = d.y;
  }
  if (GLOBAL && (m.flags & JJA_SLICE_BLOCK)) {
    // implicit indices: one word per lane
    r.idx[0] = gbl(static_cast<const uint32_t*>(P.ent_off) + m.ibase)[lane];
  } else if (GLOBAL && (m.flags & JJA_SLICE_NARROW)) {
    // lane base, then 8 16-bit offsets in one 16-byte load
    const uint32_t* ip = static_cast<const uint32_t*>(P.ent_off) + m.ibase;
    r.idx[0] = gbl(ip)[lane];
    const vuint4 o = gbl(reinterpret_cast<const vuint4*>(ip + kWave))[lane];
    r.idx[1] = o.x;
    r.idx[2] = o.y;
    r.idx[3] = o.z;
    r.idx[4] = o.w;
  } else if constexpr (GLOBAL) {
    const GAS vuint4* __restrict__ op =
        gbl(reinterpret_cast<const vuint4*>(static_cast<const uint32_t*>(P.ent_off) + m.ibase)) + 2 * lane;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const vuint4 o = op[q];
      r.idx[4 * q] = o.x;
      r.idx[4 * q + 1] = o.y;
      r.idx[4 * q + 2] = o.z;
      r.idx[4 * q + 3] = o.w;
    }
  } else {
    const vuint4 o = *(gbl(reinterpret_cast<const vuint4*>(static_cast<const uint16_t*>(P.ent_off) + m.ibase)) + lane);
    r.idx[0] = o.x;
    r.idx[1] = o.y;
    r.idx[2] = o.z;
    r.idx[3] = o.w;
  }
}

// Gather, accumulate, and (flag FINISH) reduce and write one record's rows.
template <class Z>
__device__ __forceinline__ void process_slice(const Z& zs, int32_t Nf, int lane, const jja_slice& m,
                                              const SliceRegs<Z::kGlobal>& r, double& acc) {
  // iterations at or past the record's count have zero values and index 0:
  // their gathers read one broadcast address and contribute exact zeros
  double zv[kMaxIter];
#pragma unroll
  for (int i = 0; i < kMaxIter; ++i) zv[i] = zs.gather(record_index(r, m.flags, i));
#pragma unroll
  for (int i = 0; i < kMaxIter; ++i) acc = __builtin_fma(r.v[i], zv[i], acc);
  if (m.flags & JJA_SLICE_FINISH) {
    const int G = m.flags & 0xff;
    acc = group_sum_last(acc, G);
    const int slot = lane >> __builtin_ctz(static_cast<unsigned>(G));
    if (slot < (m.flags >> 16) && r.row >= 0 && (lane & (G - 1)) == G - 1) {
      if (m.flags & JJA_SLICE_BLOCK)
        zs.wr(r.row, acc);
      else
        zs.wr(Nf + r.row - m.r0, zs.rd(r.row) - acc);
    }
    acc = 0.0;
  }
  if (m.flags & JJA_SLICE_BARRIER)
    zs.barrier();
  else if (m.flags & JJA_SLICE_WG_BARRIER)
    zs.wg_barrier();
}

// The whole forward + backward solve. Each wave walks its own record program,
// software-pipelined in three register buffers used in turn: the loads of
// records k+1 and k+2 are in flight while record k is processed (the
// scheduling barriers keep the compiler from hoisting a record's LDS address
// use above the next record's loads, which would make it wait for all loads),
// and descriptors are read ahead. A wave's executed record count is a multiple
// of three and four never-executed records follow it, so the loop has a single
// exit and every lookahead load stays inside the program. Waves meet only at
// BARRIER records (one per phase, the same number in every wave).
template <class Z>
__device__ __forceinline__ void solve(const DevProblem& P, const Z& zs) {
  constexpr bool GLOBAL = Z::kGlobal;
  const int lane = zs.vt & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(zs.vt / kWave);
  const CAS int32_t* wp = (const CAS int32_t*)(P.wave_ptr) + wave;
  const int32_t k0 = __builtin_amdgcn_readfirstlane(wp[0]);
  const int32_t nk = __builtin_amdgcn_readfirstlane(wp[1]) - k0 - 4;
  const jja_slice* __restrict__ prog = P.solve_slices + k0;
  jja_slice m0 = load_rec(prog, 0), m1 = load_rec(prog, 1), m2 = load_rec(prog, 2);
  const __amdgpu_buffer_rsrc_t vres = value_resource(P);
  SliceRegs<GLOBAL> A, B, C;
  load_slice(P, vres, m0, lane, A);
  load_slice(P, vres, m1, lane, B);
  double acc = 0.0;
  for (int32_t k = 0; k < nk; k += 3) {
    const jja_slice m3 = load_rec(prog, k + 3);
    load_slice(P, vres, m2, lane, C);
    __builtin_amdgcn_sched_barrier(0);
    process_slice(zs, P.Nf, lane, m0, A, acc);
    const jja_slice m4 = load_rec(prog, k + 4);
    load_slice(P, vres, m3, lane, A);
    __builtin_amdgcn_sched_barrier(0);
    process_slice(zs, P.Nf, lane, m1, B, acc);
    const jja_slice m5 = load_rec(prog, k + 5);
    load_slice(P, vres, m4, lane, B);
    __builtin_amdgcn_sched_barrier(0);
    process_slice(zs, P.Nf, lane, m2, C, acc);
    m0 = m3;
    m1 = m4;
    m2 = m5;
  }
}

// ------------------------------------------------------------ lattice solve
// The face matrix of a rectangular lattice of R x C faces (positions r*C + c)
// with one junction weight per direction is the Kronecker sum
// g (T_R (x) I_C) + g_c (I_R (x) T_C), T_n = tridiag(-1, 2, -1), so
// K^-1 = (I (x) S) M^-1 (I (x) S) with the orthonormal symmetric sine matrix S
// (DST-I) and M tridiagonal along r for each sine mode (Hockney, J. ACM 12
// (1965) 95-113). Per step: fold every row (S[C-1-i][j] = (-1)^j S[i][j]),
// transform (out[r][j] = sum_{i<H} F_j[r][i] S[i][j], H = ceil(C/2)), one
// tridiagonal solve per mode along r, then fold and transform again. Each
// transform output is summed in a fixed order (one thread in ascending i with
// fused multiply-adds in LDS; one matrix-core chain in a global store) and each
// block of a mode's recurrence runs in one thread with a blocking fixed by R, C
// and the part count, so results do not depend on the block size or the launch
// length; the two face stores agree to rounding.

// Stores the lattice solve runs on: one workgroup's (LDS or global) and a
// split column's (global, column barriers)
template <class Z>
struct RunsLattice {
  static constexpr bool value = false;
};
template <bool G>
struct RunsLattice<FaceStore<G>> {
  static constexpr bool value = true;
};
template <>
struct RunsLattice<SplitStore> {
  static constexpr bool value = true;
};

// elements of the fold each thread loads before it computes (global store)
constexpr int kFoldBatch = 4;

// E[r][i] = v[r][i] + v[r][C-1-i], O[r][i] = v[r][i] - v[r][C-1-i] for i < H
// (the middle of an odd row pairs with itself: E = v, O = 0) and 0 for
// H <= i < EF (EF = lat_ef). An LDS store keeps them by row (EF = ES; E at
// lat_t0 + r*ES + i, O at lat_t0 + (R + r)*ES + i: a lane reads 16 bytes of its
// own row per load). A global store keeps them in blocks of four i by row
// (EF = ES rounded up to 4; E at lat_t0 + ((i/4)*R + r)*4 + i%4, O at
// lat_t0 + R*EF + the same): a matrix-core operand, 16 rows x 4 i, is 512
// contiguous bytes, and a wave of the fold reads 32-byte runs of 16 rows.
// Either way element e of the loop lands at lat_t0 + e and lat_t0 + R*EF + e.
template <class Z>
__device__ __forceinline__ void lattice_fold(const DevProblem& P, const Z& zs) {
  constexpr int KB = Z::kGlobal ? kFoldBatch : 1;
  const int32_t R = P.lat_rows, C = P.lat_cols, ES = P.lat_es, H = (C + 1) >> 1;
  const int32_t n = R * P.lat_ef;
  const int64_t e0 = P.lat_t0, o0 = P.lat_t0 + static_cast<int64_t>(R) * P.lat_ef;
  // KB elements per thread at a time: all their loads are issued first
  for (int32_t e = zs.vt; e < n; e += KB * zs.vbs) {
    double x[KB], y[KB];
    int32_t ii[KB];
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      const int32_t ek = min(e + k * zs.vbs, n - 1);
      int32_t r, i;
      if constexpr (Z::kGlobal) {
        const int32_t blk = (ek >> 2) / R;
        r = (ek >> 2) - blk * R;
        i = 4 * blk + (ek & 3);
      } else {
        r = ek / ES;
        i = ek - r * ES;
      }
      ii[k] = i;
      // a padding slot (i >= H) reads the row's first element and drops it
      const int64_t row = static_cast<int64_t>(r) * C;
      const int32_t ic = i < H ? i : 0;
      x[k] = zs.rd(row + ic);
      y[k] = zs.rd(row + (C - 1 - ic));
    }
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      const int32_t ek = e + k * zs.vbs;
      if (ek >= n) break;
      const int32_t i = ii[k];
      double a = 0.0, b = 0.0;
      if (i < H) {
        if (2 * i + 1 == C) {
          a = x[k];
        } else {
          a = x[k] + y[k];
          b = x[k] - y[k];
        }
      }
      zs.wr(e0 + ek, a);
      zs.wr(o0 + ek, b);
    }
  }
}

// out[r][j] = sum_{i<ES} (j even ? E : O)[r][i] S[i][j] into the face vector
// of an LDS store. A wave takes 64 rows (one per lane) and JJA_LATTICE_CHUNK
// consecutive modes from an even j0, so a mode's parity is its slot's; the
// table's values are wave-uniform scalar loads. Each output adds its terms in
// ascending i with fused multiply-adds.
template <class Z>
__device__ __forceinline__ void lattice_transform_lds(const DevProblem& P, const Z& zs) {
  constexpr int MC = JJA_LATTICE_CHUNK;
  const int32_t R = P.lat_rows, C = P.lat_cols, ES = P.lat_es, CP = P.lat_cp;
  const int lane = zs.vt & (kWave - 1);
  const int32_t wave = __builtin_amdgcn_readfirstlane(zs.vt / kWave);
  const int32_t NW = zs.vbs / kWave;
  const int32_t nrb = (R + kWave - 1) / kWave, tasks = nrb * (CP / MC);
  const CAS double* __restrict__ S = (const CAS double*)P.lat_sine;
  for (int32_t task = wave; task < tasks; task += NW) {
    const int32_t rb = task % nrb, j0 = (task / nrb) * MC;
    const int32_t r = rb * kWave + lane;
    const int64_t e = P.lat_t0 + static_cast<int64_t>(r < R ? r : R - 1) * ES;
    const int64_t o = e + static_cast<int64_t>(R) * ES;
    double acc[MC];
#pragma unroll
    for (int q = 0; q < MC; ++q) acc[q] = 0.0;
    for (int32_t i = 0; i < ES; i += 2) {
      const vdouble2 ev = zs.rd2(e + i), ov = zs.rd2(o + i);
      const CAS double* s0 = S + static_cast<int64_t>(i) * CP + j0;
      const CAS double* s1 = s0 + CP;
#pragma unroll
      for (int q = 0; q < MC; ++q) {
        acc[q] = __builtin_fma((q & 1) ? ov.x : ev.x, s0[q], acc[q]);
        acc[q] = __builtin_fma((q & 1) ? ov.y : ev.y, s1[q], acc[q]);
      }
    }
    if (r < R) {
      const int64_t row = static_cast<int64_t>(r) * C;
#pragma unroll
      for (int q = 0; q < MC; ++q)
        if (j0 + q < C) zs.wr(row + j0 + q, acc[q]);
    }
  }
}

// The same sums for a global store, on the matrix cores: for each parity p,
// out[r][2m+p] = sum_i F_p[r][i] S[i][2m+p] (F_0 = E, F_1 = O) is a product
// of an R x ES and an ES x (modes of parity p) matrix. A wave takes a tile of
// 32 rows x 32 modes of one parity as 2 x 2 blocks of 16 x 16 and runs
// v_mfma_f64_16x16x4f64 over i in steps of 4; lane l holds A[i][k] =
// F[r0 + l%16][i0 + l/16] and B[k][j] = S[i0 + l/16][2(m0 + l%16) + p], and
// result register q of lane l is D[l/16 + 4q][l%16]. E and O are kept in
// blocks of four i by row, so each A load is 512 contiguous bytes; B comes from
// the table with even and odd modes apart (lat_sine_p), four 128-byte rows.
// kMfmaBatch steps of loads are issued before their products. Rows and modes
// past the lattice load clamped (finite) operands and are not stored; the
// folded rows are zero for i >= H, and table rows past ES are read clamped.
// The matrix cores round differently from a chain of fused multiply-adds: the
// two stores agree to rounding.
constexpr int kMfmaBatch = 4;
constexpr int kMfmaRowBlocks = 2;   // 16-row blocks per wave tile
constexpr int kMfmaModeBlocks = 2;  // 16-mode blocks per wave tile
template <class Z>
__device__ __forceinline__ void lattice_transform_mfma(const DevProblem& P, const Z& zs) {
  constexpr int U = kMfmaBatch, NR = kMfmaRowBlocks, NC = kMfmaModeBlocks;
  constexpr int TR = 16 * NR, TC = 16 * NC;  // tile rows, tile modes (of one parity)
  const int32_t R = P.lat_rows, C = P.lat_cols, ES = P.lat_es, EF = P.lat_ef, CP = P.lat_cp;
  const int lane = zs.vt & (kWave - 1);
  const int32_t wave = __builtin_amdgcn_readfirstlane(zs.vt / kWave);
  const int32_t NW = zs.vbs / kWave;
  const int32_t n_even = (C + 1) >> 1, n_odd = C >> 1;
  const int32_t nrt = (R + TR - 1) / TR, nte = (n_even + TC - 1) / TC, nto = (n_odd + TC - 1) / TC;
  const int32_t tasks = nrt * (nte + nto);
  const int lr = lane & 15, lk = lane >> 4;
  const GAS double* __restrict__ S = gbl(P.lat_sine_p);
  const int32_t half = CP >> 1;
  for (int32_t task = wave; task < tasks; task += NW) {
    const int32_t rt = task % nrt, mt = task / nrt;
    const int par = mt < nte ? 0 : 1;
    const int32_t m0 = TC * (par ? mt - nte : mt);
    const int32_t r0 = TR * rt;
    // operand addresses of this lane (clamped into the lattice)
    const GAS double* fa[NR];
    const GAS double* sb[NC];
#pragma unroll
    for (int b = 0; b < NR; ++b) {
      const int32_t r = min(r0 + 16 * b + lr, R - 1);
      fa[b] = zs.p + P.lat_t0 + static_cast<int64_t>(par) * EF * R + 4 * r + lk;
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) sb[c] = S + par * half + min(m0 + 16 * c + lr, half - 1);
    vdouble4 acc[NR][NC];
#pragma unroll
    for (int b = 0; b < NR; ++b)
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[b][c] = vdouble4{0.0, 0.0, 0.0, 0.0};
    for (int32_t i0 = 0; i0 < EF; i0 += 4 * U) {
      double av[U][NR], bv[U][NC];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t i4 = i0 + 4 * u;                     // this step's first i (a multiple of 4)
        const int32_t blk = min(i4, EF - 4) >> 2;
        const int32_t ic = min(i4 + lk, ES - 1);
#pragma unroll
        for (int b = 0; b < NR; ++b) {
          const double v = fa[b][static_cast<int64_t>(blk) * R * 4];
          av[u][b] = i4 < EF ? v : 0.0;
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) bv[u][c] = sb[c][static_cast<int64_t>(ic) * CP];
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int b = 0; b < NR; ++b)
#pragma unroll
          for (int c = 0; c < NC; ++c)
            acc[b][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u][b], bv[u][c], acc[b][c], 0, 0, 0);
    }
#pragma unroll
    for (int b = 0; b < NR; ++b)
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int32_t j = 2 * (m0 + 16 * c + lr) + par;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int32_t r = r0 + 16 * b + lk + 4 * q;
          if (r < R && j < C) zs.wr(static_cast<int64_t>(r) * C + j, acc[b][c][q]);
        }
      }
  }
}

template <class Z>
__device__ __forceinline__ void lattice_transform(const DevProblem& P, const Z& zs) {
  if constexpr (Z::kGlobal)
    lattice_transform_mfma(P, zs);
  else
    lattice_transform_lds(P, zs);
}

// Mode k's tridiagonal solve along r, in place in the face vector. Both
// sweeps are first-order linear recurrences with the multiplier
// a_r = g inv_r: forward d_r = (y_r + g d_{r-1}) inv_r = y_r inv_r + a_r d_{r-1},
// backward x_r = d_r - ((-g) inv_r) x_{r+1} = d_r + a_r x_{r+1}. The rows of a
// mode are cut into B blocks (B from R, C and the part count only, so results
// do not depend on the block size): each block runs its recurrence from a
// zero start while keeping the running product of its multipliers, and then
// adds its product times the value carried in from the blocks before it. A
// block's thread chains those blocks' (value, product) ends itself, in block
// order, so each carry is the same sum whoever forms it, and no phase or
// barrier is spent on a separate chain. Three phases: forward from zero;
// forward carries, then backward from zero on the same rows; backward carries.
constexpr int kTriRows = 8;  // rows (or block ends) held in registers at a time
// blocks per mode: JJA_LATTICE_TRI_THREADS per part of the column, over C
__device__ __forceinline__ int32_t lattice_blocks(const DevProblem& P) {
  return max(1, min(P.lat_rows, JJA_LATTICE_TRI_THREADS * P.parts / P.lat_cols));
}

// The value carried into block b of mode k: dir +1 chains the ends of blocks
// 0 .. b-1 upward (cin = endp cin + endv from 0), dir -1 the ends of blocks
// B-1 .. b+1 downward.
template <class Z>
__device__ __forceinline__ double lattice_carry(const Z& zs, int32_t C, int32_t B, int64_t endv, int64_t endp,
                                                int32_t b, int32_t k, int dir) {
  constexpr int TB = kTriRows;
  const int32_t n = dir > 0 ? b : B - 1 - b, first = dir > 0 ? 0 : B - 1;
  double cin = 0.0;
  for (int32_t i0 = 0; i0 < n; i0 += TB) {
    double ev[TB], ep[TB];
#pragma unroll
    for (int q = 0; q < TB; ++q) {
      const int64_t at = static_cast<int64_t>(first + dir * min(i0 + q, n - 1)) * C + k;
      ev[q] = zs.rd(endv + at);
      ep[q] = zs.rd(endp + at);
    }
#pragma unroll
    for (int q = 0; q < TB; ++q)
      if (i0 + q < n) cin = __builtin_fma(ep[q], cin, ev[q]);
  }
  return cin;
}

// The tridiagonal solves of a split column with each part's modes in its own
// LDS, followed by the second fold. Part p takes the mode pairs (i, C-1-i) for
// i in [p*MH, p*MH + MH) (MH = lat_mp / 2; the middle mode of an odd C has no
// partner) with all R rows (Y[r][slot]: the lower modes in slots 0..nh-1, their
// partners after them; R*lat_mp doubles), cuts each mode into BP = lat_bp
// blocks and runs the same blocked recurrences as lattice_tridiag between
// workgroup barriers; the block ends (forward and backward, value and product,
// 4*BP*lat_mp doubles) stay in LDS, so no column barrier is needed inside.
// Holding both modes of a pair, the part then writes their folded rows
// (E = x_i + x_{C-1-i}, O = x_i - x_{C-1-i}, the same sums as lattice_fold) in
// the global store's blocked layout, in place of a separate fold pass.
__device__ __forceinline__ void lattice_tridiag_parts(const DevProblem& P, const SplitStore& zs) {
  constexpr int TB = kTriRows;
  const int32_t R = P.lat_rows, C = P.lat_cols, CP = P.lat_cp, MP = P.lat_mp, B = P.lat_bp;
  const int32_t H = (C + 1) >> 1, EF = P.lat_ef, MH = MP >> 1;
  const int32_t nt = static_cast<int32_t>(blockDim.x), tid = static_cast<int32_t>(threadIdx.x);
  const int32_t part = __builtin_amdgcn_readfirstlane(zs.vt / nt);
  const int64_t e0 = P.lat_t0, o0 = P.lat_t0 + static_cast<int64_t>(R) * EF;
  // folded-row padding (H <= i < EF) is zero; part 0 writes it
  if (part == 0)
    for (int32_t e = tid; e < R * (EF - H); e += nt) {
      const int32_t i = H + e / R, r = e - (e / R) * R;
      const int64_t at = (static_cast<int64_t>(i >> 2) * R + r) * 4 + (i & 3);
      zs.wr(e0 + at, 0.0);
      zs.wr(o0 + at, 0.0);
    }
  const int32_t i0 = part * MH, nh = min(MH, H - i0);
  if (nh <= 0) return;  // a part past the last pair (the whole workgroup)
  const int32_t nu = (2 * (i0 + nh - 1) + 1 == C) ? nh - 1 : nh;  // partners (the middle has none)
  const int32_t nm = nh + nu;
  auto mode = [&](int32_t kk) { return kk < nh ? i0 + kk : C - 1 - (i0 + kk - nh); };
  const GAS double* __restrict__ inv = gbl(P.lat_inv);
  const double g = P.lat_g;
  const int32_t m = (R + B - 1) / B;
  LAS double* Y = zs.lds;
  LAS double* fv = Y + static_cast<int32_t>(R) * MP;
  LAS double* fp = fv + B * MP;
  LAS double* bv = fp + B * MP;
  LAS double* bp = bv + B * MP;
  for (int32_t e = tid; e < R * nm; e += nt) {
    const int32_t r = e / nm, kk = e - r * nm;
    Y[r * MP + kk] = zs.rd(static_cast<int64_t>(r) * C + mode(kk));
  }
  __syncthreads();
  auto sweep = [&](int32_t kk, int32_t first, int32_t count, int32_t dir, auto&& f) {
    const int32_t k = mode(kk);
    for (int32_t c0 = 0; c0 < count; c0 += TB) {
      double y[TB], a[TB];
#pragma unroll
      for (int q = 0; q < TB; ++q) {
        const int32_t r = first + dir * min(c0 + q, count - 1);
        y[q] = Y[r * MP + kk];
        a[q] = inv[static_cast<int64_t>(r) * CP + k];
      }
#pragma unroll
      for (int q = 0; q < TB; ++q)
        if (c0 + q < count) f(first + dir * (c0 + q), y[q], a[q]);
    }
  };
  // the value carried into block b (dir +1: from blocks 0..b-1, upward; dir
  // -1: from blocks B-1..b+1, downward), chained in block order
  auto carry = [&](const LAS double* ev, const LAS double* ep, int32_t b, int32_t kk, int dir) {
    const int32_t n = dir > 0 ? b : B - 1 - b, first = dir > 0 ? 0 : B - 1;
    double cin = 0.0;
    for (int32_t i = 0; i < n; ++i) {
      const int32_t at = (first + dir * i) * MP + kk;
      cin = __builtin_fma(ep[at], cin, ev[at]);
    }
    return cin;
  };
  // forward from zero
  for (int32_t e = tid; e < B * nm; e += nt) {
    const int32_t b = e / nm, kk = e - b * nm, r0 = b * m, n = min(m, R - r0);
    double d = 0.0, pr = 1.0;
    if (n > 0)
      sweep(kk, r0, n, 1, [&](int32_t r, double y, double iv) {
        d = __builtin_fma(g, d, y) * iv;
        pr = pr * (g * iv);
        Y[r * MP + kk] = d;
      });
    fv[b * MP + kk] = d;
    fp[b * MP + kk] = pr;
  }
  __syncthreads();
  // forward carries, then backward from zero (the same thread's rows)
  for (int32_t e = tid; e < B * nm; e += nt) {
    const int32_t b = e / nm, kk = e - b * nm, r0 = b * m, n = min(m, R - r0);
    double x = 0.0, pr = 1.0;
    if (n > 0) {
      if (b > 0) {
        const double cin = carry(fv, fp, b, kk, 1);
        double pf = 1.0;
        sweep(kk, r0, n, 1, [&](int32_t r, double y, double iv) {
          pf = pf * (g * iv);
          Y[r * MP + kk] = __builtin_fma(pf, cin, y);
        });
      }
      sweep(kk, r0 + n - 1, n, -1, [&](int32_t r, double dv, double iv) {
        const double a = g * iv;
        x = __builtin_fma(a, x, dv);
        pr = pr * a;
        Y[r * MP + kk] = x;
      });
    }
    bv[b * MP + kk] = x;
    bp[b * MP + kk] = pr;
  }
  __syncthreads();
  // backward carries (the last block has none)
  for (int32_t e = tid; e < B * nm; e += nt) {
    const int32_t b = e / nm, kk = e - b * nm, r0 = b * m, n = min(m, R - r0);
    if (n <= 0 || b == B - 1) continue;
    const double cin = carry(bv, bp, b, kk, -1);
    double pr = 1.0;
    sweep(kk, r0 + n - 1, n, -1, [&](int32_t r, double x, double iv) {
      pr = pr * (g * iv);
      Y[r * MP + kk] = __builtin_fma(pr, cin, x);
    });
  }
  __syncthreads();
  // the pairs' folded rows into the global store's blocked layout
  for (int32_t e = tid; e < R * nh; e += nt) {
    const int32_t r = e / nh, j = e - r * nh, i = i0 + j;
    const double x = Y[r * MP + j];
    double a = x, b = 0.0;
    if (2 * i + 1 != C) {
      const double y = Y[r * MP + nh + j];
      a = x + y;
      b = x - y;
    }
    const int64_t at = (static_cast<int64_t>(i >> 2) * R + r) * 4 + (i & 3);
    zs.wr(e0 + at, a);
    zs.wr(o0 + at, b);
  }
}

template <class Z>
__device__ __forceinline__ void lattice_tridiag(const DevProblem& P, const Z& zs) {
  if constexpr (std::is_same<Z, SplitStore>::value) {
    if (P.lat_mp) {
      lattice_tridiag_parts(P, zs);
      return;
    }
  }
  constexpr int TB = kTriRows;
  const int32_t R = P.lat_rows, C = P.lat_cols, CP = P.lat_cp;
  const GAS double* __restrict__ inv = gbl(P.lat_inv);
  const double g = P.lat_g;
  const int32_t B = lattice_blocks(P), m = (R + B - 1) / B;
  // forward and backward block ends (value, product) in the free folded-row
  // area: the second phase reads forward ends while other blocks write their
  // backward ends
  const int64_t fv = P.lat_t0, fp = fv + static_cast<int64_t>(B) * C;
  const int64_t bv = fp + static_cast<int64_t>(B) * C, bp = bv + static_cast<int64_t>(B) * C;
  auto zi = [&](int32_t r, int32_t k) { return static_cast<int64_t>(r) * C + k; };
  // one pass over `count` rows of mode k from `first` in direction dir (+1
  // up, -1 down): f(r, y, a) for each row, rows loaded TB at a time
  auto sweep = [&](int32_t k, int32_t first, int32_t count, int32_t dir, auto&& f) {
    for (int32_t c0 = 0; c0 < count; c0 += TB) {
      double y[TB], a[TB];
#pragma unroll
      for (int q = 0; q < TB; ++q) {
        const int32_t r = first + dir * min(c0 + q, count - 1);
        y[q] = zs.rd(zi(r, k));
        a[q] = inv[static_cast<int64_t>(r) * CP + k];
      }
#pragma unroll
      for (int q = 0; q < TB; ++q)
        if (c0 + q < count) f(first + dir * (c0 + q), y[q], a[q]);
    }
  };
  // forward from zero (an empty block at the end passes its carry through:
  // value 0, product 1)
  for (int32_t e = zs.vt; e < B * C; e += zs.vbs) {
    const int32_t b = e / C, k = e - b * C, r0 = b * m, n = min(m, R - r0);
    double d = 0.0, pr = 1.0;
    if (n > 0)
      sweep(k, r0, n, 1, [&](int32_t r, double y, double iv) {
        d = __builtin_fma(g, d, y) * iv;
        pr = pr * (g * iv);
        zs.wr(zi(r, k), d);
      });
    zs.wr(fv + e, d);
    zs.wr(fp + e, pr);
  }
  zs.barrier();
  // forward carries into the rows, then backward from zero, top row down (the
  // same thread wrote the rows it reads back)
  for (int32_t e = zs.vt; e < B * C; e += zs.vbs) {
    const int32_t b = e / C, k = e - b * C, r0 = b * m, n = min(m, R - r0);
    double x = 0.0, pr = 1.0;
    if (n > 0) {
      if (b > 0) {
        const double cin = lattice_carry(zs, C, B, fv, fp, b, k, 1);
        double pf = 1.0;
        sweep(k, r0, n, 1, [&](int32_t r, double y, double iv) {
          pf = pf * (g * iv);
          zs.wr(zi(r, k), __builtin_fma(pf, cin, y));
        });
      }
      sweep(k, r0 + n - 1, n, -1, [&](int32_t r, double dv, double iv) {
        const double a = g * iv;
        x = __builtin_fma(a, x, dv);
        pr = pr * a;
        zs.wr(zi(r, k), x);
      });
    }
    zs.wr(bv + e, x);
    zs.wr(bp + e, pr);
  }
  zs.barrier();
  // backward carries into the rows
  for (int32_t e = zs.vt; e < B * C; e += zs.vbs) {
    const int32_t b = e / C, k = e - b * C, r0 = b * m, n = min(m, R - r0);
    if (n <= 0 || b == B - 1) continue;
    const double cin = lattice_carry(zs, C, B, bv, bp, b, k, -1);
    double pr = 1.0;
    sweep(k, r0 + n - 1, n, -1, [&](int32_t r, double x, double iv) {
      pr = pr * (g * iv);
      zs.wr(zi(r, k), __builtin_fma(pr, cin, x));
    });
  }
}

template <class Z>
__device__ __forceinline__ void solve_lattice(const DevProblem& P, const Z& zs) {
  // a split column's part-local tridiagonal solve also writes the second fold
  bool folded = false;
  if constexpr (std::is_same<Z, SplitStore>::value) folded = P.lat_mp != 0;
  lattice_fold(P, zs);
  zs.barrier();
  lattice_transform(P, zs);
  zs.barrier();
  lattice_tridiag(P, zs);
  zs.barrier();
  if (!folded) {
    lattice_fold(P, zs);
    zs.barrier();
  }
  lattice_transform(P, zs);
  zs.barrier();
}

// K^-1 b of the face vector in place: the lattice solve or the factor program.
template <class Z>
__device__ __forceinline__ void solve_face_system(const DevProblem& P, const Z& zs) {
  if constexpr (RunsLattice<Z>::value) {
    if (P.lat_rows) {
      solve_lattice(P, zs);
      return;
    }
  }
  solve(P, zs);
}

// Two columns of one problem through the same record (their face stores za and
// zb): each column's gathers, sums and writes are exactly those of
// process_slice, so each column's results are bit-identical to a solve of its
// own; the record's values and indices are loaded once for both.
template <class Z>
__device__ __forceinline__ void process_slice_pair(const Z& za, const Z& zb, int32_t Nf, int lane,
                                                   const jja_slice& m, const SliceRegs<Z::kGlobal>& r,
                                                   double& acca, double& accb) {
  double va[kMaxIter], vb[kMaxIter];
#pragma unroll
  for (int i = 0; i < kMaxIter; ++i) {
    const uint32_t ix = record_index(r, m.flags, i);
    va[i] = za.gather(ix);
    vb[i] = zb.gather(ix);
  }
#pragma unroll
  for (int i = 0; i < kMaxIter; ++i) {
    acca = __builtin_fma(r.v[i], va[i], acca);
    accb = __builtin_fma(r.v[i], vb[i], accb);
  }
  if (m.flags & JJA_SLICE_FINISH) {
    const int G = m.flags & 0xff;
    acca = group_sum_last(acca, G);
    accb = group_sum_last(accb, G);
    const int slot = lane >> __builtin_ctz(static_cast<unsigned>(G));
    if (slot < (m.flags >> 16) && r.row >= 0 && (lane & (G - 1)) == G - 1) {
      if (m.flags & JJA_SLICE_BLOCK) {
        za.wr(r.row, acca);
        zb.wr(r.row, accb);
      } else {
        za.wr(Nf + r.row - m.r0, za.rd(r.row) - acca);
        zb.wr(Nf + r.row - m.r0, zb.rd(r.row) - accb);
      }
    }
    acca = 0.0;
    accb = 0.0;
  }
  if (m.flags & JJA_SLICE_BARRIER)
    za.barrier();
  else if (m.flags & JJA_SLICE_WG_BARRIER)
    za.wg_barrier();
}

// solve() for two columns that share one set of threads and one barrier (za's).
template <class Z>
__device__ __forceinline__ void solve_pair(const DevProblem& P, const Z& za, const Z& zb) {
  constexpr bool GLOBAL = Z::kGlobal;
  const int lane = za.vt & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(za.vt / kWave);
  const CAS int32_t* wp = (const CAS int32_t*)(P.wave_ptr) + wave;
  const int32_t k0 = __builtin_amdgcn_readfirstlane(wp[0]);
  const int32_t nk = __builtin_amdgcn_readfirstlane(wp[1]) - k0 - 4;
  const jja_slice* __restrict__ prog = P.solve_slices + k0;
  jja_slice m0 = load_rec(prog, 0), m1 = load_rec(prog, 1), m2 = load_rec(prog, 2);
  const __amdgpu_buffer_rsrc_t vres = value_resource(P);
  SliceRegs<GLOBAL> A, B, C;
  load_slice(P, vres, m0, lane, A);
  load_slice(P, vres, m1, lane, B);
  double acca = 0.0, accb = 0.0;
  for (int32_t k = 0; k < nk; k += 3) {
    const jja_slice m3 = load_rec(prog, k + 3);
    load_slice(P, vres, m2, lane, C);
    __builtin_amdgcn_sched_barrier(0);
    process_slice_pair(za, zb, P.Nf, lane, m0, A, acca, accb);
    const jja_slice m4 = load_rec(prog, k + 4);
    load_slice(P, vres, m3, lane, A);
    __builtin_amdgcn_sched_barrier(0);
    process_slice_pair(za, zb, P.Nf, lane, m1, B, acca, accb);
    const jja_slice m5 = load_rec(prog, k + 5);
    load_slice(P, vres, m4, lane, B);
    __builtin_amdgcn_sched_barrier(0);
    process_slice_pair(za, zb, P.Nf, lane, m2, C, acca, accb);
    m0 = m3;
    m1 = m4;
    m2 = m5;
  }
}

// One input term of one column at one step: element i is p[i * si] * amp.
// amp is exactly 1 for a term without amplitude; a term that is absent reads
// an exact +0.0 from the plan's zero page (si = 0), which leaves every sum it
// enters bit-identical to skipping it.
struct TermView {
  const GAS double* p;
  int64_t si;
  double amp;
};

__device__ __forceinline__ TermView term_view(const jja_source& s, bool present, const double* zero,
                                              int64_t w, int64_t t) {
  TermView v{gbl(zero), 0, 1.0};
  if (present && s.base) {
    const int64_t tt = t - s.t0;
    v.p = gbl(s.base) + w * s.s_w + tt * s.s_t;
    v.si = s.s_i;
    if (s.amp) v.amp = gbl(s.amp)[w * s.a_w + tt * s.a_t];
  }
  return v;
}

__device__ __forceinline__ double term_at(const TermView& v, int32_t i) {
  return v.p[static_cast<int64_t>(i) * v.si] * v.amp;
}

// Face phase: b_p = sum +-u_j in ascending j, minus 2 pi f (time_evolution.py:560-569).
// A face longer than one record continues in the next record (flag FINISH on
// its last one). Records go in groups of kFaceGroup: the code, row and face
// loads of the group are issued together, then its u gathers and flux values,
// so a group costs two memory round trips.
struct FaceRec {
  int32_t cd[kFaceIter];  // +-junction codes, -1 past the face's length
  int32_t pos;            // solve position written at FINISH, or -1
  int32_t fid;            // face index of that position (flux input), or -1
  bool finish;
};

__device__ __forceinline__ void face_rec_load(const DevProblem& P, int32_t k, bool present, int tid,
                                              int32_t vbs, FaceRec& r) {
  const jja_slice m = load_rec(P.face_slices, k);
  const bool live = present && tid < (m.flags >> 16);
  const GAS int32_t* __restrict__ code = gbl(P.face_code) + m.base + tid;
#pragma unroll
  for (int i = 0; i < kFaceIter; ++i) r.cd[i] = (live && i < m.n_iter) ? code[i * vbs] : -1;
  r.finish = present && (m.flags & JJA_SLICE_FINISH);
  const bool rowed = live && r.finish;
  r.pos = rowed ? gbl(P.slice_rows)[m.rowbase + tid] : -1;
  r.fid = rowed ? gbl(P.face_fid)[m.rowbase + tid] : -1;
}

// The junction values u_j the face phase sums: u itself (an LDS store's
// evolution, whose junction pass writes u, and the adjoint), or, for a global
// store, formed from the step's x as the junction pass would have: x_j / c0 (a
// product with the reciprocal), less theta_s with voltage sources; one write
// stream fewer per junction. rc0u is 1/c0 when every junction has the same
// (read once).
struct JuncSource {
  const GAS double* v;
  const GAS double* ths;
  bool scaled;
  double rc0u;
};

// MODE 0: u itself; 1: x / c0 with one c0; 2: x / c0 per junction; 3: x / c0
// per junction less theta_s
template <int MODE>
__device__ __forceinline__ double junc_value(const DevProblem& P, const JuncSource& s, int32_t j) {
  if constexpr (MODE == 0) return s.v[j];
  if constexpr (MODE == 1) return s.v[j] * s.rc0u;
  if constexpr (MODE == 2) return s.v[j] * rc0_at(P, j);
  return s.v[j] * rc0_at(P, j) - s.ths[j];
}

template <int BS, int MODE, class Z>
__device__ __forceinline__ void face_phase_impl(const DevProblem& P, const jja_source* f, const JuncSource& us,
                                                const Z& zs, int64_t w, int64_t t) {
  const int tid = zs.vt;
  // flux of face q at this step is fv.p[q * fv.si] * fv.amp; an absent flux
  // reads an exact zero, and acc - 2 pi * 0 leaves acc unchanged
  const TermView fv = term_view(f ? *f : jja_source{}, f && !f->zero, P.zero, w, t);
  double acc = 0.0;
  const int32_t n = P.n_face_slices;
  for (int32_t k = 0; k < n; k += kFaceGroup) {
    FaceRec r[kFaceGroup];
#pragma unroll
    for (int g = 0; g < kFaceGroup; ++g)
      face_rec_load(P, k + g < n ? k + g : k, k + g < n, tid, zs.vbs, r[g]);
    double uu[kFaceGroup][kFaceIter], fq[kFaceGroup];
#pragma unroll
    for (int g = 0; g < kFaceGroup; ++g) {
#pragma unroll
      for (int i = 0; i < kFaceIter; ++i)
        uu[g][i] = r[g].cd[i] >= 0 ? junc_value<MODE>(P, us, r[g].cd[i] >> 1) : 0.0;
      fq[g] = r[g].fid >= 0 ? term_at(fv, r[g].fid) : 0.0;
    }
#pragma unroll
    for (int g = 0; g < kFaceGroup; ++g) {
#pragma unroll
      for (int i = 0; i < kFaceIter; ++i)
        if (r[g].cd[i] >= 0) acc = (r[g].cd[i] & 1) ? acc - uu[g][i] : acc + uu[g][i];
      if (r[g].finish) {
        if (r[g].pos >= 0) {
          acc = acc - kTwoPi * fq[g];
          zs.wr(r[g].pos, acc);
          if (r[g].pos < P.copy_rows) zs.wr(P.Nf + r[g].pos, acc);  // first stage: scratch = b
        }
        acc = 0.0;
      }
    }
  }
}

template <int BS, class Z>
__device__ __forceinline__ void face_phase(const DevProblem& P, const jja_source* f, JuncSource us, const Z& zs,
                                           int64_t w, int64_t t) {
  if (!us.scaled) {
    face_phase_impl<BS, 0>(P, f, us, zs, w, t);
  } else if (us.ths) {
    face_phase_impl<BS, 3>(P, f, us, zs, w, t);
  } else if (P.coef_si == 0) {
    us.rc0u = rc0_at(P, 0);
    face_phase_impl<BS, 1>(P, f, us, zs, w, t);
  } else {
    face_phase_impl<BS, 2>(P, f, us, zs, w, t);
  }
}

// Everything one junction pass needs that is the same for all its junctions.
struct JunctionStep {
  GAS double* th_n;            // slot of step t: holds x_t, receives theta_t
  const GAS double* th_k[4];   // theta_{tx-1-k}, k = 0..S-1 (k = 0 unused with UPDATE)
  GAS double* th_x;            // slot of step tx: receives x_tx
  GAS double* ths;             // voltage-source phase (NULL without Vs)
  GAS double* u;               // LDS store: receives u = x_tx/c0 - theta_s (global store: NULL)
  GAS double* th_out;          // stored theta_t row or NULL
  GAS double* I_out;           // stored current row or NULL
  TermView Is_t, Vs_t;         // step t (update)
  TermView Is_x, sig_x;        // step tx (prep); sig_x is sigma, or T when from_T
  const GAS double* noise;     // reference normals of step tx, or NULL
  const GAS double* Ic;        // critical currents (the run's, else the plan's)
  bool is_zero, t_zero, from_T;
  double dt;
  uint64_t seed;
  uint32_t noise_col;
  uint64_t tx;
  int32_t vt, vbs;             // the thread's index in the column, the column's thread count
};

// ---------------------------------------------------------------- general
// One junction, all options: with UPDATE finish step t (time_evolution.py:
// 560-580): y = A^T J from the face vector, theta_t = (y - x)/c0, stored
// outputs, theta_s += Vs dt; with PREP form x of step tx (time_evolution.py:
// 533-558) and write x to the history slot, and for an LDS store u = x/c0 -
// theta_s to the scratch (a global store's face phase forms u from x). All
// loads are issued before any arithmetic.
template <int SW, bool UPDATE, bool PREP, class Z>
__device__ __forceinline__ void junction_one(const DevProblem& P, const JunctionStep& J,
                                             const Z& zs, int32_t j, double zj) {
  const double rc0 = rc0_at(P, j);
  vint2 code{-1, -1};
  double xo = 0.0, isn = 0.0, thsv = 0.0, vsv = 0.0;
  if (UPDATE) {
    code = gbl(reinterpret_cast<const vint2*>(P.junc_code))[j];
    xo = J.th_n[j];
    if (J.I_out) isn = term_at(J.Is_t, j);
    if (J.ths) vsv = term_at(J.Vs_t, j);
  }
  if (J.ths) thsv = J.ths[j];
  double t[4] = {0.0, 0.0, 0.0, 0.0}, c[SW], Ic = 0.0, isx = 0.0, sg = 0.0, rv = 0.0;
  if (PREP) {
#pragma unroll
    for (int k = UPDATE ? 1 : 0; k < SW - 1; ++k) t[k] = J.th_k[k][j];
    Ic = J.Ic[j];
#pragma unroll
    for (int k = 1; k < SW; ++k) c[k] = coef_at(P, k, j);
    isx = term_at(J.Is_x, j);
    sg = term_at(J.sig_x, j);
    if (J.from_T) rv = gbl(P.Rv)[j];
    if (J.noise) zj = J.noise[j];
  }
  if (UPDATE) {
    double y = 0.0;
    if (code.x >= 0) y = (code.x & 1) ? y - zs.rd(code.x >> 1) : y + zs.rd(code.x >> 1);
    if (code.y >= 0) y = (code.y & 1) ? y - zs.rd(code.y >> 1) : y + zs.rd(code.y >> 1);
    t[0] = (y - xo) * rc0;
    J.th_n[j] = t[0];
    if (J.th_out) J.th_out[j] = t[0];
    if (J.I_out) J.I_out[j] = y + isn;
    if (J.ths) {
      thsv += vsv * J.dt;
      J.ths[j] = thsv;
    }
  }
  if (PREP) {
    double X;
    if constexpr (SW == 3) {
      X = Ic * cpr_sum(P, 2.0 * t[0] - t[1]) + c[1] * t[0] + c[2] * t[1];
    } else if constexpr (SW == 4) {
      X = Ic * cpr_sum(P, 3.0 * t[0] - 3.0 * t[1] + t[2]) + c[1] * t[0] + c[2] * t[1] + c[3] * t[2];
    } else {
      X = Ic * cpr_sum(P, 4.0 * t[0] - 6.0 * t[1] + 4.0 * t[2] - t[3]) + c[1] * t[0] + c[2] * t[1] +
          c[3] * t[2] + c[4] * t[3];
    }
    double eta = 0.0;
    if (!J.t_zero) eta = (J.from_T ? sqrt(2.0 * sg * rv) : sg) * zj;
    const double x = J.is_zero ? eta + X : (eta - isx) + X;
    J.th_x[j] = x;
    if constexpr (!Z::kGlobal) {
      double uj = x * rc0;
      if (J.ths) uj = uj - thsv;
      J.u[j] = uj;
    }
  }
}

// Deviate of this lane's junction from a Philox pair computed by another lane
// of the wave: lane `src` holds (z0, z1) for junctions (2i, 2i+1); `second`
// picks z1.
__device__ __forceinline__ double take_normal(double z0, double z1, int src, bool second) {
  const long long b0 = __builtin_bit_cast(long long, z0), b1 = __builtin_bit_cast(long long, z1);
  const int a = src << 2;
  const int lo0 = __builtin_amdgcn_ds_bpermute(a, static_cast<int>(b0));
  const int hi0 = __builtin_amdgcn_ds_bpermute(a, static_cast<int>(b0 >> 32));
  const int lo1 = __builtin_amdgcn_ds_bpermute(a, static_cast<int>(b1));
  const int hi1 = __builtin_amdgcn_ds_bpermute(a, static_cast<int>(b1 >> 32));
  const int lo = second ? lo1 : lo0, hi = second ? hi1 : hi0;
  return __builtin_bit_cast(double, static_cast<long long>(static_cast<unsigned>(lo)) |
                                        (static_cast<long long>(hi) << 32));
}

// Normals for the junctions k*BS + tid and (k+1)*BS + tid of this lane. Each
// Philox pair (junctions 2i, 2i+1) is drawn once: lanes 0..31 of a wave draw
// the 32 pairs its 64 junctions of iteration k form, lanes 32..63 those of
// iteration k+1, and every lane then fetches its own deviate. All lanes of the
// wave take part, whether or not their junctions exist.
template <int BS>
__device__ __forceinline__ void wave_normals(const JunctionStep& J, int32_t k, double& zA, double& zB) {
  const int lane = J.vt & 63;
  const int32_t wave_base = J.vt - lane;
  const int32_t it = lane < 32 ? k : k + 1;
  const int32_t j_even = it * J.vbs + wave_base + 2 * (lane & 31);
  double z0, z1;
  normal_pair(J.seed, static_cast<uint32_t>(j_even), J.noise_col, J.tx, &z0, &z1);
  zA = take_normal(z0, z1, lane >> 1, lane & 1);
  zB = take_normal(z0, z1, 32 + (lane >> 1), lane & 1);
}

// The junction pass, general form: thread tid takes junctions k*BS + tid
// (consecutive lanes, consecutive junctions), two iterations at a time.
template <int BS, int SW, bool UPDATE, bool PREP, class Z>
__device__ __forceinline__ void junction_loop(const DevProblem& P, const JunctionStep& J, const Z& zs) {
  const int32_t Nj = P.Nj;
  const int32_t tid = J.vt, vbs = J.vbs;
  const bool gauss = PREP && !J.t_zero && !J.noise;
  for (int32_t k = 0; k * vbs < Nj; k += 2) {
    double zA = 0.0, zB = 0.0;
    if (gauss) wave_normals<BS>(J, k, zA, zB);
    const int32_t jA = k * vbs + tid, jB = jA + vbs;
    if (jA < Nj) junction_one<SW, UPDATE, PREP, Z>(P, J, zs, jA, zA);
    // keep the two junctions apart: interleaving them would double the live
    // registers and spill
    __builtin_amdgcn_sched_barrier(0);
    if (jB < Nj) junction_one<SW, UPDATE, PREP, Z>(P, J, zs, jB, zB);
  }
}

// ------------------------------------------------------------ common step
// The usual step (stencil 3, no voltage sources, nothing stored at step t,
// noise amplitude given per junction, Philox or no noise): the same
// arithmetic as junction_one, each lane taking a junction pair, with the loads
// of both junctions issued first and the pair's Philox draw computed while
// they are in flight. (Loading the next pair's face codes an iteration ahead,
// so that the face-store gathers could go out with the other loads, made the
// split kernel spill and was 3.6% slower on SquareArray(512,512).)
struct JFast {
  vint2 code;
  double rc0, xo, t2, c1, c2, Ic, isx, sg;
};

// UNI: 1/c0, c1, c2 and sigma are one value each (uniform R and C, constant
// T): read once before the loop (scalar registers), not per junction
template <bool UNI>
__device__ __forceinline__ void jfast_load(const DevProblem& P, const JunctionStep& J, int32_t j, JFast& L) {
  L.code = gbl(reinterpret_cast<const vint2*>(P.junc_code))[j];
  L.xo = J.th_n[j];
  L.t2 = J.th_k[1][j];
  L.Ic = J.Ic[j];
  L.isx = term_at(J.Is_x, j);
  if constexpr (!UNI) {
    L.rc0 = rc0_at(P, j);
    L.c1 = coef_at(P, 1, j);
    L.c2 = coef_at(P, 2, j);
    L.sg = term_at(J.sig_x, j);
  }
}

template <class Z>
__device__ __forceinline__ void jfast_compute(const DevProblem& P, const JunctionStep& J, const Z& zs, int32_t j,
                                              const JFast& L, double z) {
  double y = 0.0;
  if (L.code.x >= 0) y = (L.code.x & 1) ? y - zs.rd(L.code.x >> 1) : y + zs.rd(L.code.x >> 1);
  if (L.code.y >= 0) y = (L.code.y & 1) ? y - zs.rd(L.code.y >> 1) : y + zs.rd(L.code.y >> 1);
  const double t1 = (y - L.xo) * L.rc0;
  J.th_n[j] = t1;
  const double X = L.Ic * cpr_sum(P, 2.0 * t1 - L.t2) + L.c1 * t1 + L.c2 * L.t2;
  double eta = 0.0;
  if (!J.t_zero) eta = L.sg * z;
  const double x = J.is_zero ? eta + X : (eta - L.isx) + X;
  J.th_x[j] = x;
  if constexpr (!Z::kGlobal) J.u[j] = x * L.rc0;
}

template <int BS, bool UNI, class Z>
__device__ __forceinline__ void junction_loop_fast(const DevProblem& P, const JunctionStep& J, const Z& zs) {
  const int32_t Nj = P.Nj;
  double u_rc0 = 0.0, u_c1 = 0.0, u_c2 = 0.0, u_sg = 0.0;
  if constexpr (UNI) {
    u_rc0 = rc0_at(P, 0);
    u_c1 = coef_at(P, 1, 0);
    u_c2 = coef_at(P, 2, 0);
    u_sg = term_at(J.sig_x, 0);
  }
  const int32_t tid = J.vt, vbs = J.vbs;
  const int32_t last = Nj - 1;
  const bool gauss = !J.t_zero;
  // a lane takes the junction pair (2q, 2q+1): its own Philox draw gives both
  // deviates (the pair's counter, as in the general loop); a pair past the
  // end (odd Nj) reads the last junction and does not use it
  for (int32_t q = tid; 2 * q < Nj; q += vbs) {
    const int32_t jA = 2 * q, jB = jA + 1;
    JFast A, B;
    jfast_load<UNI>(P, J, jA, A);
    jfast_load<UNI>(P, J, min(jB, last), B);
    if constexpr (UNI) {
      A.rc0 = B.rc0 = u_rc0;
      A.c1 = B.c1 = u_c1;
      A.c2 = B.c2 = u_c2;
      A.sg = B.sg = u_sg;
    }
    __builtin_amdgcn_sched_barrier(0);
    // the Philox draw runs while the loads are in flight
    double zA = 0.0, zB = 0.0;
    if (gauss) normal_pair(J.seed, static_cast<uint32_t>(jA), J.noise_col, J.tx, &zA, &zB);
    jfast_compute(P, J, zs, jA, A, zA);
    if (jB < Nj) jfast_compute(P, J, zs, jB, B, zB);
  }
}

// Set up and run one junction pass (finish step t if `update`, prepare step tx
// if `prep`) for column w.
template <int BS, int SW, class Z>
__device__ __forceinline__ void junction_pass(const DevProblem& P, const jja_problem_run& R,
                                              GAS double* __restrict__ th, GAS double* __restrict__ ths,
                                              GAS double* __restrict__ u, const Z& zs,
                                              int64_t w, size_t slot_stride, int64_t t, bool update,
                                              int64_t tx, bool prep, uint32_t noise_col) {
  constexpr int32_t S = SW - 1;
  auto slot = [&](int64_t m) -> size_t {
    int64_t r = m % S;
    return static_cast<size_t>(r < 0 ? r + S : r) * slot_stride;
  };
  JunctionStep J;
  J.vt = zs.vt;
  J.vbs = zs.vbs;
  J.th_n = th + slot(t);
#pragma unroll
  for (int k = 0; k < 4; ++k) J.th_k[k] = th + (k < S ? slot(tx - 1 - k) : 0);
  J.th_x = th + slot(tx);
  J.ths = ths;
  J.u = u;
  const size_t col = static_cast<size_t>(w) * P.Nj;
  const int32_t tho = (update && R.th_slot) ? gbl(R.th_slot)[t - R.slot_t0] : -1;
  const int32_t Io = (update && R.I_slot) ? gbl(R.I_slot)[t - R.slot_t0] : -1;
  J.th_out = tho >= 0 ? gblw(R.th_out) + static_cast<size_t>(tho) * slot_stride + col : nullptr;
  J.I_out = Io >= 0 ? gblw(R.I_out) + static_cast<size_t>(Io) * slot_stride + col : nullptr;
  J.Is_t = term_view(R.Is, update, P.zero, w, t);
  J.Vs_t = term_view(R.Vs, update && ths, P.zero, w, t);
  J.Is_x = term_view(R.Is, prep, P.zero, w, tx);
  J.t_zero = R.T.zero;
  J.from_T = !R.sigma.base;
  J.sig_x = term_view(J.from_T ? R.T : R.sigma, prep && !R.T.zero, P.zero, w, tx);
  J.noise = (prep && !R.T.zero && R.noise)
                ? gbl(R.noise) + static_cast<size_t>(tx - R.noise_t0) * slot_stride + col
                : nullptr;
  J.Ic = gbl(R.Ic ? R.Ic : P.Ic);
  J.is_zero = R.Is.zero;
  J.dt = R.dt;
  J.seed = R.seed;
  J.noise_col = noise_col;
  J.tx = static_cast<uint64_t>(tx);
  const bool fast = SW == 3 && update && prep && !ths && !J.th_out && !J.I_out && !J.noise && !J.from_T;
  if (fast && P.coef_si == 0 && J.sig_x.si == 0) junction_loop_fast<BS, true, Z>(P, J, zs);
  else if (fast) junction_loop_fast<BS, false, Z>(P, J, zs);
  else if (update && prep) junction_loop<BS, SW, true, true, Z>(P, J, zs);
  else if (update) junction_loop<BS, SW, true, false, Z>(P, J, zs);
  else if (prep) junction_loop<BS, SW, false, true, Z>(P, J, zs);
}

// The face phase's junction values of step t: x_t, in history slot t mod
// (SW - 1), written by the junction pass before it.
__device__ __forceinline__ JuncSource step_source(const GAS double* th, const GAS double* ths,
                                                 size_t slot_stride, int SW, int64_t t) {
  return JuncSource{th + static_cast<size_t>(t % (SW - 1)) * slot_stride, ths, true, 0.0};
}

// n_steps steps of column w of one problem, by the threads the store names.
template <int BS, int SW, class Z>
__device__ __forceinline__ void evolve_steps(const DevProblem& P, const jja_problem_run& R, int64_t w,
                                             const Z& zs, int32_t n_steps, int64_t step0) {
  const size_t slot_stride = static_cast<size_t>(R.W) * P.Nj;
  GAS double* __restrict__ th = gblw(R.theta) + static_cast<size_t>(w) * P.Nj;
  GAS double* __restrict__ ths = R.theta_s ? gblw(R.theta_s) + static_cast<size_t>(w) * P.Nj : nullptr;
  GAS double* __restrict__ u = Z::kGlobal ? nullptr : gblw(R.u) + static_cast<size_t>(w) * P.Nj;
  const uint32_t noise_col = static_cast<uint32_t>(R.column_offset + w);

  junction_pass<BS, SW>(P, R, th, ths, u, zs, w, slot_stride, step0, false, step0, true, noise_col);
  zs.sync();
  for (int32_t n = 0; n < n_steps; ++n) {
    const int64_t t = step0 + n;
    if (P.Nf > 0) {  // a circuit without faces has nothing to solve (K is 0 x 0)
      face_phase<BS>(P, &R.f, Z::kGlobal ? step_source(th, ths, slot_stride, SW, t) : JuncSource{u, nullptr, false, 0.0},
                     zs, w, t);
      zs.barrier();
      solve_face_system(P, zs);
    }
    // finish step t and prepare step t+1 in one pass
    junction_pass<BS, SW>(P, R, th, ths, u, zs, w, slot_stride, t, true, t + 1, n + 1 < n_steps, noise_col);
    zs.sync();
  }
}

// Columns wa and wa+1 (two) or wa alone of one problem, worked by one set of
// threads: the face phases and junction passes run column by column, the solve
// streams each record once for both, and za's barriers serve both.
template <int BS, int SW>
__device__ __forceinline__ void evolve_steps_pair(const DevProblem& P, const jja_problem_run& R, int64_t wa,
                                                  bool two, const SplitStore& za, const SplitStore& zb,
                                                  int32_t n_steps, int64_t step0) {
  const size_t slot_stride = static_cast<size_t>(R.W) * P.Nj;
  const int64_t wb = wa + 1;
  const size_t ca = static_cast<size_t>(wa) * P.Nj, cb = two ? static_cast<size_t>(wb) * P.Nj : ca;
  // column b's arrays alias column a's when it is alone (and are then unused)
  GAS double* tha = gblw(R.theta) + ca;
  GAS double* thb = gblw(R.theta) + cb;
  GAS double* thsa = R.theta_s ? gblw(R.theta_s) + ca : nullptr;
  GAS double* thsb = R.theta_s ? gblw(R.theta_s) + cb : nullptr;
  const uint32_t na = static_cast<uint32_t>(R.column_offset + wa), nb = static_cast<uint32_t>(R.column_offset + wb);

  junction_pass<BS, SW>(P, R, tha, thsa, nullptr, za, wa, slot_stride, step0, false, step0, true, na);
  if (two) junction_pass<BS, SW>(P, R, thb, thsb, nullptr, zb, wb, slot_stride, step0, false, step0, true, nb);
  za.sync();
  for (int32_t n = 0; n < n_steps; ++n) {
    const int64_t t = step0 + n;
    if (P.Nf > 0) {
      face_phase<BS>(P, &R.f, step_source(tha, thsa, slot_stride, SW, t), za, wa, t);
      if (two) face_phase<BS>(P, &R.f, step_source(thb, thsb, slot_stride, SW, t), zb, wb, t);
      za.barrier();
      if (two)
        solve_pair(P, za, zb);
      else
        solve(P, za);
    }
    const bool prep = n + 1 < n_steps;
    junction_pass<BS, SW>(P, R, tha, thsa, nullptr, za, wa, slot_stride, t, true, t + 1, prep, na);
    if (two) junction_pass<BS, SW>(P, R, thb, thsb, nullptr, zb, wb, slot_stride, t, true, t + 1, prep, nb);
    za.sync();
  }
}

// Column w of one problem in one workgroup: face vector by solve position,
// stage scratch after it.
template <int BS, int SW, bool GLOBAL>
__device__ __forceinline__ void evolve_column(const DevProblem& P, const jja_problem_run& R, int64_t w,
                                              double* lds, int32_t n_steps, int64_t step0) {
  FaceStore<GLOBAL> zs;
  if constexpr (GLOBAL)
    zs.p = gblw(R.zbuf) + static_cast<size_t>(w) * (static_cast<size_t>(P.Nf) + P.max_stage_rows);
  else
    zs.p = lds;
  zs.vt = threadIdx.x;
  zs.vbs = BS;
  evolve_steps<BS, SW>(P, R, w, zs, n_steps, step0);
}

// A column shared by P.parts workgroups (part `part` of the work item's column
// w): face vector in global memory, a column barrier wherever the one-workgroup
// kernel has a workgroup barrier.
__device__ __forceinline__ SplitStore split_store(const DevProblem& P, GAS double* zbuf, int64_t w,
                                                  int32_t part, uint32_t* ctrs, const WorkItem& wi,
                                                  int32_t* err, uint32_t wait_limit, int32_t* stopped,
                                                  int BS, LAS double* lds) {
  SplitStore zs;
  zs.p = zbuf + static_cast<size_t>(w) * (static_cast<size_t>(P.Nf) + P.max_stage_rows);
  zs.vt = part * BS + static_cast<int32_t>(threadIdx.x);
  zs.vbs = P.parts * BS;
  zs.ctr = ctrs + 2 * static_cast<int64_t>(wi.slot);
  zs.xcds = zs.ctr + 1;
  zs.parts = static_cast<uint32_t>(P.parts);
  zs.wait_limit = wait_limit;
  zs.err = err;
  zs.stopped = stopped;
  zs.gen = 0;
  zs.one_xcd = false;
  zs.lds = lds;
  if (threadIdx.x == 0 && P.parts > 1)
    __hip_atomic_fetch_or(zs.xcds, 1u << xcd_id(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return zs;
}

// PAIR: a work item is a pair of columns (2q, 2q+1; the last one alone when W is
// odd), its barrier words those of the pair.
template <int BS, int SW, bool PAIR>
__global__ __launch_bounds__(BS) void evolve_split_kernel(const DevProblem* __restrict__ probs,
                                                          const jja_problem_run* __restrict__ runs,
                                                          const WorkItem* __restrict__ work,
                                                          int32_t n_steps, int64_t step0, uint32_t* ctrs,
                                                          int32_t* err, uint32_t wait_limit) {
  extern __shared__ double split_lds[];
  __shared__ int32_t stopped;
  if (threadIdx.x == 0) stopped = 0;
  __syncthreads();
  const WorkItem wi = work[blockIdx.x];
  const DevProblem& P = probs[wi.problem];
  const jja_problem_run& R = runs[wi.problem];
  const int64_t w = wi.column / P.parts;
  const int32_t part = wi.column % P.parts;
  if constexpr (PAIR) {
    const int64_t wa = 2 * w;
    const bool two = wa + 1 < R.W;
    const SplitStore za = split_store(P, gblw(R.zbuf), wa, part, ctrs, wi, err, wait_limit, &stopped, BS,
                                        (LAS double*)split_lds);
    SplitStore zb = za;
    if (two) zb.p = za.p + (static_cast<size_t>(P.Nf) + P.max_stage_rows);
    evolve_steps_pair<BS, SW>(P, R, wa, two, za, zb, n_steps, step0);
  } else {
    const SplitStore zs = split_store(P, gblw(R.zbuf), w, part, ctrs, wi, err, wait_limit, &stopped, BS,
                                        (LAS double*)split_lds);
    evolve_steps<BS, SW>(P, R, w, zs, n_steps, step0);
  }
}

// One workgroup per work item (problem, column). STORE: 0 face stores in LDS,
// 1 in global memory, 2 either, as each problem's index width says (a batch of
// small and large circuits in one launch).
template <int BS, int SW, int STORE>
__global__ __launch_bounds__(BS) void evolve_kernel(const DevProblem* __restrict__ probs,
                                                    const jja_problem_run* __restrict__ runs,
                                                    const WorkItem* __restrict__ work,
                                                    int32_t n_steps, int64_t step0) {
  extern __shared__ double lds[];
  const WorkItem wi = work[blockIdx.x];
  const DevProblem& P = probs[wi.problem];
  const jja_problem_run& R = runs[wi.problem];
  if constexpr (STORE == 2) {
    if (P.index_bytes == 4)
      evolve_column<BS, SW, true>(P, R, wi.column, lds, n_steps, step0);
    else
      evolve_column<BS, SW, false>(P, R, wi.column, lds, n_steps, step0);
  } else {
    evolve_column<BS, SW, STORE == 1>(P, R, wi.column, lds, n_steps, step0);
  }
}

// ------------------------------------------------------------------ adjoint
// Adjoint of theta_k wherever a run stores it: lam[m mod 2].
struct AdjCol {
  GAS double* lam;  // slot m mod 2 at lam + (m & 1) * stride
  const GAS double* g;  // this column of the upstream gradient (slot s at g + s * stride)
  GAS double* u;
  const GAS double* hist;
  int64_t hist_m0;
  size_t stride;
  const GAS double* Ic;
};

__device__ __forceinline__ double adj_g(const jja_adjoint_run& A, const AdjCol& C, int64_t m, int32_t j) {
  if (!A.g || m < A.g_lo || m >= A.g_hi) return 0.0;
  const int32_t sl = gbl(A.g_slot)[m - A.slot_t0];
  return sl >= 0 ? C.g[static_cast<size_t>(sl) * C.stride + j] : 0.0;
}

// lambda_k += g_k (complete) and u = lambda_k / c0: the first step of a sweep.
template <class Z>
__device__ __forceinline__ void adjoint_prime(const DevProblem& P, const jja_adjoint_run& A,
                                              const AdjCol& C, const Z& zs, int64_t k) {
  GAS double* __restrict__ lk = C.lam + (k & 1) * C.stride;
  for (int32_t j = zs.vt; j < P.Nj; j += zs.vbs) {
    const double l = lk[j] + adj_g(A, C, k, j);
    lk[j] = l;
    C.u[j] = l * rc0_at(P, j);
  }
}

// After the solve of step k (face vector J' = K^-1 A D lambda_k in LDS):
// mu_k = (A^T J' - lambda_k)/c0 and the updates of lambda_{k-1}, lambda_{k-2}
// and of the gradients; then lambda_{k-1} gets g_{k-1} and u = lambda_{k-1}/c0
// prepares step k-1.
template <class Z>
__device__ __forceinline__ void adjoint_junctions(const DevProblem& P, const jja_adjoint_run& A,
                                                  const AdjCol& C, const Z& zs, int64_t k, int64_t w) {
  const int32_t Nj = P.Nj;
  const GAS vint2* __restrict__ jcode = gbl(reinterpret_cast<const vint2*>(P.junc_code));
  GAS double* __restrict__ lk = C.lam + (k & 1) * C.stride;
  GAS double* __restrict__ lk1 = C.lam + ((k - 1) & 1) * C.stride;
  const GAS double* __restrict__ h1 = C.hist + static_cast<size_t>(k - 1 - C.hist_m0) * C.stride;
  const GAS double* __restrict__ h2 = C.hist + static_cast<size_t>(k - 2 - C.hist_m0) * C.stride;
  GAS double* __restrict__ gIc = gblw(A.grad_Ic) + static_cast<size_t>(w) * Nj;
  GAS double* __restrict__ gIs = gblw(A.grad_Is) + static_cast<size_t>(w) * Nj;
  for (int32_t j = zs.vt; j < Nj; j += zs.vbs) {
    const vint2 code = jcode[j];
    const double rc0 = rc0_at(P, j), c1 = coef_at(P, 1, j), c2 = coef_at(P, 2, j);
    const double lam = lk[j], lprev = lk1[j], t1 = h1[j], t2 = h2[j], Ic = C.Ic[j];
    const double gj = adj_g(A, C, k - 1, j);
    const double gi = gIc[j], gs = gIs[j];
    double y = 0.0;
    if (code.x >= 0) y = (code.x & 1) ? y - zs.rd(code.x >> 1) : y + zs.rd(code.x >> 1);
    if (code.y >= 0) y = (code.y & 1) ? y - zs.rd(code.y >> 1) : y + zs.rd(code.y >> 1);
    const double mu = (y - lam) * rc0;
    double sn, cs;
    jja_sincos(2.0 * t1 - t2, &sn, &cs);
    const double a = 2.0 * Ic * cs + c1, b = c2 - Ic * cs;
    gIc[j] = gi + sn * mu;
    gIs[j] = gs - mu;
    lk[j] = b * mu;  // slot k mod 2 now holds lambda_{k-2}'s partial sum
    const double l1 = (lprev + a * mu) + gj;
    lk1[j] = l1;
    C.u[j] = l1 * rc0;
  }
}

// The column of work item wi: lambda, history, gradients and upstream
// gradient of column w.
__device__ __forceinline__ AdjCol adjoint_column(const DevProblem& P, const jja_adjoint_run& A, int64_t w) {
  const size_t stride = static_cast<size_t>(A.W) * P.Nj;
  const size_t col = static_cast<size_t>(w) * P.Nj;
  AdjCol C;
  C.lam = gblw(A.lam) + col;
  C.g = A.g ? gbl(A.g) + col : nullptr;
  C.u = gblw(A.u) + col;
  C.hist = gbl(A.hist) + col;
  C.hist_m0 = A.hist_m0;
  C.stride = stride;
  C.Ic = gbl(A.Ic ? A.Ic : P.Ic);
  return C;
}

// Steps k_hi down to k_lo of the adjoint sweep of column w.
template <int BS, class Z>
__device__ __forceinline__ void adjoint_steps(const DevProblem& P, const jja_adjoint_run& A, int64_t w,
                                              const Z& zs, int64_t k_hi, int64_t k_lo) {
  const AdjCol C = adjoint_column(P, A, w);
  GAS double* __restrict__ gf = gblw(A.grad_f) + static_cast<size_t>(w) * P.Nf;
  if (A.prime) adjoint_prime(P, A, C, zs, k_hi);
  zs.sync();
  for (int64_t k = k_hi; k >= k_lo; --k) {
    if (P.Nf > 0) {
      face_phase<BS>(P, nullptr, JuncSource{C.u, nullptr, false, 0.0}, zs, w, k);
      zs.barrier();
      solve_face_system(P, zs);
    }
    for (int32_t p = zs.vt; p < P.Nf; p += zs.vbs) gf[p] = gf[p] - kTwoPi * zs.rd(p);
    adjoint_junctions(P, A, C, zs, k, w);
    zs.sync();
  }
}

// adjoint_steps for columns wa and wa+1 (two) or wa alone, on one set of
// threads: per-column face phases and junction updates, one shared solve, za's
// barriers.
template <int BS>
__device__ __forceinline__ void adjoint_steps_pair(const DevProblem& P, const jja_adjoint_run& A, int64_t wa,
                                                   bool two, const SplitStore& za, const SplitStore& zb,
                                                   int64_t k_hi, int64_t k_lo) {
  const int64_t wb = two ? wa + 1 : wa;
  const AdjCol Ca = adjoint_column(P, A, wa), Cb = adjoint_column(P, A, wb);
  // column b's gradient aliases column a's when it is alone (and is then unused)
  GAS double* gfa = gblw(A.grad_f) + static_cast<size_t>(wa) * P.Nf;
  GAS double* gfb = gblw(A.grad_f) + static_cast<size_t>(wb) * P.Nf;
  if (A.prime) {
    adjoint_prime(P, A, Ca, za, k_hi);
    if (two) adjoint_prime(P, A, Cb, zb, k_hi);
  }
  za.sync();
  for (int64_t k = k_hi; k >= k_lo; --k) {
    if (P.Nf > 0) {
      face_phase<BS>(P, nullptr, JuncSource{Ca.u, nullptr, false, 0.0}, za, wa, k);
      if (two) face_phase<BS>(P, nullptr, JuncSource{Cb.u, nullptr, false, 0.0}, zb, wb, k);
      za.barrier();
      if (two)
        solve_pair(P, za, zb);
      else
        solve(P, za);
    }
    for (int32_t p = za.vt; p < P.Nf; p += za.vbs) gfa[p] = gfa[p] - kTwoPi * za.rd(p);
    if (two)
      for (int32_t p = zb.vt; p < P.Nf; p += zb.vbs) gfb[p] = gfb[p] - kTwoPi * zb.rd(p);
    adjoint_junctions(P, A, Ca, za, k, wa);
    if (two) adjoint_junctions(P, A, Cb, zb, k, wb);
    za.sync();
  }
}

template <int BS, bool GLOBAL>
__global__ __launch_bounds__(BS) void adjoint_kernel(const DevProblem* __restrict__ probs,
                                                     const jja_adjoint_run* __restrict__ runs,
                                                     const WorkItem* __restrict__ work, int64_t k_hi,
                                                     int64_t k_lo) {
  extern __shared__ double lds[];
  const WorkItem wi = work[blockIdx.x];
  const DevProblem& P = probs[wi.problem];
  const jja_adjoint_run& A = runs[wi.problem];
  const int64_t w = wi.column;
  FaceStore<GLOBAL> zs;
  if constexpr (GLOBAL)
    zs.p = gblw(A.zbuf) + static_cast<size_t>(w) * (static_cast<size_t>(P.Nf) + P.max_stage_rows);
  else
    zs.p = lds;
  zs.vt = threadIdx.x;
  zs.vbs = BS;
  adjoint_steps<BS>(P, A, w, zs, k_hi, k_lo);
}

template <int BS, bool PAIR>
__global__ __launch_bounds__(BS) void adjoint_split_kernel(const DevProblem* __restrict__ probs,
                                                           const jja_adjoint_run* __restrict__ runs,
                                                           const WorkItem* __restrict__ work, int64_t k_hi,
                                                           int64_t k_lo, uint32_t* ctrs, int32_t* err,
                                                           uint32_t wait_limit) {
  extern __shared__ double split_lds[];
  __shared__ int32_t stopped;
  if (threadIdx.x == 0) stopped = 0;
  __syncthreads();
  const WorkItem wi = work[blockIdx.x];
  const DevProblem& P = probs[wi.problem];
  const jja_adjoint_run& A = runs[wi.problem];
  const int64_t w = wi.column / P.parts;
  const int32_t part = wi.column % P.parts;
  if constexpr (PAIR) {
    const int64_t wa = 2 * w;
    const bool two = wa + 1 < A.W;
    const SplitStore za = split_store(P, gblw(A.zbuf), wa, part, ctrs, wi, err, wait_limit, &stopped, BS,
                                        (LAS double*)split_lds);
    SplitStore zb = za;
    if (two) zb.p = za.p + (static_cast<size_t>(P.Nf) + P.max_stage_rows);
    adjoint_steps_pair<BS>(P, A, wa, two, za, zb, k_hi, k_lo);
  } else {
    const SplitStore zs = split_store(P, gblw(A.zbuf), w, part, ctrs, wi, err, wait_limit, &stopped, BS,
                                        (LAS double*)split_lds);
    adjoint_steps<BS>(P, A, w, zs, k_hi, k_lo);
  }
}

// n[col][f] = -sum_j A[f,j] rint(theta_j / 2pi) (time_evolution.py:734-755): one
// workgroup per phase vector, one thread per face position (a workgroup takes
// the slots of all parts of a split plan in turn).
template <int BS>
__global__ __launch_bounds__(BS) void vortex_kernel(const DevProblem* __restrict__ probs, int32_t p,
                                                    const double* __restrict__ theta,
                                                    int32_t* __restrict__ n_out) {
  const DevProblem& P = probs[p];
  const int64_t col = blockIdx.x;
  const GAS double* __restrict__ th = gbl(theta) + col * P.Nj;
  const int64_t vbs = static_cast<int64_t>(P.parts) * BS;
  for (int32_t part = 0; part < P.parts; ++part) {
    const int32_t vt = part * BS + static_cast<int32_t>(threadIdx.x);
    int64_t acc = 0;
    for (int32_t k = 0; k < P.n_face_slices; ++k) {
      const jja_slice m = load_rec(P.face_slices, k);
      const GAS int32_t* __restrict__ code = gbl(P.face_code) + m.base + vt;
      for (int32_t it = 0; it < m.n_iter; ++it) {
        const int32_t cd = code[it * vbs];
        if (cd < 0) continue;
        const int64_t kk = static_cast<int64_t>(rint(th[cd >> 1] / kTwoPi));
        acc += (cd & 1) ? kk : -kk;
      }
      if (m.flags & JJA_SLICE_FINISH) {
        const int32_t pos = gbl(P.slice_rows)[m.rowbase + vt];
        if (pos >= 0) gblw(n_out)[col * P.Nf + gbl(P.face_of_pos)[pos]] = static_cast<int32_t>(acc);
        acc = 0;
      }
    }
  }
}

// K x = b for `cols` right-hand sides with problem p's solve program: one
// workgroup per column loads b into the face store by solve position (and the
// first stage's scratch, which equals b), runs the staged forward and backward
// solves, and writes x back in face order. b and x are [cols][Nf] with leading
// dimensions ldb and ldx.
template <int BS, bool GLOBAL>
__global__ __launch_bounds__(BS) void spd_solve_kernel(const DevProblem* __restrict__ probs, int32_t p,
                                                       const double* __restrict__ b, int64_t ldb,
                                                       double* __restrict__ x, int64_t ldx, double* zbuf) {
  extern __shared__ double lds[];
  const DevProblem& P = probs[p];
  const int64_t col = blockIdx.x;
  FaceStore<GLOBAL> zs;
  if constexpr (GLOBAL)
    zs.p = gblw(zbuf) + col * (static_cast<int64_t>(P.Nf) + P.max_stage_rows);
  else
    zs.p = lds;
  zs.vt = threadIdx.x;
  zs.vbs = BS;
  const GAS double* __restrict__ bc = gbl(b) + col * ldb;
  GAS double* __restrict__ xc = gblw(x) + col * ldx;
  const GAS int32_t* __restrict__ fop = gbl(P.face_of_pos);
  for (int32_t q = threadIdx.x; q < P.Nf; q += BS) {
    const double v = bc[fop[q]];
    zs.wr(q, v);
    if (q < P.copy_rows) zs.wr(P.Nf + q, v);
  }
  zs.sync();
  solve_face_system(P, zs);
  zs.sync();
  for (int32_t q = threadIdx.x; q < P.Nf; q += BS) xc[fop[q]] = zs.rd(q);
}

struct Plan {
  int device = 0;
  std::vector<void*> allocations;
  std::vector<DevProblem> problems;
  std::vector<int64_t> lds_bytes;     // per problem; 0 for a global face store
  std::vector<int64_t> zbuf_doubles;  // per problem; 0 for an LDS face store
  std::vector<int64_t> cost;          // per problem: entries one step of a column reads
  std::vector<int32_t> max_cols;      // per problem: the most columns a run may have
  DevProblem* d_problems = nullptr;
  jja_problem_run* d_runs = nullptr;
  WorkItem* d_work = nullptr;
  size_t work_cap = 0;
  int64_t lds_limit = 0;
  int32_t block_size = 1024;
  int32_t stencil_width = 3;
  jja_adjoint_run* d_adj = nullptr;
  int32_t parts = 1;                 // workgroups per column (block_size is one workgroup's)
  bool pairs = false;                // split launches work two columns per work item
  uint32_t* d_ctrs = nullptr;        // split launches: barrier words per column (counter, XCD set)
  size_t ctr_cap = 0;
  int32_t* d_err = nullptr;          // split launches: [0] a column barrier gave up,
                                     // [1] workgroups whose column ran on one XCD
  int32_t n_cu = 0;                  // compute units of the device
  uint32_t wait_limit = kWaitLimit;  // split launches: polls before a column barrier gives up
  int64_t split_lds = 0;             // split launches: dynamic LDS bytes (part-local lattice solves)
  // what the device copies of the launch descriptors hold: a launch uploads
  // only what changed since the previous one
  std::vector<jja_problem_run> h_runs;
  std::vector<jja_adjoint_run> h_adj;
  std::vector<WorkItem> h_work;
  bool runs_valid = false, adj_valid = false, work_valid = false;
  // the stream of the previous launch and an event recorded after it
  hipEvent_t done = nullptr;
  hipStream_t last_stream = nullptr;
  bool launched = false;
  // per kernel: the dynamic LDS limit set on it, and its resident workgroups per CU
  std::vector<std::pair<const void*, int>> lds_attr;
  std::vector<std::pair<const void*, int>> per_cu;
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
  if (plan->done) (void)hipEventDestroy(plan->done);
  delete plan;
}

// Launches of a plan share its device descriptors, so they run on one stream
// at a time: a launch on another stream waits for nothing and is refused while
// the previous launch has not finished.
int claim_stream(Plan* plan, hipStream_t stream) {
  if (!plan->launched || stream == plan->last_stream) return JJA_OK;
  const hipError_t e = hipEventQuery(plan->done);
  if (e == hipErrorNotReady)
    return fail(JJA_EBUSY, "the plan's previous launch, on another stream, has not finished");
  if (e != hipSuccess) return fail(JJA_EHIP, std::string("hipEventQuery: ") + hipGetErrorString(e));
  return JJA_OK;
}

int mark_launched(Plan* plan, hipStream_t stream) {
  JJA_HIP(hipEventRecord(plan->done, stream));
  plan->last_stream = stream;
  plan->launched = true;
  return JJA_OK;
}

// Copies n host descriptors to the plan's device array when they differ from
// what it holds (pageable source: staged before the call returns, ordered after
// earlier work on the stream).
template <typename T>
int upload_if_changed(T* dev, const T* host, size_t n, std::vector<T>& shadow, bool& valid, hipStream_t stream) {
  if (valid && shadow.size() == n && std::memcmp(shadow.data(), host, n * sizeof(T)) == 0) return JJA_OK;
  JJA_HIP(hipMemcpyAsync(dev, host, n * sizeof(T), hipMemcpyHostToDevice, stream));
  shadow.assign(host, host + n);
  valid = true;
  return JJA_OK;
}

// Raise a kernel's dynamic LDS limit to at least lds bytes. The limit belongs
// to the kernel on a device, not to a plan, so the largest value any plan has
// asked for is kept process-wide and never lowered (a plan that set a smaller
// one would otherwise break a later launch of another plan).
int raise_lds_attr(const void* kern, int lds) {
  static std::mutex mu;
  static std::vector<std::pair<std::pair<int, const void*>, int>> limits;
  int dev = 0;
  JJA_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  for (auto& a : limits)
    if (a.first.first == dev && a.first.second == kern) {
      if (a.second >= lds) return JJA_OK;
      JJA_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
      a.second = lds;
      return JJA_OK;
    }
  JJA_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  limits.push_back({{dev, kern}, lds});
  return JJA_OK;
}

// The dynamic LDS limit of a kernel with a dynamic LDS face store.
int set_lds_limit(Plan* plan, const void* kern, int lds) {
  for (auto& a : plan->lds_attr)
    if (a.first == kern) {
      if (a.second >= lds) return JJA_OK;
      if (int rc = raise_lds_attr(kern, lds)) return rc;
      a.second = lds;
      return JJA_OK;
    }
  hipFuncAttributes attr;
  JJA_HIP(hipFuncGetAttributes(&attr, kern));
  // the solve reads LDS at the plan's absolute byte offsets: the dynamic LDS
  // array must start at address 0, i.e. the kernel must have no static LDS
  if (attr.sharedSizeBytes != 0) return fail(JJA_EUNSUPPORTED, "kernel with a dynamic LDS face store has static LDS");
  if (int rc = raise_lds_attr(kern, lds)) return rc;
  plan->lds_attr.push_back({kern, lds});
  return JJA_OK;
}

bool pow2_le(int32_t g, int32_t limit) { return g >= 1 && g <= limit && (g & (g - 1)) == 0; }

int32_t parts_of(const jja_problem_desc& d) { return d.parts > 0 ? d.parts : 1; }

// The problems of one plan share the stencil width and the workgroup size
// (block_size / parts); the part count may differ between problems, but if any
// column is split, every problem keeps its face store in global memory (one
// launch runs them all).
int check_plan_shapes(int32_t n, const jja_problem_desc* descs) {
  bool split = false;
  for (int32_t p = 0; p < n; ++p) split = split || parts_of(descs[p]) > 1;
  for (int32_t p = 0; p < n; ++p) {
    const jja_problem_desc& d = descs[p];
    if (d.stencil_width != descs[0].stencil_width)
      return fail(JJA_EINVAL, "jja_plan_create: all problems of a plan need the same stencil width");
    if (parts_of(d) <= 0 || d.block_size % parts_of(d) ||
        d.block_size / parts_of(d) != descs[0].block_size / parts_of(descs[0]))
      return fail(JJA_EINVAL, "jja_plan_create: all problems of a plan need the same block_size/parts");
    if (split && d.index_bytes != 4)
      return fail(JJA_EINVAL, "jja_plan_create: with a split column, every problem needs the global face store");

    if ((d.column_pairs != 0) != (descs[0].column_pairs != 0))
      return fail(JJA_EINVAL, "jja_plan_create: column_pairs must be the same for every problem of a plan");
  }
  return JJA_OK;
}

// Every index the kernels follow is checked here, once, on the host.
int validate(const jja_problem_desc& d) {
  const int32_t Nj = d.n_junctions, Nf = d.n_faces, BS = d.block_size, ms = d.max_stage_rows;
  if (Nj < 0 || Nf < 0 || (Nj == 0 && Nf == 0) || d.stencil_width < 3 || d.stencil_width > 5 || d.n_face_slices < 0 ||
      d.n_solve_slices < 0 || ms < 0)
    return fail(JJA_EINVAL, "jja_plan_create: bad sizes");
  const int32_t parts = d.parts > 0 ? d.parts : 1;
  if (!pow2_le(parts, 32)) return fail(JJA_EINVAL, "jja_plan_create: parts must be a power of two <= 32");
  const int32_t wg = BS / parts;
  if (BS % parts || (wg != 64 && wg != 128 && wg != 256 && wg != 512 && wg != 1024))
    return fail(JJA_EINVAL, "jja_plan_create: block_size/parts must be 64, 128, 256, 512 or 1024");
  if (parts > 1 && d.index_bytes != 4)
    return fail(JJA_EINVAL, "jja_plan_create: a column split over workgroups needs the global face store");
  if (d.column_pairs != 0 && (d.column_pairs != 1 || parts < 2 || wg > 512))
    return fail(JJA_EINVAL, "jja_plan_create: column_pairs needs split columns (parts >= 2) and workgroups "
                            "of at most 512 threads");
  if (d.max_columns < 1 || d.max_columns > (1 << 20))
    return fail(JJA_EINVAL, "jja_plan_create: max_columns must be 1 .. 2^20");
  if (d.cpr_terms < 0 || d.cpr_terms > JJA_CPR_MAX_TERMS)
    return fail(JJA_EINVAL, "jja_plan_create: cpr_terms must be 0 .. " + std::to_string(JJA_CPR_MAX_TERMS));
  if (Nj > 0 && (!d.Ic || !d.Rv || !d.coef || !d.junc_code))
    return fail(JJA_EINVAL, "jja_plan_create: missing junction arrays");
  for (int64_t e = 0; e < 2 * static_cast<int64_t>(Nj); ++e)
    if (d.junc_code[e] < -1 || (d.junc_code[e] >> 1) >= Nf)
      return fail(JJA_EINVAL, "junc_code out of range");
  if (Nf == 0)
    return (d.n_face_slices || d.n_solve_slices || d.lattice_rows) ? fail(JJA_EINVAL, "slices without faces")
                                                                   : JJA_OK;
  if (!d.face_of_pos || !d.face_code || !d.face_slices || !d.slice_rows ||
      (d.n_solve_slices && (!d.solve_slices || !d.ent_val || !d.ent_off)))
    return fail(JJA_EINVAL, "jja_plan_create: missing face or solve arrays");
  if (d.index_bytes != 2 && d.index_bytes != 4) return fail(JJA_EINVAL, "index_bytes must be 2 or 4");
  if (d.index_bytes == 2 && static_cast<int64_t>(Nf) + ms > 65536)
    return fail(JJA_EINVAL, "face vector exceeds 16-bit indices");
  if (static_cast<int64_t>(Nf) + ms > INT32_MAX) return fail(JJA_EINVAL, "face vector too large");
  std::vector<char> seen(Nf, 0);
  for (int32_t q = 0; q < Nf; ++q) {
    const int32_t f = d.face_of_pos[q];
    if (f < 0 || f >= Nf || seen[f]) return fail(JJA_EINVAL, "face_of_pos is not a permutation");
    seen[f] = 1;
  }
  auto check_slice = [&](const jja_slice& m, int64_t n_ent, bool paired) -> int {
    const int32_t G = m.flags & 0xff;
    if (!pow2_le(G, std::min(64, BS))) return fail(JJA_EINVAL, "slice lanes per row must be a power of two <= 64");
    if (m.n_iter < 0 || m.n_iter > kMaxIter) return fail(JJA_EINVAL, "slice n_iter must be 0..8");
    // face records: n_iter*BS codes; solve records: ceil(n_iter/2) pairs of BS doubles
    const int64_t extent = paired ? 2 * static_cast<int64_t>((m.n_iter + 1) / 2) * BS
                                  : static_cast<int64_t>(m.n_iter) * BS;
    if (m.base < 0 || m.base + extent > n_ent) return fail(JJA_EINVAL, "slice entries out of range");
    if (m.rowbase < 0 || m.rowbase + BS / G > d.n_slice_rows) return fail(JJA_EINVAL, "slice rows out of range");
    if ((m.flags >> 16) < 0 || (m.flags >> 16) > BS / G) return fail(JJA_EINVAL, "slice row count out of range");
    if (m.flags & 0xf800) return fail(JJA_EINVAL, "unknown slice flags");
    return JJA_OK;
  };
  for (int32_t k = 0; k < d.n_face_slices; ++k) {
    const jja_slice& m = d.face_slices[k];
    if (int rc = check_slice(m, d.n_face_entries, false)) return rc;
    if ((m.flags & 0xff) != 1) return fail(JJA_EINVAL, "face slices use one lane per row");
    if (m.n_iter > kFaceIter) return fail(JJA_EINVAL, "face slices hold at most 4 codes per thread");
    for (int32_t i = 0; i < BS; ++i) {
      const int32_t r = d.slice_rows[m.rowbase + i];
      if (r < -1 || r >= Nf) return fail(JJA_EINVAL, "face slice row out of range");
    }
    for (int64_t e = m.base; e < m.base + static_cast<int64_t>(m.n_iter) * BS; ++e)
      if (d.face_code[e] < -1 || (d.face_code[e] >> 1) >= Nj) return fail(JJA_EINVAL, "face_code out of range");
  }
  if (d.lattice_rows != 0) {
    const int64_t R = d.lattice_rows, C = d.lattice_cols, H = (C + 1) / 2, ES = H + (H & 1);
    if (R < 2 || C < 2 || R * C != Nf) return fail(JJA_EINVAL, "lattice rows * columns must equal n_faces (>= 2 each)");
    if (d.column_pairs) return fail(JJA_EINVAL, "a lattice solve does not pair columns");
    if (!d.lattice_sine || !d.lattice_inv) return fail(JJA_EINVAL, "jja_plan_create: missing lattice tables");
    if (d.n_solve_slices != 0 || d.face_copy_rows != 0)
      return fail(JJA_EINVAL, "a lattice solve has no solve records and no face copy rows");
    const int64_t B = std::max<int64_t>(1, std::min<int64_t>(R, int64_t{JJA_LATTICE_TRI_THREADS} * parts / C));
    const int64_t EF = d.index_bytes == 4 ? (ES + 3) & ~int64_t{3} : ES;
    if (ms < (Nf & 1) + std::max(2 * R * EF, 4 * B * C))
      return fail(JJA_EINVAL, "max_stage_rows too small for the lattice solve");
    if (!(d.lattice_g > 0.0) || !std::isfinite(d.lattice_g)) return fail(JJA_EINVAL, "lattice_g must be finite and > 0");
    return JJA_OK;
  }
  // solve programs: one per wave of 64 lanes; each runs a multiple of three
  // records and loads four more ahead, all of which must be in range
  const int32_t NW = BS / kWave;
  const int32_t* wp = d.solve_wave_ptr;
  if (!wp || wp[0] != 0 || wp[NW] != d.n_solve_slices)
    return fail(JJA_EINVAL, "solve_wave_ptr must cover the solve records");
  if (d.n_entries < kMaxIter * kWave || d.n_index < kMaxIter * kWave)
    return fail(JJA_EINVAL, "solve arrays too short");
  if (d.n_entries * static_cast<int64_t>(sizeof(double)) >= (int64_t{1} << 31))
    return fail(JJA_EUNSUPPORTED, "solve values of one problem exceed 2 GiB");
  if (d.face_copy_rows < 0 || d.face_copy_rows > Nf || d.face_copy_rows > ms)
    return fail(JJA_EINVAL, "face_copy_rows out of range");
  const int64_t lds_end = static_cast<int64_t>(Nf) + ms;
  int32_t barriers = -1;
  std::string seq0;  // column ('c') and workgroup ('w') barriers of wave 0, in order
  for (int32_t w = 0; w < NW; ++w) {
    if (wp[w + 1] - wp[w] < 4 || (wp[w + 1] - wp[w] - 4) % 3)
      return fail(JJA_EINVAL, "a wave program needs a record count divisible by 3 and 4 lookahead records");
    int32_t nb = 0;
    std::string seq;
    for (int32_t k = wp[w]; k < wp[w + 1]; ++k) {
      const jja_slice& m = d.solve_slices[k];
      const int32_t G = m.flags & 0xff, rows = m.flags >> 16;
      if (!pow2_le(G, kWave)) return fail(JJA_EINVAL, "solve record lanes per row must be a power of two <= 64");
      if (m.n_iter < 0 || m.n_iter > kMaxIter || rows < 0 || rows > kWave / G || (m.flags & 0xe000))
        return fail(JJA_EINVAL, "solve record fields out of range");
      if ((m.flags & JJA_SLICE_NARROW) && (d.index_bytes != 4 || (m.flags & JJA_SLICE_BLOCK)))
        return fail(JJA_EINVAL, "NARROW solve records need 4-byte indices and an off-block phase");
      if (m.base < 0 || (m.base % 2) || m.base + kMaxIter * kWave > d.n_entries || m.ibase < 0 ||
          (m.ibase % (16 / d.index_bytes)) || m.ibase + kMaxIter * kWave > d.n_index || m.rowbase < 0 ||
          m.rowbase + kWave / G > d.n_slice_rows)
        return fail(JJA_EINVAL, "solve record data out of range or misaligned");
      if (k >= wp[w + 1] - 4) continue;  // lookahead only: loaded, never executed
      if (m.flags & JJA_SLICE_BARRIER) {
        ++nb;
        seq.push_back('c');
      } else if (m.flags & JJA_SLICE_WG_BARRIER) {
        seq.push_back('w');
      }
      const bool block = m.flags & JJA_SLICE_BLOCK;
      for (int32_t i = 0; i < rows; ++i) {
        const int32_t r = d.slice_rows[m.rowbase + i];
        if (r == -1) continue;
        if (r < 0 || r >= Nf || (!block && (r - m.r0 < 0 || r - m.r0 >= ms)))
          return fail(JJA_EINVAL, "solve slice row out of range");
      }
      if (m.flags & JJA_SLICE_NARROW) {
        // lane base, then two 16-bit offsets per word
        const uint32_t* ip = static_cast<const uint32_t*>(d.ent_off) + m.ibase;
        for (int32_t l = 0; l < kWave; ++l) {
          uint32_t dmax = 0;
          for (int32_t q = 0; q < kMaxIter / 2; ++q) {
            const uint32_t wd = ip[kWave + 4 * l + q];
            dmax = std::max(dmax, std::max(wd & 0xffffu, wd >> 16));
          }
          if (static_cast<int64_t>(ip[l]) + dmax >= lds_end) return fail(JJA_EINVAL, "solve LDS offset out of range");
        }
      } else if (block && d.index_bytes == 4) {
        // one word per lane: base | len << 28; entries base + i*G, i < len
        for (int32_t l = 0; l < kWave; ++l) {
          const uint32_t wd = static_cast<const uint32_t*>(d.ent_off)[m.ibase + l];
          const int64_t len = wd >> 28, base = wd & 0x0FFFFFFFu;
          if (len > kMaxIter || (len > 0 && base + (len - 1) * G >= lds_end))
            return fail(JJA_EINVAL, "solve LDS offset out of range");
        }
      } else {
        for (int64_t e = m.ibase; e < m.ibase + kMaxIter * kWave; ++e) {
          const int64_t ix = d.index_bytes == 2 ? static_cast<const uint16_t*>(d.ent_off)[e]
                                                : static_cast<const uint32_t*>(d.ent_off)[e];
          if (ix >= lds_end) return fail(JJA_EINVAL, "solve LDS offset out of range");
        }
      }
    }
    if (barriers < 0) barriers = nb;
    if (nb != barriers) return fail(JJA_EINVAL, "wave programs disagree on their barrier count");
    // every wave meets the same barriers in the same order (a workgroup
    // barrier is a __syncthreads of its workgroup, and so is part of a column one)
    if (w == 0)
      seq0 = seq;
    else if (seq != seq0)
      return fail(JJA_EINVAL, "wave programs disagree on their barrier sequence");
  }
  return JJA_OK;
}

template <int BS, int SW, int STORE>
int launch_evolve(Plan* plan, const WorkItem* work, int32_t n_work, int64_t lds, int32_t n_steps,
                  int64_t step0, hipStream_t stream) {
  auto* kern = evolve_kernel<BS, SW, STORE>;
  if (int rc = set_lds_limit(plan, reinterpret_cast<const void*>(kern), static_cast<int>(lds))) return rc;
  hipLaunchKernelGGL(kern, dim3(n_work), dim3(BS), static_cast<size_t>(lds), stream,
                     plan->d_problems, plan->d_runs, work, n_steps, step0);
  JJA_HIP(hipGetLastError());
  return JJA_OK;
}

template <int BS, bool GLOBAL>
int launch_adjoint(Plan* plan, const WorkItem* work, int32_t n_work, int64_t lds, int64_t k_hi,
                   int64_t k_lo, hipStream_t stream) {
  auto* kern = adjoint_kernel<BS, GLOBAL>;
  if (int rc = set_lds_limit(plan, reinterpret_cast<const void*>(kern), static_cast<int>(lds))) return rc;
  hipLaunchKernelGGL(kern, dim3(n_work), dim3(BS), static_cast<size_t>(lds), stream,
                     plan->d_problems, plan->d_adj, work, k_hi, k_lo);
  JJA_HIP(hipGetLastError());
  return JJA_OK;
}

template <int BS>
int launch_vortex(Plan* plan, int32_t p, const double* theta, int32_t cols, int32_t* n_out,
                  hipStream_t stream) {
  hipLaunchKernelGGL(vortex_kernel<BS>, dim3(cols), dim3(BS), 0, stream, plan->d_problems, p, theta,
                     n_out);
  JJA_HIP(hipGetLastError());
  return JJA_OK;
}

// Copies the work items of one launch to the plan's array (sized at plan
// creation) when they differ from what it holds.
int upload_work(Plan* plan, const std::vector<WorkItem>& work, hipStream_t stream) {
  if (work.size() > plan->work_cap) return fail(JJA_EINVAL, "launch has more work items than the plan was sized for");
  return upload_if_changed(plan->d_work, work.data(), work.size(), plan->h_work, plan->work_valid, stream);
}

template <int BS, int STORE>
int launch_evolve_sw(Plan* plan, const WorkItem* work, int32_t n, int64_t lds, int32_t n_steps,
                     int64_t step0, hipStream_t stream) {
  switch (plan->stencil_width) {
    case 3: return launch_evolve<BS, 3, STORE>(plan, work, n, lds, n_steps, step0, stream);
    case 4: return launch_evolve<BS, 4, STORE>(plan, work, n, lds, n_steps, step0, stream);
    default: return launch_evolve<BS, 5, STORE>(plan, work, n, lds, n_steps, step0, stream);
  }
}

template <int STORE>
int launch_evolve_bs(Plan* plan, const WorkItem* work, int32_t n, int64_t lds, int32_t n_steps,
                     int64_t step0, hipStream_t stream) {
  switch (plan->block_size) {
    case 64: return launch_evolve_sw<64, STORE>(plan, work, n, lds, n_steps, step0, stream);
    case 128: return launch_evolve_sw<128, STORE>(plan, work, n, lds, n_steps, step0, stream);
    case 256: return launch_evolve_sw<256, STORE>(plan, work, n, lds, n_steps, step0, stream);
    case 512: return launch_evolve_sw<512, STORE>(plan, work, n, lds, n_steps, step0, stream);
    default: return launch_evolve_sw<1024, STORE>(plan, work, n, lds, n_steps, step0, stream);
  }
}

// The work items of a launch, largest problems first: workgroups are
// dispatched in index order as compute units free up, so this is a
// longest-first list schedule of uneven problems (the sort is stable: the
// columns of a problem, and problems of equal cost, keep their order).
void order_work(const Plan* plan, std::vector<WorkItem>& work) {
  std::stable_sort(work.begin(), work.end(), [plan](const WorkItem& a, const WorkItem& b) {
    return plan->cost[a.problem] > plan->cost[b.problem];
  });
}

// Work items of a split launch, from one item per column: every part of every
// column (its problem's part count), each column with its own barrier counter
// (slot). All workgroups of a cooperative launch are resident at once, so the
// order only decides placement: workgroups go to the eight XCDs in turn by
// block index (block b on XCD b mod 8), and the columns are packed onto the
// XCDs (largest part count first, each to the XCD with the most free blocks)
// so that the parts of a column share one L2. If a column does not fit, the
// items stay in column order.
std::vector<WorkItem> split_work(const Plan* plan, std::vector<WorkItem> cols) {
  order_work(plan, cols);
  constexpr size_t kXcd = 8;
  const size_t nc = cols.size();
  size_t total = 0;
  for (const WorkItem& c : cols) total += static_cast<size_t>(plan->problems[c.problem].parts);
  std::vector<WorkItem> out(total);
  // free block indices of each XCD, in ascending order
  std::vector<std::vector<size_t>> slots(kXcd);
  for (size_t b = 0; b < total; ++b) slots[b % kXcd].push_back(b);
  std::vector<size_t> next(kXcd, 0);
  std::vector<size_t> order(nc);
  for (size_t c = 0; c < nc; ++c) order[c] = c;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    return plan->problems[cols[a].problem].parts > plan->problems[cols[b].problem].parts;
  });
  bool packed = true;
  for (size_t c : order) {
    const size_t G = static_cast<size_t>(plan->problems[cols[c].problem].parts);
    size_t best = kXcd;
    for (size_t x = 0; x < kXcd; ++x)
      if (slots[x].size() - next[x] >= G && (best == kXcd || slots[x].size() - next[x] > slots[best].size() - next[best]))
        best = x;
    if (best == kXcd) {
      packed = false;
      break;
    }
    for (size_t g = 0; g < G; ++g)
      out[slots[best][next[best]++]] = {cols[c].problem, static_cast<int32_t>(cols[c].column * G + g),
                                        static_cast<int32_t>(c)};
  }
  if (!packed) {
    out.clear();
    for (size_t c = 0; c < nc; ++c) {
      const int32_t G = plan->problems[cols[c].problem].parts;
      for (int32_t g = 0; g < G; ++g)
        out.push_back({cols[c].problem, cols[c].column * G + g, static_cast<int32_t>(c)});
    }
  }
  return out;
}

// Work items of a problem with W columns: one per column, or one per pair of
// columns (the last one alone when W is odd) in a plan with column pairs.
int32_t items_of(const Plan* plan, int32_t W) { return plan->pairs ? (W + 1) / 2 : W; }

// Barrier words of n work items (arrival counter, XCD set), zeroed on the
// stream before a split launch (the array is sized at plan creation).
int reset_counters(Plan* plan, size_t n, hipStream_t stream) {
  n *= 2;
  if (n > plan->ctr_cap) return fail(JJA_EINVAL, "launch has more columns than the plan was sized for");
  JJA_HIP(hipMemsetAsync(plan->d_ctrs, 0, n * sizeof(uint32_t), stream));
  return JJA_OK;
}

// Cooperative launch of a split kernel: all n workgroups resident at once (the
// column barriers need every part of a column running), or an error. Each
// workgroup gets the plan's split_lds bytes of dynamic LDS.
template <typename... Args>
int launch_cooperative(Plan* plan, const void* kern, int BS, int32_t n, hipStream_t stream, Args... args) {
  const int lds = static_cast<int>(plan->split_lds);
  int per_cu = -1;
  for (const auto& a : plan->per_cu)
    if (a.first == kern) per_cu = a.second;
  if (lds > 0)
    if (int rc = raise_lds_attr(kern, lds)) return rc;
  if (per_cu < 0) {
    JJA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, BS, static_cast<size_t>(lds)));
    plan->per_cu.push_back({kern, per_cu});
  }
  if (static_cast<int64_t>(per_cu) * plan->n_cu < n)
    return fail(JJA_EUNSUPPORTED, "split launch of " + std::to_string(n) + " workgroups; the device holds " +
                                      std::to_string(static_cast<int64_t>(per_cu) * plan->n_cu) + " at once");
  void* argv[] = {static_cast<void*>(&args)...};
  JJA_HIP(hipLaunchCooperativeKernel(kern, dim3(n), dim3(BS), argv, static_cast<size_t>(lds), stream));
  return JJA_OK;
}

template <int BS>
int launch_evolve_split_bs(Plan* plan, int32_t n, int32_t n_steps, int64_t step0, hipStream_t stream) {
  const DevProblem* probs = plan->d_problems;
  const jja_problem_run* runs = plan->d_runs;
  const WorkItem* work = plan->d_work;
  const void* kern = plan->pairs ? reinterpret_cast<const void*>(evolve_split_kernel<BS, 3, BS <= 512>)
                                 : reinterpret_cast<const void*>(evolve_split_kernel<BS, 3, false>);
  return launch_cooperative(plan, kern, BS, n, stream, probs, runs, work, n_steps, step0, plan->d_ctrs,
                            plan->d_err, plan->wait_limit);
}

int launch_evolve_split(Plan* plan, int32_t n, int32_t n_steps, int64_t step0, hipStream_t stream) {
  if (plan->stencil_width != 3) return fail(JJA_EUNSUPPORTED, "split columns: stencil width 3 only");
  switch (plan->block_size) {
    case 256: return launch_evolve_split_bs<256>(plan, n, n_steps, step0, stream);
    case 512: return launch_evolve_split_bs<512>(plan, n, n_steps, step0, stream);
    case 1024: return launch_evolve_split_bs<1024>(plan, n, n_steps, step0, stream);
    default: return fail(JJA_EUNSUPPORTED, "split columns: workgroups of 256, 512 or 1024 threads");
  }
}

template <int BS>
int launch_adjoint_split_bs(Plan* plan, int32_t n, int64_t k_hi, int64_t k_lo, hipStream_t stream) {
  const DevProblem* probs = plan->d_problems;
  const jja_adjoint_run* runs = plan->d_adj;
  const WorkItem* work = plan->d_work;
  const void* kern = plan->pairs ? reinterpret_cast<const void*>(adjoint_split_kernel<BS, BS <= 512>)
                                 : reinterpret_cast<const void*>(adjoint_split_kernel<BS, false>);
  return launch_cooperative(plan, kern, BS, n, stream, probs, runs, work, k_hi, k_lo, plan->d_ctrs,
                            plan->d_err, plan->wait_limit);
}

int launch_adjoint_split(Plan* plan, int32_t n, int64_t k_hi, int64_t k_lo, hipStream_t stream) {
  switch (plan->block_size) {
    case 256: return launch_adjoint_split_bs<256>(plan, n, k_hi, k_lo, stream);
    case 512: return launch_adjoint_split_bs<512>(plan, n, k_hi, k_lo, stream);
    case 1024: return launch_adjoint_split_bs<1024>(plan, n, k_hi, k_lo, stream);
    default: return fail(JJA_EUNSUPPORTED, "split columns: workgroups of 256, 512 or 1024 threads");
  }
}

template <bool GLOBAL>
int launch_adjoint_bs(Plan* plan, const WorkItem* work, int32_t n, int64_t lds, int64_t k_hi,
                      int64_t k_lo, hipStream_t stream) {
  switch (plan->block_size) {
    case 64: return launch_adjoint<64, GLOBAL>(plan, work, n, lds, k_hi, k_lo, stream);
    case 128: return launch_adjoint<128, GLOBAL>(plan, work, n, lds, k_hi, k_lo, stream);
    case 256: return launch_adjoint<256, GLOBAL>(plan, work, n, lds, k_hi, k_lo, stream);
    case 512: return launch_adjoint<512, GLOBAL>(plan, work, n, lds, k_hi, k_lo, stream);
    default: return launch_adjoint<1024, GLOBAL>(plan, work, n, lds, k_hi, k_lo, stream);
  }
}

template <int BS>
int launch_spd_solve(Plan* plan, int32_t p, const double* b, int64_t ldb, double* x, int64_t ldx,
                     int32_t cols, double* zbuf, hipStream_t stream) {
  const bool global = plan->zbuf_doubles[p] > 0;
  const int lds = global ? 8 : static_cast<int>(std::max<int64_t>(plan->lds_bytes[p], 8));
  const void* kern = global ? reinterpret_cast<const void*>(spd_solve_kernel<BS, true>)
                            : reinterpret_cast<const void*>(spd_solve_kernel<BS, false>);
  if (int rc = set_lds_limit(plan, kern, lds)) return rc;
  if (global)
    hipLaunchKernelGGL((spd_solve_kernel<BS, true>), dim3(cols), dim3(BS), static_cast<size_t>(lds), stream,
                       plan->d_problems, p, b, ldb, x, ldx, zbuf);
  else
    hipLaunchKernelGGL((spd_solve_kernel<BS, false>), dim3(cols), dim3(BS), static_cast<size_t>(lds), stream,
                       plan->d_problems, p, b, ldb, x, ldx, zbuf);
  JJA_HIP(hipGetLastError());
  return JJA_OK;
}

// scratch's device address on each GPU, taken once per device with that device
// current (the array's name is a host-side stand-in, not a device address)
constexpr int kMaxDevices = 64;
double* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_ptr(int device, double** out) {
  if (device < 0 || device >= kMaxDevices) return fail(JJA_EINVAL, "jja_solve_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = -1;
    JJA_HIP(hipGetDevice(&cur));
    if (cur != device) return fail(JJA_EINVAL, "jja_solve_noalloc: the plan's device is not the current device");
    void* ptr = nullptr;
    JJA_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(JJA_EHIP, "jja_solve_noalloc: no device address for the scratch array");
    g_scratch_ptr[device] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[device];
  return JJA_OK;
}

}  // namespace

extern "C" {

const char* jja_last_error(void) { return g_error.c_str(); }

int jja_plan_validate(int32_t n_problems, const jja_problem_desc* descs) {
  if (n_problems <= 0 || !descs) return fail(JJA_EINVAL, "jja_plan_validate: invalid arguments");
  if (int rc = check_plan_shapes(n_problems, descs)) return rc;
  for (int32_t p = 0; p < n_problems; ++p)
    if (int rc = validate(descs[p])) return rc;
  return JJA_OK;
}

int jja_plan_create(int32_t n_problems, const jja_problem_desc* descs, jja_plan** out) {
  if (!out || n_problems <= 0 || !descs) return fail(JJA_EINVAL, "jja_plan_create: invalid arguments");
  *out = nullptr;
  if (int rc = check_plan_shapes(n_problems, descs)) return rc;
  Plan* plan = new Plan();
  plan->parts = 1;
  for (int32_t p = 0; p < n_problems; ++p) plan->parts = std::max(plan->parts, parts_of(descs[p]));
  plan->pairs = descs[0].column_pairs != 0;
  plan->block_size = descs[0].block_size / parts_of(descs[0]);
  plan->stencil_width = descs[0].stencil_width;
  int rc = JJA_OK;
  hipDeviceProp_t prop;
  if (hipGetDevice(&plan->device) != hipSuccess ||
      hipGetDeviceProperties(&prop, plan->device) != hipSuccess) {
    delete plan;
    return fail(JJA_EHIP, "jja_plan_create: no HIP device");
  }
  plan->lds_limit = static_cast<int64_t>(prop.sharedMemPerBlock);
  plan->n_cu = prop.multiProcessorCount;
  for (int32_t p = 0; p < n_problems && rc == JJA_OK; ++p) {
    const jja_problem_desc& d = descs[p];
    rc = validate(d);
    if (rc) break;
    DevProblem dp{};
    dp.Nj = d.n_junctions;
    dp.Nf = d.n_faces;
    dp.sw = d.stencil_width;
    dp.BS = d.block_size;
    dp.parts = parts_of(d);
    dp.n_face_slices = d.n_face_slices;
    dp.n_solve_slices = d.n_solve_slices;
    dp.max_stage_rows = d.max_stage_rows;
    dp.copy_rows = d.face_copy_rows;
    dp.index_bytes = d.index_bytes;
    dp.val_bytes = static_cast<int32_t>(d.n_entries * static_cast<int64_t>(sizeof(double)));
    dp.cpr_terms = d.cpr_terms;
    for (int k = 0; k < JJA_CPR_MAX_TERMS; ++k) dp.cpr_coef[k] = d.cpr_coef[k];
    if (d.lattice_rows) {
      const int32_t R = d.lattice_rows, C = d.lattice_cols, H = (C + 1) / 2;
      dp.lat_rows = R;
      dp.lat_cols = C;
      dp.lat_es = H + (H & 1);
      dp.lat_ef = d.index_bytes == 4 ? (dp.lat_es + 3) & ~3 : dp.lat_es;
      // a split column solves its tridiagonal systems part by part in LDS
      // when a part's modes and block ends fit
      const int32_t np = parts_of(d);
      if (np > 1) {
        const int32_t mp = 2 * ((H + np - 1) / np);  // mode pairs (i, C-1-i) per part, two slots each
        const int32_t bp = std::max(1, std::min(R, JJA_LATTICE_TRI_THREADS / mp));
        const int64_t need = (static_cast<int64_t>(R) * mp + 4 * static_cast<int64_t>(bp) * mp) * 8;
        if (need <= kSplitLdsMax) {
          dp.lat_mp = mp;
          dp.lat_bp = bp;
          plan->split_lds = std::max(plan->split_lds, need);
        }
      }
      dp.lat_cp = (C + JJA_LATTICE_CHUNK - 1) / JJA_LATTICE_CHUNK * JJA_LATTICE_CHUNK;
      dp.lat_t0 = static_cast<int64_t>(d.n_faces) + (d.n_faces & 1);
      dp.lat_g = d.lattice_g;
      if ((rc = upload(plan, d.lattice_sine, static_cast<size_t>(dp.lat_es) * dp.lat_cp, &dp.lat_sine))) break;
      {
        // mode j of row i at i*CP + (j & 1)*(CP/2) + j/2 (CP is a multiple of 14, so even)
        const size_t cp = dp.lat_cp, half = cp / 2, n = static_cast<size_t>(dp.lat_es) * cp;
        std::vector<double> split(n);
        for (size_t i = 0; i < static_cast<size_t>(dp.lat_es); ++i)
          for (size_t j = 0; j < cp; ++j) split[i * cp + (j & 1) * half + j / 2] = d.lattice_sine[i * cp + j];
        if ((rc = upload(plan, split.data(), n, &dp.lat_sine_p))) break;
      }
      if ((rc = upload(plan, d.lattice_inv, static_cast<size_t>(R) * dp.lat_cp, &dp.lat_inv))) break;
    }
    const size_t Nj = d.n_junctions, Nf = d.n_faces;
    if ((rc = upload(plan, d.Ic, Nj, &dp.Ic))) break;
    if ((rc = upload(plan, d.Rv, Nj, &dp.Rv))) break;
    if ((rc = upload(plan, d.coef, Nj * d.stencil_width, &dp.coef))) break;
    dp.coef_si = 0;  // every coefficient row one value (uniform R and C): read one address
    for (int32_t k = 0; k < d.stencil_width && !dp.coef_si; ++k)
      for (size_t j = 1; j < Nj; ++j)
        if (d.coef[k * Nj + j] != d.coef[k * Nj]) {
          dp.coef_si = 1;
          break;
        }
    {
      std::vector<double> rc0(Nj);
      for (size_t j = 0; j < Nj; ++j) rc0[j] = 1.0 / d.coef[j];
      if ((rc = upload(plan, rc0.data(), Nj, &dp.rc0))) break;
    }
    if ((rc = upload(plan, d.junc_code, 2 * Nj, &dp.junc_code))) break;
    if (Nf) {
      if ((rc = upload(plan, d.face_of_pos, Nf, &dp.face_of_pos))) break;
      if ((rc = upload(plan, d.face_slices, static_cast<size_t>(d.n_face_slices), &dp.face_slices))) break;
      if ((rc = upload(plan, d.solve_slices, static_cast<size_t>(d.n_solve_slices), &dp.solve_slices))) break;
      if ((rc = upload(plan, d.solve_wave_ptr, static_cast<size_t>(d.block_size / kWave + 1), &dp.wave_ptr))) break;
      if ((rc = upload(plan, d.face_code, static_cast<size_t>(d.n_face_entries), &dp.face_code))) break;
      if ((rc = upload(plan, d.slice_rows, static_cast<size_t>(d.n_slice_rows), &dp.slice_rows))) break;
      {
        // face index of every row slot a face record finishes (the flux input)
        std::vector<int32_t> fid(static_cast<size_t>(d.n_slice_rows), -1);
        for (int32_t k = 0; k < d.n_face_slices; ++k) {
          const jja_slice& m = d.face_slices[k];
          if (!(m.flags & JJA_SLICE_FINISH)) continue;
          for (int32_t i = 0; i < (m.flags >> 16); ++i) {
            const int32_t pos = d.slice_rows[m.rowbase + i];
            if (pos >= 0) fid[static_cast<size_t>(m.rowbase) + i] = d.face_of_pos[pos];
          }
        }
        if ((rc = upload(plan, fid.data(), fid.size(), &dp.face_fid))) break;
      }
      if ((rc = upload(plan, d.ent_val, static_cast<size_t>(d.n_entries), &dp.ent_val))) break;
      if (d.index_bytes == 2) {
        const uint16_t* o = nullptr;
        if ((rc = upload(plan, static_cast<const uint16_t*>(d.ent_off), static_cast<size_t>(d.n_index), &o))) break;
        dp.ent_off = o;
      } else {
        const uint32_t* o = nullptr;
        if ((rc = upload(plan, static_cast<const uint32_t*>(d.ent_off), static_cast<size_t>(d.n_index), &o))) break;
        dp.ent_off = o;
      }
    }
    {
      static const double zeros[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      if ((rc = upload(plan, zeros, 8, &dp.zero))) break;
    }
    plan->problems.push_back(dp);
    const int64_t store = (static_cast<int64_t>(d.n_faces) + d.max_stage_rows) * static_cast<int64_t>(sizeof(double));
    plan->lds_bytes.push_back(d.index_bytes == 2 ? store : 0);
    // work of one column step in factor-entry units (a lattice solve: its multiply-adds)
    const int64_t solve_work = d.lattice_rows ? static_cast<int64_t>(d.n_faces) * (d.lattice_cols + 1) : d.n_entries;
    plan->cost.push_back(solve_work + d.n_face_entries + 8 * static_cast<int64_t>(d.n_junctions));
    plan->zbuf_doubles.push_back(d.index_bytes == 2 ? 0 : static_cast<int64_t>(d.n_faces) + d.max_stage_rows);
    plan->max_cols.push_back(d.max_columns);
  }
  if (rc == JJA_OK) {
    const DevProblem* dprob = nullptr;
    rc = upload(plan, plan->problems.data(), plan->problems.size(), &dprob);
    plan->d_problems = const_cast<DevProblem*>(dprob);
  }
  // launch descriptors, work lists and barrier words for the largest runs
  // the descriptors allow: launches allocate nothing
  size_t items = 0, blocks = 0;
  for (size_t p = 0; p < plan->max_cols.size(); ++p) {
    const size_t it = static_cast<size_t>(items_of(plan, plan->max_cols[p]));
    items += it;
    blocks += it * static_cast<size_t>(plan->problems[p].parts);
  }
  auto alloc = [&](size_t bytes, void** out) {
    if (rc != JJA_OK) return;
    void* d = nullptr;
    if (hipMalloc(&d, bytes) != hipSuccess) {
      rc = fail(JJA_ENOMEM, "jja_plan_create: device allocation failed");
      return;
    }
    plan->allocations.push_back(d);
    *out = d;
  };
  alloc(sizeof(jja_problem_run) * n_problems, reinterpret_cast<void**>(&plan->d_runs));
  alloc(sizeof(jja_adjoint_run) * n_problems, reinterpret_cast<void**>(&plan->d_adj));
  alloc(sizeof(WorkItem) * blocks, reinterpret_cast<void**>(&plan->d_work));
  plan->work_cap = blocks;
  if (plan->parts > 1) {
    alloc(2 * sizeof(uint32_t) * items, reinterpret_cast<void**>(&plan->d_ctrs));
    plan->ctr_cap = 2 * items;
  }
  if (rc == JJA_OK && hipEventCreateWithFlags(&plan->done, hipEventDisableTiming) != hipSuccess)
    rc = fail(JJA_EHIP, "jja_plan_create: hipEventCreateWithFlags failed");
  if (rc == JJA_OK && plan->parts > 1) {
    void* d = nullptr;
    if (hipMalloc(&d, 2 * sizeof(int32_t)) != hipSuccess || hipMemset(d, 0, 2 * sizeof(int32_t)) != hipSuccess) {
      rc = fail(JJA_ENOMEM, "jja_plan_create: device allocation failed");
    } else {
      plan->allocations.push_back(d);
      plan->d_err = static_cast<int32_t*>(d);
    }
  }
  if (rc != JJA_OK) {
    free_plan(plan);
    return rc;
  }
  *out = reinterpret_cast<jja_plan*>(plan);
  return JJA_OK;
}

int jja_plan_status(jja_plan* p) {
  Plan* plan = reinterpret_cast<Plan*>(p);
  if (!plan) return fail(JJA_EINVAL, "jja_plan_status: invalid arguments");
  if (!plan->d_err) return JJA_OK;
  int32_t e = 0;
  JJA_HIP(hipMemcpy(&e, plan->d_err, sizeof(e), hipMemcpyDeviceToHost));
  if (!e) return JJA_OK;
  JJA_HIP(hipMemset(plan->d_err, 0, sizeof(int32_t)));
  return fail(JJA_EHIP, "a column barrier of a split launch gave up waiting; its results are invalid");
}

int64_t jja_plan_one_xcd_parts(jja_plan* p) {
  Plan* plan = reinterpret_cast<Plan*>(p);
  if (!plan) return fail(JJA_EINVAL, "jja_plan_one_xcd_parts: invalid arguments");
  if (!plan->d_err) return 0;
  int32_t n = 0;
  JJA_HIP(hipMemcpy(&n, plan->d_err + 1, sizeof(n), hipMemcpyDeviceToHost));
  JJA_HIP(hipMemset(plan->d_err + 1, 0, sizeof(int32_t)));
  return n;
}

int jja_plan_set_wait_limit(jja_plan* p, uint32_t polls) {
  Plan* plan = reinterpret_cast<Plan*>(p);
  if (!plan) return fail(JJA_EINVAL, "jja_plan_set_wait_limit: invalid arguments");
  plan->wait_limit = polls;
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

int64_t jja_plan_zbuf_doubles(const jja_plan* p, int32_t i) {
  const Plan* plan = reinterpret_cast<const Plan*>(p);
  if (!plan || i < 0 || i >= static_cast<int32_t>(plan->zbuf_doubles.size())) return 0;
  return plan->zbuf_doubles[i];
}

int jja_evolve(jja_plan* p, const jja_problem_run* runs, int32_t n_steps, int64_t step0,
               void* stream_ptr) {
  Plan* plan = reinterpret_cast<Plan*>(p);
  if (!plan || !runs || n_steps < 0 || step0 < 0) return fail(JJA_EINVAL, "jja_evolve: invalid arguments");
  if (n_steps == 0) return JJA_OK;
  const int32_t np = static_cast<int32_t>(plan->problems.size());
  std::vector<WorkItem> work, work_g;
  int64_t lds = 0;
  for (int32_t i = 0; i < np; ++i) {
    const jja_problem_run& r = runs[i];
    if (plan->problems[i].Nj == 0)
      return fail(JJA_EINVAL, "jja_evolve: problem " + std::to_string(i) + " is a matrix plan without junctions");
    if (r.W <= 0 || !r.theta) return fail(JJA_EINVAL, "jja_evolve: run needs W > 0 and theta");
    if (r.W > plan->max_cols[i])
      return fail(JJA_EINVAL, "jja_evolve: problem " + std::to_string(i) + " has more columns than its max_columns");
    if (r.dt <= 0) return fail(JJA_EINVAL, "jja_evolve: dt must be positive");
    if (!r.Is.base || !r.Vs.base || !r.f.base || !r.T.base)
      return fail(JJA_EINVAL, "jja_evolve: every source needs a base array");
    if (!r.Vs.zero && !r.theta_s) return fail(JJA_EINVAL, "jja_evolve: non-zero Vs needs theta_s");
    const bool global = plan->zbuf_doubles[i] > 0;
    if (global && !r.zbuf)
      return fail(JJA_EINVAL, "jja_evolve: problem " + std::to_string(i) + " needs a global face store (zbuf)");
    // an LDS face store's junction pass writes u for its face phase (a split
    // plan's columns all use the global store, faceless ones included)
    if (!global && plan->parts == 1 && !r.u)
      return fail(JJA_EINVAL, "jja_evolve: problem " + std::to_string(i) + " (LDS face store) needs u");
    if (!global && plan->lds_bytes[i] > plan->lds_limit)
      return fail(JJA_EUNSUPPORTED, "jja_evolve: face vector of problem " + std::to_string(i) +
                                        " needs " + std::to_string(plan->lds_bytes[i]) +
                                        " bytes of LDS; limit " + std::to_string(plan->lds_limit));
    if (!global) lds = std::max(lds, plan->lds_bytes[i]);
    for (int32_t w = 0; w < items_of(plan, r.W); ++w) (global ? work_g : work).push_back({i, w});
  }
  hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
  if (int rc = claim_stream(plan, stream)) return rc;
  if (int rc = upload_if_changed(plan->d_runs, runs, static_cast<size_t>(np), plan->h_runs, plan->runs_valid, stream))
    return rc;
  int rc = JJA_OK;
  if (plan->parts > 1) {
    // every problem of a split plan has the global face store (validated; a
    // circuit without faces has none and needs none): all columns, one launch
    work.insert(work.end(), work_g.begin(), work_g.end());
    const std::vector<WorkItem> split = split_work(plan, work);
    if ((rc = reset_counters(plan, work.size(), stream))) return rc;
    if ((rc = upload_work(plan, split, stream))) return rc;
    rc = launch_evolve_split(plan, static_cast<int32_t>(split.size()), n_steps, step0, stream);
  } else {
    // one launch: a kernel for one kind of face store, or one that takes either
    const int32_t n_lds = static_cast<int32_t>(work.size()), n_g = static_cast<int32_t>(work_g.size());
    work.insert(work.end(), work_g.begin(), work_g.end());
    order_work(plan, work);
    if ((rc = upload_work(plan, work, stream))) return rc;
    if (lds < 8) lds = 8;
    const int32_t n = n_lds + n_g;
    if (n_g == 0)
      rc = launch_evolve_bs<0>(plan, plan->d_work, n, lds, n_steps, step0, stream);
    else if (n_lds == 0)
      rc = launch_evolve_bs<1>(plan, plan->d_work, n, 8, n_steps, step0, stream);
    else
      rc = launch_evolve_bs<2>(plan, plan->d_work, n, lds, n_steps, step0, stream);
  }
  if (rc) return rc;
  return mark_launched(plan, stream);
}

int jja_adjoint(jja_plan* p, const jja_adjoint_run* runs, int64_t k_hi, int64_t k_lo, void* stream_ptr) {
  Plan* plan = reinterpret_cast<Plan*>(p);
  if (!plan || !runs || k_lo < 0 || k_hi < k_lo) return fail(JJA_EINVAL, "jja_adjoint: invalid arguments");
  if (plan->stencil_width != 3) return fail(JJA_EUNSUPPORTED, "jja_adjoint: stencil width 3 only");
  for (const DevProblem& P : plan->problems)
    if (P.cpr_terms != 0) return fail(JJA_EUNSUPPORTED, "jja_adjoint: the default current-phase relation only");
  const int32_t np = static_cast<int32_t>(plan->problems.size());
  std::vector<WorkItem> work, work_g;
  int64_t lds = 0;
  for (int32_t i = 0; i < np; ++i) {
    const jja_adjoint_run& r = runs[i];
    if (r.W <= 0 || !r.hist || !r.lam || !r.u || !r.grad_Ic || !r.grad_Is || !r.grad_f)
      return fail(JJA_EINVAL, "jja_adjoint: run needs W > 0, hist, lam, u and the gradient arrays");
    if (r.W > plan->max_cols[i])
      return fail(JJA_EINVAL, "jja_adjoint: problem " + std::to_string(i) + " has more columns than its max_columns");
    if (r.hist_m0 > k_lo - 2) return fail(JJA_EINVAL, "jja_adjoint: hist must start at k_lo - 2 or earlier");
    if (r.g && (!r.g_slot || r.g_lo < r.slot_t0 || r.g_hi < r.g_lo))
      return fail(JJA_EINVAL, "jja_adjoint: g needs g_slot covering [g_lo, g_hi)");
    const bool global = plan->zbuf_doubles[i] > 0;
    if (global && !r.zbuf)
      return fail(JJA_EINVAL, "jja_adjoint: problem " + std::to_string(i) + " needs a global face store (zbuf)");
    if (!global && plan->lds_bytes[i] > plan->lds_limit)
      return fail(JJA_EUNSUPPORTED, "jja_adjoint: face vector of problem " + std::to_string(i) + " exceeds LDS");
    if (!global) lds = std::max(lds, plan->lds_bytes[i]);
    for (int32_t w = 0; w < items_of(plan, r.W); ++w) (global ? work_g : work).push_back({i, w});
  }
  hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
  if (int rc = claim_stream(plan, stream)) return rc;
  if (int rc = upload_if_changed(plan->d_adj, runs, static_cast<size_t>(np), plan->h_adj, plan->adj_valid, stream))
    return rc;
  int rc = JJA_OK;
  if (plan->parts > 1) {
    work.insert(work.end(), work_g.begin(), work_g.end());
    const std::vector<WorkItem> split = split_work(plan, work);
    if ((rc = reset_counters(plan, work.size(), stream))) return rc;
    if ((rc = upload_work(plan, split, stream))) return rc;
    rc = launch_adjoint_split(plan, static_cast<int32_t>(split.size()), k_hi, k_lo, stream);
  } else {
    const int32_t n_lds = static_cast<int32_t>(work.size()), n_g = static_cast<int32_t>(work_g.size());
    order_work(plan, work);
    order_work(plan, work_g);
    work.insert(work.end(), work_g.begin(), work_g.end());
    if ((rc = upload_work(plan, work, stream))) return rc;
    if (lds < 8) lds = 8;
    if (n_lds) rc = launch_adjoint_bs<false>(plan, plan->d_work, n_lds, lds, k_hi, k_lo, stream);
    if (!rc && n_g) rc = launch_adjoint_bs<true>(plan, plan->d_work + n_lds, n_g, 8, k_hi, k_lo, stream);
  }
  if (rc) return rc;
  return mark_launched(plan, stream);
}

int jja_solve(jja_plan* p, int32_t i, const double* b, int64_t ldb, double* x, int64_t ldx, int32_t cols,
              double* zbuf, void* stream_ptr) {
  Plan* plan = reinterpret_cast<Plan*>(p);
  if (!plan || i < 0 || i >= static_cast<int32_t>(plan->problems.size()) || cols < 0 || (cols && (!b || !x)))
    return fail(JJA_EINVAL, "jja_solve: invalid arguments");
  const DevProblem& P = plan->problems[i];
  if (P.parts != 1) return fail(JJA_EUNSUPPORTED, "jja_solve: the plan's columns are split over workgroups");
  if (ldb < P.Nf || ldx < P.Nf) return fail(JJA_EINVAL, "jja_solve: leading dimension below n");
  if (cols == 0 || P.Nf == 0) return JJA_OK;
  const bool global = plan->zbuf_doubles[i] > 0;
  if (global && !zbuf) return fail(JJA_EINVAL, "jja_solve: the problem needs a global face store (zbuf)");
  if (!global && plan->lds_bytes[i] > plan->lds_limit)
    return fail(JJA_EUNSUPPORTED, "jja_solve: the face vector exceeds LDS");
  hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
  switch (plan->block_size) {
    case 64: return launch_spd_solve<64>(plan, i, b, ldb, x, ldx, cols, zbuf, stream);
    case 128: return launch_spd_solve<128>(plan, i, b, ldb, x, ldx, cols, zbuf, stream);
    case 256: return launch_spd_solve<256>(plan, i, b, ldb, x, ldx, cols, zbuf, stream);
    case 512: return launch_spd_solve<512>(plan, i, b, ldb, x, ldx, cols, zbuf, stream);
    default: return launch_spd_solve<1024>(plan, i, b, ldb, x, ldx, cols, zbuf, stream);
  }
}

int jja_solve_noalloc(jja_plan* p, int32_t i, const double* b, int64_t ldb, double* x, int64_t ldx,
                      int32_t cols, void* stream_ptr) {
  Plan* plan = reinterpret_cast<Plan*>(p);
  if (!plan || i < 0 || i >= static_cast<int32_t>(plan->problems.size()) || cols < 0)
    return fail(JJA_EINVAL, "jja_solve_noalloc: invalid arguments");
  double* zbuf = nullptr;
  if (plan->zbuf_doubles[i] > 0 && cols > 0) {
    const int64_t need = static_cast<int64_t>(cols) * plan->zbuf_doubles[i];
    if (need > kScratchDoubles)
      return fail(JJA_EINVAL, "jja_solve_noalloc: " + std::to_string(cols) + " columns need " + std::to_string(need) +
                                  " doubles of scratch; the limit is " + std::to_string(kScratchDoubles));
    if (int rc = scratch_ptr(plan->device, &zbuf)) return rc;
  }
  return jja_solve(p, i, b, ldb, x, ldx, cols, zbuf, stream_ptr);
}

int jja_vortex(jja_plan* p, int32_t i, const double* theta, int32_t cols, int32_t* n_out,
               void* stream_ptr) {
  Plan* plan = reinterpret_cast<Plan*>(p);
  if (!plan || i < 0 || i >= static_cast<int32_t>(plan->problems.size()) || !theta || !n_out || cols < 0)
    return fail(JJA_EINVAL, "jja_vortex: invalid arguments");
  if (cols == 0 || plan->problems[i].Nf == 0) return JJA_OK;
  hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
  switch (plan->block_size) {
    case 64: return launch_vortex<64>(plan, i, theta, cols, n_out, stream);
    case 128: return launch_vortex<128>(plan, i, theta, cols, n_out, stream);
    case 256: return launch_vortex<256>(plan, i, theta, cols, n_out, stream);
    case 512: return launch_vortex<512>(plan, i, theta, cols, n_out, stream);
    default: return launch_vortex<1024>(plan, i, theta, cols, n_out, stream);
  }
}

}  // extern "C"
