// Device multifrontal LDL^T for the quasi-definite KKT matrix (gfx950).
//
// Fronts: supernode s owns a dense nr x nr front F_s, lower triangle stored
// packed by columns.  Its first nc columns hold, after the partial
// factorization, L (strictly below the diagonal) and D (diagonal); the
// trailing lower (nr-nc)^2 block holds the update matrix passed to the parent.
//
// Factorization schedule, per level of the assembly tree (leaves first):
//   k_assemble      zero-fill, scatter K entries (+ diagonal bump), extend-add
//                   of the children's update matrices (pull, deterministic)
//   k_factor_small  fronts with nr <= 64, factored whole in LDS
//   k_panel         64-column sub-panel, one workgroup per 64-row chunk below
//                   it: LDL^T of the diagonal block in LDS (every workgroup of
//                   the sub-panel factors it, the first one stores it), then
//                   L21 = R L11^-T D^-1 for its chunk
//   k_trail         lower-triangle trailing update C -= L D L^T on FP64 MFMA
//                   (v_mfma_f64_16x16x4f64), 128x128 tiles
// Triangular solves: one ticketed launch per sweep over the whole tree with
// bounded waits on per-block flags (forward: row blocks, backward: column
// blocks), then an in-block substitution.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <numeric>

#include "kkt_device.hpp"

#include <utility>

namespace cx {

namespace {

constexpr int SMALL_NR = 64;  // fronts up to this size are factored in LDS
constexpr int NB = 64;        // sub-panel width
// outer panel width: the K depth of the big trailing updates; wider for the
// largest fronts, where the pass count over the trailing triangle dominates
inline int panel_width(int nr) { return nr > 32768 ? 1024 : 512; }
constexpr int TM = 128;       // trailing-update tile
constexpr long SMALL_TRAIL_TILES = 512;  // below this many 128-tiles a launch uses 64-tiles
constexpr int KC = 32;        // K chunk staged in LDS per step
constexpr int SB = 64;        // solve block
constexpr int LDSP = NB + 1;  // padded LDS row stride for 64x64 blocks
constexpr long SPIN_LIMIT = 1L << 21;  // ~0.2 s of polling

typedef double v4d __attribute__((ext_vector_type(4)));

// Fronts are stored lower-packed by columns: column j of an nr-row front holds
// rows j..nr-1 and starts colofs(nr, j) elements after the front's base; a
// front takes nr(nr+1)/2 elements.  FP32 and FP64 factors share one buffer
// (the FP32 factor uses its first half).
__host__ __device__ inline long colofs(int nr, int j) { return (long)j * nr - ((long)j * (j - 1)) / 2; }
// element (i, j), i >= j
__host__ __device__ inline long fpos(int nr, int i, int j) { return colofs(nr, j) + (i - j); }

// stats[0]: an exact zero pivot; stats[1]: pivots corrected inline;
// stats[2]: in the checked mode, a pivot with |d| <= eps (or NaN) was met.
// Such a factorization is rejected whatever follows (its minimum |D| is at
// most eps), so the remaining launches of it return at once.
__device__ inline double pivot_fix(double d, double sgn, int mode, double eps, double delta,
                                   int* stats, bool count) {
  if (mode == PIVOT_REGULARIZE) {
    // regularizePivot: src/conicxx/kkt/regularized_ldlt.cpp:132-139
    if (d * sgn >= eps) return d;
    d += sgn * delta;
    if (d * sgn < eps) d = sgn * delta;
    if (count) atomicAdd(&stats[1], 1);
    return d;
  }
  if (d == 0.0 && count) atomicOr(&stats[0], 1);  // QDLDL_factor reports an exact zero pivot
  if (count && !(fabs(d) > eps)) atomicOr(&stats[2], 1);
  return d;
}

// true once a launch of the current factorization has flagged a failing pivot
__device__ inline bool factor_stopped(const int* stats) {
  return __hip_atomic_load(&stats[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

__device__ inline double load_acq(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Cross-workgroup protocol of the solve sweeps.  Values that other
// workgroups of the same launch read (xp, U) are written with agent-coherent
// atomic stores; the writer then waits for its stores to complete
// (s_waitcnt) before it publishes a flag or bumps a counter.  Readers poll the
// flag with relaxed loads and read the values with agent-coherent loads.  No
// agent-scope acquire/release fence is issued (on this GPU each one writes back
// or invalidates the XCD's L2).  Waits are bounded: after SPIN_LIMIT polls, or
// at once when another workgroup has flagged a failure, the wait gives up and
// flags the sweep as failed.  Flags and counters live one per 128-byte line
// (FSTRIDE ints apart).
constexpr int FSTRIDE = 32;

__device__ inline void store_shared(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void stores_complete() {
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ inline bool wait_flag(int* f, int want, int* err) {
  long spins = 0;
  while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != want) {
    if (++spins > SPIN_LIMIT ||
        ((spins & 63) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
      atomicOr(err, 1);
      return false;
    }
    __builtin_amdgcn_s_sleep(2);
  }
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  return true;
}

// Hand-off of a block's values to the one task whose next step waits on
// them: the values are also written to a hand-off buffer that was filled
// with HANDOFF_EMPTY (a signalling-NaN pattern, which no arithmetic result
// has) before the sweep, and the waiting wave polls the values themselves,
// skipping the writer's store-completion wait and the flag round trip.
constexpr long long HANDOFF_EMPTY = 0x7FF4C0C0A5A5DEADLL;

// Wave-wide: every active lane's value at p has been written.  Returns it.
__device__ inline double wait_value(const double* p, bool active, int* err) {
  long spins = 0;
  double v = 0.0;
  for (;;) {
    v = active ? load_acq(p) : 0.0;
    if (!__any(active && __double_as_longlong(v) == HANDOFF_EMPTY)) break;
    if (++spins > SPIN_LIMIT ||
        ((spins & 63) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
      if ((threadIdx.x & 63) == 0) atomicOr(err, 1);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  return v;
}

__device__ inline bool wait_count(int* f, int want, int* err) {
  long spins = 0;
  while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < want) {
    if (++spins > SPIN_LIMIT ||
        ((spins & 63) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
      atomicOr(err, 1);
      return false;
    }
    __builtin_amdgcn_s_sleep(2);
  }
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  return true;
}

// Right-looking LDL^T of an n x n block held row-major in LDS (stride LDSP),
// nc pivot columns; all 256 threads participate.  Pivots at their diagonal
// positions afterwards, L strictly below.
template <typename T>
__device__ void ldl_lds(T* A, int n, int nc, const double* sgn, int mode, double eps,
                        double delta, int* stats, bool count, T* colk) {
  const int tid = threadIdx.x;
  for (int k = 0; k < nc; ++k) {
    if (tid == 0) A[k * LDSP + k] = (T)pivot_fix((double)A[k * LDSP + k], sgn[k], mode, eps, delta, stats, count);
    __syncthreads();
    const T dinv = T(1) / A[k * LDSP + k];
    for (int i = k + 1 + tid; i < n; i += blockDim.x) {
      T c = A[i * LDSP + k];
      colk[i] = c;
      A[i * LDSP + k] = c * dinv;
    }
    __syncthreads();
    // trailing update, thread (row i, column group g): columns k+1+g, k+5+g, ...
    for (int i = k + 1 + (tid & 63); i < n; i += 64) {
      const T lik = A[i * LDSP + k];
      for (int j = k + 1 + (tid >> 6); j <= i; j += 4) A[i * LDSP + j] -= lik * colk[j];
    }
    __syncthreads();
  }
}

// Wave-level broadcast of lane l's value.
template <typename T>
__device__ inline T lane_bcast(T v, int l);
template <>
__device__ inline double lane_bcast<double>(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}
template <>
__device__ inline float lane_bcast<float>(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Step K of ldl_wave (K a template parameter, so that every register index
// and loop bound below is a compile-time constant).
template <typename T, int K>
__device__ __forceinline__ void ldl_wave_step(T (&a)[64], int nc, double sg, int lane, int mode, double eps,
                                              double delta, int* stats, bool count, T* colk) {
  if (K >= nc) return;
  const T d = (T)pivot_fix((double)lane_bcast(a[K], K), lane_bcast(sg, K), mode, eps, delta, stats,
                           count && lane == 0);
  const T c = a[K];  // unscaled entry (lane, K)
  colk[lane] = c;
  __builtin_amdgcn_wave_barrier();
  const T l = c / d;
  a[K] = (lane == K) ? d : (lane > K ? l : a[K]);
  // Unpredicated: lane i only needs j <= i, and for j > i the registers hold
  // entries above the diagonal that are never stored.  Lanes <= K update only
  // such entries (with values read from their own upper registers), so
  // nothing they compute reaches the factor.
  // colk[j] (entry (j, K) = l_jK d_K) is read in groups of 16, the reads of
  // the next group issued before the FMAs of the current one; the scheduling
  // barriers keep the reads from being sunk next to their uses, where the LDS
  // latency would be exposed at every pair of columns.
  constexpr int J = K + 1;
  T cb[2][16];
#pragma unroll
  for (int q = 0; q < 16; ++q)
    if (J + q < 64) cb[0][q] = colk[J + q];
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const int j0 = J + 16 * g;
    if (j0 < 64) {
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (j0 + 16 + q < 64) cb[(g + 1) & 1][q] = colk[j0 + 16 + q];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (j0 + q < 64) a[j0 + q] -= l * cb[g & 1][q];
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  __builtin_amdgcn_wave_barrier();  // every read of colk before the next write
}

template <typename T, int... Ks>
__device__ __forceinline__ void ldl_wave_steps(std::integer_sequence<int, Ks...>, T (&a)[64], int nc, double sg,
                                               int lane, int mode, double eps, double delta, int* stats,
                                               bool count, T* colk) {
  (ldl_wave_step<T, Ks>(a, nc, sg, lane, mode, eps, delta, stats, count, colk), ...);
}

// pivot_fix with its statistics accumulated in registers (written once, by
// ldl_wave_full, after the last step): no branch or atomic in a step
struct PivotAcc {
  int zero = 0, stop = 0, nreg = 0;
};
__device__ inline double pivot_rule(double d, double sgn, int mode, double eps, double delta, PivotAcc& pa) {
  if (mode == PIVOT_REGULARIZE) {  // regularizePivot: src/conicxx/kkt/regularized_ldlt.cpp:132-139
    const bool ok = d * sgn >= eps;
    double dd = d + sgn * delta;
    dd = (dd * sgn < eps) ? sgn * delta : dd;
    pa.nreg += ok ? 0 : 1;
    return ok ? d : dd;
  }
  pa.zero |= d == 0.0 ? 1 : 0;
  pa.stop |= !(fabs(d) > eps) ? 1 : 0;
  return d;
}

// The 64 steps of a full block (n = nc = 64) with the same operations, in
// the same order per entry, as ldl_wave_step, software-pipelined for latency.
// Step K is split into
//   A: the pivot d (readlane of lane K, pivot_rule) and the multiplier
//      l = c / d, with c the lane's entry (lane, K) and the next two columns'
//      entries c1, c2 taken from lanes K+1, K+2 by readlane;
//   B: the entries of columns K+1 and K+2 (the next pivots' columns), and c
//      written to one of two LDS buffers for the far columns;
//   D: the far columns K+3..63, reading c from LDS in chunks of 16, the next
//      chunk's reads issued before the current chunk's FMAs.
// They are emitted A0 B0 | A1 D0 B1 | A2 D1 B2 | ...: step K+1's pivot and
// division (which need only B of step K) run while step K's far FMAs issue.
// Every entry still takes its updates in step order (column K+2 takes D of
// step K-1 before B of step K), so the factor is the same.
template <typename T>
struct LStep {
  T d, c, c1, c2, l;
};

// An empty asm statement with the value as a register in/out operand: no
// instruction, but the value must exist at this point, so the compiler cannot
// sink the update that produced it down to its last use (which would keep
// every step's multiplier and column values live at once).
template <typename T>
__device__ __forceinline__ void pin(T& v) {
  __asm__ volatile("" : "+v"(v));
}

template <typename T, int K>
__device__ __forceinline__ void lstep_a(T (&a)[64], double sg, int mode, double eps, double delta, PivotAcc& pa,
                                        LStep<T>& s) {
  s.d = (T)pivot_rule((double)lane_bcast(a[K], K), lane_bcast(sg, K), mode, eps, delta, pa);
  s.c = a[K];
  s.c1 = s.c2 = T(0);
  if constexpr (K + 1 < 64) s.c1 = lane_bcast(s.c, K + 1);
  if constexpr (K + 2 < 64) s.c2 = lane_bcast(s.c, K + 2);
  s.l = s.c / s.d;
  pin(s.l);  // the division stays in the region of the previous step's far FMAs
}

// B of step K also publishes column K of L: the multipliers go to the
// block's LDS rows (A, stride LDSP), and every fourth step prog becomes K + 1,
// so a wave running the forward substitution of the rows below the block
// (which needs row i of L complete, i.e. steps 0..i-1) can follow the
// factorization a few rows behind.  Both are LDS stores of this one wave,
// which the LDS executes in issue order, so a reader that sees prog = K + 1
// sees columns 0..K; the signal fence keeps the compiler from reordering the
// two stores (a release store would also wait for this wave's LDS reads in
// flight, which the software pipelining keeps busy).
template <typename T, int K>
__device__ __forceinline__ void lstep_b(T (&a)[64], int lane, const LStep<T>& s, T* colk, T* A, int* prog) {
  a[K] = (lane == K) ? s.d : (lane > K ? s.l : a[K]);
  pin(a[K]);
  A[lane * LDSP + K] = a[K];  // (lanes above the diagonal write entries nobody reads)
  if constexpr ((K & 3) == 3) {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (lane == 0) __hip_atomic_store(prog, K + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  if constexpr (K + 1 < 64) a[K + 1] -= s.l * s.c1, pin(a[K + 1]);
  if constexpr (K + 2 < 64) a[K + 2] -= s.l * s.c2, pin(a[K + 2]);
  if constexpr (K + 3 < 64) colk[(K & 1) * 64 + lane] = s.c;
}

// Group K (1 <= K < 64): A of step K, D of step K-1, B of step K.  The
// first chunk of D's column values was read at the end of the previous group
// (v0), so A's division shares its scheduling region with FMAs whose
// operands are already in registers; the group ends by reading the first
// chunk of its own step's far columns for the next group.
template <typename T, int K>
__device__ __forceinline__ void lstep_group(T (&a)[64], double sg, int lane, int mode, double eps, double delta,
                                            PivotAcc& pa, T* colk, LStep<T>& prev, T (&v0)[16], T* A,
                                            int* prog) {
  constexpr int J = K + 2;  // first far column of step K-1
  const T* cb = colk + ((K - 1) & 1) * 64;
  LStep<T> cur;
  lstep_a<T, K>(a, sg, mode, eps, delta, pa, cur);
  if constexpr (J < 64) {
    T v[2][16];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[0][q] = v0[q];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int j0 = J + 16 * g;
      if (j0 < 64) {
#pragma unroll
        for (int q = 0; q < 16; ++q)
          if (j0 + 16 + q < 64) v[(g + 1) & 1][q] = cb[j0 + 16 + q];
#pragma unroll
        for (int q = 0; q < 16; ++q)
          if (j0 + q < 64) a[j0 + q] -= prev.l * v[g & 1][q], pin(a[j0 + q]);
        // one chunk in flight at a time
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  lstep_b<T, K>(a, lane, cur, colk, A, prog);
  prev = cur;
  if constexpr (K + 3 < 64) {
    const T* nb = colk + (K & 1) * 64;
#pragma unroll
    for (int q = 0; q < 16; ++q)
      if (K + 3 + q < 64) v0[q] = nb[K + 3 + q];
  }
  __builtin_amdgcn_sched_barrier(0);
}

template <typename T, int... Ks>
__device__ __forceinline__ void lstep_groups(std::integer_sequence<int, Ks...>, T (&a)[64], double sg, int lane,
                                             int mode, double eps, double delta, PivotAcc& pa, T* colk,
                                             LStep<T>& prev, T (&v0)[16], T* A, int* prog) {
  (lstep_group<T, Ks + 1>(a, sg, lane, mode, eps, delta, pa, colk, prev, v0, A, prog), ...);
}

template <typename T>
__device__ __forceinline__ void ldl_full(T (&a)[64], double sg, int lane, int mode, double eps, double delta,
                                         PivotAcc& pa, T* colk, T* A, int* prog) {
  LStep<T> prev;
  T v0[16];
  lstep_a<T, 0>(a, sg, mode, eps, delta, pa, prev);
  lstep_b<T, 0>(a, lane, prev, colk, A, prog);
#pragma unroll
  for (int q = 0; q < 16; ++q) v0[q] = colk[3 + q];
  __builtin_amdgcn_sched_barrier(0);
  lstep_groups<T>(std::make_integer_sequence<int, 63>{}, a, sg, lane, mode, eps, delta, pa, colk, prev, v0, A,
                  prog);
}

// The same partial LDL^T as ldl_lds for n <= 64 rows, by one wave with no
// workgroup barrier: lane i keeps row i in registers, pivot k is read from
// lane k, and the unscaled column k goes through LDS (colk) so that every lane
// reads entry j of it as a broadcast.  Called by the first wave only (threads
// 0..63); the caller synchronizes the workgroup before and after.  A full
// block (n = nc = 64) publishes each column of L as it is computed (prog =
// number of finished steps, in LDS; see lstep_b).
template <typename T>
__device__ __forceinline__ void ldl_wave(T* A, int n, int nc, const double* sgn, int mode, double eps, double delta,
                         int* stats, bool count, T* colk, int* prog) {
  const int lane = threadIdx.x & 63;
  const double sg = lane < nc ? sgn[lane] : 1.0;  // pivot signs, read from lane k at step k
  // each path loads and stores the block itself: no register array lives
  // across the branch
  if (n == 64 && nc == 64) {
    // a full block: colk holds two 64-entry buffers
    T a[64];
#pragma unroll
    for (int j = 0; j < 64; ++j) a[j] = (j <= lane) ? A[lane * LDSP + j] : T(0);
    PivotAcc pa;
    ldl_full<T>(a, sg, lane, mode, eps, delta, pa, colk, A, prog);
    if (count && lane == 0) {
      if (pa.zero) atomicOr(&stats[0], 1);
      if (pa.nreg) atomicAdd(&stats[1], pa.nreg);
      if (pa.stop) atomicOr(&stats[2], 1);
    }
#pragma unroll
    for (int j = 0; j < 64; ++j)
      if (j <= lane) A[lane * LDSP + j] = a[j];
  } else {
    T a[64];
#pragma unroll
    for (int j = 0; j < 64; ++j) a[j] = (lane < n && j <= lane) ? A[lane * LDSP + j] : T(0);
    ldl_wave_steps<T>(std::make_integer_sequence<int, 64>{}, a, nc, sg, lane, mode, eps, delta, stats, count, colk);
#pragma unroll
    for (int j = 0; j < 64; ++j)
      if (lane < n && j <= lane) A[lane * LDSP + j] = a[j];
  }
}

// Forward substitution with the unit-lower part of a 64 x 64 block held
// row-major in LDS (stride LDSP), one right-hand side per lane:
// y_i = b_i - sum_{t<i} L(i, t) y_t, four accumulator chains per row (t mod 4).
// The row entries L(i, t), which do not depend on y, are read in chunks of IC
// (row i takes ceil(i/IC) chunks), each chunk's reads issued before the FMAs
// of the previous one, across row boundaries; chunk indices, rows and register
// indices are compile-time constants.  Chunks of 8 keep the substitution wave
// within the register file while it waits for rows of L (16 spilled there).
constexpr int IC = 8;  // entries of a row of L per chunk
constexpr int inv_row_chunks(int i) { return (i + IC - 1) / IC; }
constexpr int inv_chunk_row(int c) {
  int i = 1;
  while (c >= inv_row_chunks(i)) c -= inv_row_chunks(i++);
  return i;
}
constexpr int inv_chunk_grp(int c) {
  int i = 1;
  while (c >= inv_row_chunks(i)) c -= inv_row_chunks(i++);
  return c;
}
constexpr int inv_nchunks() {
  int n = 0;
  for (int i = 1; i < 64; ++i) n += inv_row_chunks(i);
  return n;
}

template <typename T, int C, typename W>
__device__ __forceinline__ void inv_chunk(const T* L, const T* b, T (&y)[64], T (&acc)[4], T (&cb)[2][IC],
                                          W& wait_row) {
  constexpr int I = inv_chunk_row(C), G = inv_chunk_grp(C), T0 = IC * G;
  if constexpr (G == 0) {
    acc[0] = b[I];
    acc[1] = acc[2] = acc[3] = T(0);
  }
  if constexpr (C + 1 < inv_nchunks()) {
    constexpr int I1 = inv_chunk_row(C + 1), T1 = IC * inv_chunk_grp(C + 1);
    if constexpr (I1 != I) {
      // the accumulators must exist before the wait: otherwise the compiler
      // sinks the arithmetic below every wait and holds all of L's reads
#pragma unroll
      for (int j = 0; j < 4; ++j) pin(acc[j]);
      wait_row(I1);
    }
#pragma unroll
    for (int q = 0; q < IC; ++q)
      if (T1 + q < I1) cb[(C + 1) & 1][q] = L[I1 * LDSP + T1 + q];
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int q = 0; q < IC; ++q)
    if (T0 + q < I) acc[(T0 + q) & 3] -= cb[C & 1][q] * y[T0 + q];
  __builtin_amdgcn_sched_barrier(0);
  if constexpr (T0 + IC >= I) y[I] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
}

template <typename T, typename W, int... Cs>
__device__ __forceinline__ void inv_chunks(std::integer_sequence<int, Cs...>, const T* L, const T* b, T (&y)[64],
                                           T (&acc)[4], T (&cb)[2][IC], W& wait_row) {
  (inv_chunk<T, Cs>(L, b, y, acc, cb, wait_row), ...);
}

// b: the lane's right-hand side (64 entries, LDS).  wait_row(i) returns once
// row i of L is complete in LDS (it is called before the first read of each
// row, rows in increasing order).
template <typename T, typename W>
__device__ __forceinline__ void unit_lower_solve(const T* L, const T* b, T (&y)[64], W& wait_row) {
  T acc[4], cb[2][IC];
  y[0] = b[0];
  wait_row(1);
  cb[0][0] = L[1 * LDSP + 0];  // chunk 0: row 1, t = 0
  inv_chunks<T>(std::make_integer_sequence<int, inv_nchunks()>{}, L, b, y, acc, cb, wait_row);
}

// ---------------------------------------------------------------------------
// assembly: zero-fill + K scatter + extend-add of children update matrices
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_assemble(
    const AsmTask* __restrict__ tasks, const AsmSeg* __restrict__ segs, const long* __restrict__ fofs,
    const int* __restrict__ nrv, const int* __restrict__ ncv, const int* __restrict__ firstv,
    const int* __restrict__ kcolptr, const int* __restrict__ ksrc, const long* __restrict__ kdst,
    const signed char* __restrict__ kdiag, const double* __restrict__ Kx, double extra_p,
    double extra_a, const int* __restrict__ rmap_ofs, const int* __restrict__ rmap, T* F, const int* stats) {
  if (factor_stopped(stats)) return;
  const AsmTask Tk = tasks[blockIdx.x];
  const int s = Tk.s, nr = nrv[s], nc = ncv[s];
  const long base = fofs[s];
  const int tid = threadIdx.x;
  // 1. zero the task's columns
  const long ncolw = colofs(nr, Tk.j1) - colofs(nr, Tk.j0);
  T* Fc = F + base + colofs(nr, Tk.j0);
  for (long idx = tid; idx < ncolw; idx += blockDim.x) Fc[idx] = T(0);
  __syncthreads();
  // 2. original K entries of the pivot columns in range (distinct destinations)
  const int jend = min(Tk.j1, nc);
  if (Tk.j0 < jend) {
    const int e0 = kcolptr[firstv[s] + Tk.j0], e1 = kcolptr[firstv[s] + jend];
    for (int e = e0 + tid; e < e1; e += blockDim.x) {
      double v = Kx[ksrc[e]];
      signed char dg = kdiag[e];
      if (dg > 0) v += extra_p;
      if (dg < 0) v -= extra_a;
      F[kdst[e]] += (T)v;
    }
  }
  __syncthreads();
  // 3. children update matrices, one child at a time
  const int lane = tid & 63, wv = tid >> 6;
  for (int g = Tk.seg_beg; g < Tk.seg_end; ++g) {
    const AsmSeg S = segs[g];
    const int c = S.child, nrc = nrv[c], ncc = ncv[c];
    const int nu = nrc - ncc;
    const int* rm = rmap + rmap_ofs[c];
    // a wave per child column, or, for the few columns a task of a large
    // front gets, the whole workgroup along each column
    const bool wide = S.i1 - S.i0 < 4;
    const int i0 = wide ? S.i0 : S.i0 + wv, istep = wide ? 1 : 4;
    const int t0 = wide ? tid : lane, tstep = wide ? 256 : 64;
    for (int i = i0; i < S.i1; i += istep) {
      const int pc = rm[i];
      T* Fp = F + base + colofs(nr, pc) - pc;  // Fp[row] = element (row, pc)
      // Ucol[i'] = element (ncc + i', ncc + i) of the child, i' >= i
      const T* Ucol = F + fofs[c] + colofs(nrc, ncc + i) - i;
      for (int ip = i + t0; ip < nu; ip += tstep) Fp[rm[ip]] += Ucol[ip];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// small fronts: whole front in LDS
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_factor_small(
    const int* __restrict__ list, const long* __restrict__ fofs, const int* __restrict__ nrv,
    const int* __restrict__ ncv, const int* __restrict__ firstv, const double* __restrict__ sgn,
    T* F, double* Dg, int mode, double eps, double delta, int* stats) {
  if (factor_stopped(stats)) return;
  __shared__ T A[SMALL_NR * LDSP];
  __shared__ T colk[SMALL_NR];
  const int s = list[blockIdx.x];
  const int nr = nrv[s], nc = ncv[s], first = firstv[s];
  const long base = fofs[s];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < nr * nr; idx += blockDim.x) {
    int col = idx / nr, row = idx % nr;
    if (row >= col) A[row * LDSP + col] = F[base + fpos(nr, row, col)];
  }
  __syncthreads();
  ldl_lds(A, nr, nc, sgn + first, mode, eps, delta, stats, true, colk);
  for (int idx = tid; idx < nr * nr; idx += blockDim.x) {
    int col = idx / nr, row = idx % nr;
    if (row >= col) F[base + fpos(nr, row, col)] = A[row * LDSP + col];
  }
  for (int k = tid; k < nc; k += blockDim.x) Dg[first + k] = A[k * LDSP + k];
}

// ---------------------------------------------------------------------------
// sub-panels
// ---------------------------------------------------------------------------
constexpr int PANEL_ROWS = 64;

// Sub-panel [a, a+w) of a front: LDL^T of the w x w diagonal block, then the
// rows [r0, r1) below it, L21 = R L11^-T D^-1.  Every workgroup of the
// sub-panel factors the (assembled, updated) diagonal block itself, with the
// same deterministic code, so no launch separates the two steps; the one
// whose chunk starts right below the block (r0 == a + w; a task with no rows
// exists for a sub-panel at the bottom of its front) stores it, its D and its
// pivot statistics.
template <typename T>
__global__ __launch_bounds__(256) void k_panel(
    const PanelTask* __restrict__ tasks, const long* __restrict__ fofs, const int* __restrict__ nrv,
    const int* __restrict__ firstv, const double* __restrict__ sgn, T* F, T* Fd, double* Dg, int mode,
    double eps, double delta, int* stats) {
  if (factor_stopped(stats)) return;
  // one block in a fixed order, the broadcast column first: its reads then
  // take immediate LDS offsets (below 64 KB) instead of an address register each
  struct Lds {
    T colk[2 * NB];
    T dinv[NB];
    T L11[NB * LDSP];
    T X[PANEL_ROWS * LDSP];
    int prog;  // finished LDL^T steps of a full diagonal block
  };
  __shared__ Lds sm;
  T* colk = sm.colk;
  T* dinv = sm.dinv;
  T* L11 = sm.L11;
  T* X = sm.X;
  const PanelTask Tk = tasks[blockIdx.x];
  const int s = Tk.s, nr = nrv[s], first = firstv[s], a = Tk.a, w = Tk.w;
  const long base = fofs[s];
  const int tid = threadIdx.x;
  const bool owner = Tk.r0 == a + w;
  const int rows = Tk.r1 - Tk.r0;
  if (tid == 0) sm.prog = 0;
  {
    // lane = row, wave g = columns g, g+4, ...: all 32 loads of a thread are
    // issued from clamped (always valid, lower-triangle) addresses before any
    // is stored, then the ones in range go to LDS
    const int li = tid & 63, g = tid >> 6;
    const int iL = min(li, w - 1);                                   // diagonal-block row
    const int rX = rows > 0 ? Tk.r0 + min(li, rows - 1) : a + w - 1;  // chunk row, below every column
    T vl[16], vx[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int j = min(g + 4 * q, w - 1);
      const long c0 = base + colofs(nr, a + j) - (a + j);  // + row = element (row, a + j)
      vl[q] = F[c0 + a + max(iL, j)];
      vx[q] = F[c0 + rX];
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int j = g + 4 * q;
      if (j < w && li < w && li >= j) L11[li * LDSP + j] = vl[q];
      if (j < w && li < rows) X[li * LDSP + j] = vx[q];
    }
  }
  __syncthreads();
  // Wave 1 runs the forward substitution of the chunk's rows (below) while
  // wave 0 factors the diagonal block: a full 64-column block publishes each
  // row of L as soon as it is final and the substitution follows a row
  // behind; a narrower block publishes all rows at the end.  The arithmetic
  // is that of the sequential order.
  // (the wave index in a scalar register: uniform branches, so the two
  // waves' register sets are allocated as separate paths)
  const int wvu = __builtin_amdgcn_readfirstlane(tid >> 6);
  if (wvu == 0) {
    ldl_wave(L11, w, w, sgn + first + a, mode, eps, delta, stats, owner, colk, &sm.prog);
    if (tid == 0) __hip_atomic_store(&sm.prog, NB, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  } else if (wvu == 1 && rows > 0) {
    const int lane = tid & 63;
    int seen = 0;
    auto wait_row = [&](int i) __attribute__((always_inline)) {
      if (i > seen) {
        // wave 0 always runs to its final store of NB; the spin limit only
        // bounds a broken invariant.  The count is read into a scalar
        // register: a wave-uniform loop.
        int v;
        for (int spins = 0; spins < (int)SPIN_LIMIT; ++spins) {
          v = __builtin_amdgcn_readfirstlane(
              __hip_atomic_load(&sm.prog, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
          if (v >= i) break;
          __builtin_amdgcn_s_sleep(1);
        }
        seen = v;
      }
    };
    T y[64];
    unit_lower_solve(L11, X + lane * LDSP, y, wait_row);
    __builtin_amdgcn_wave_barrier();  // the lane's right-hand side is read: its row takes the result
#pragma unroll
    for (int k = 0; k < 64; ++k) X[lane * LDSP + k] = y[k];
  }
  __syncthreads();
  if (owner) {
    // other workgroups of this launch may still be reading the unfactored
    // block from F: the factored one is staged and stored by k_diag_store
    T* dst = Fd + (long)Tk.blk * NB * NB;
    for (int idx = tid; idx < w * w; idx += blockDim.x) {
      int j = idx / w, i = idx % w;
      if (i >= j) dst[idx] = L11[i * LDSP + j];
    }
    for (int k = tid; k < w; k += blockDim.x) Dg[first + a + k] = L11[k * LDSP + k];
  }
  if (rows <= 0) return;
  for (int k = tid; k < w; k += blockDim.x) dinv[k] = T(1) / L11[k * LDSP + k];
  __syncthreads();
  // L21 = R L11^-T D^-1 by forward substitution, lane = row r of the chunk:
  // x_k = R(r, k) - sum_{t<k} L11(k, t) x_t, L21(r, k) = x_k / d_k.  A
  // triangular solve rather than a product with an explicit inverse: on the
  // badly scaled diagonal blocks of late interior-point iterations the
  // inverse loses accuracy that the refinement cannot recover.  Columns at or
  // past w (and lanes past the chunk's rows) compute unused values.

  // column by column, lanes along the rows: coalesced stores by all waves
  for (int idx = tid; idx < rows * w; idx += blockDim.x) {
    const int k = idx / rows, rr = idx - k * rows;
    F[base + fpos(nr, Tk.r0 + rr, a + k)] = X[rr * LDSP + k] * dinv[k];
  }
}

// stores every staged diagonal block into its front, one workgroup per block
template <typename T>
__global__ __launch_bounds__(256) void k_diag_store(const PanelTask* __restrict__ slots,
                                                    const long* __restrict__ fofs,
                                                    const int* __restrict__ nrv, const T* __restrict__ Fd,
                                                    T* F, const int* stats) {
  if (factor_stopped(stats)) return;
  const PanelTask Tk = slots[blockIdx.x];
  const int s = Tk.s, nr = nrv[s], a = Tk.a, w = Tk.w;
  const long base = fofs[s];
  const T* src = Fd + (long)Tk.blk * NB * NB;
  for (int idx = threadIdx.x; idx < w * w; idx += blockDim.x) {
    int j = idx / w, i = idx % w;
    if (i >= j) F[base + fpos(nr, a + i, a + j)] = src[idx];
  }
}

// ---------------------------------------------------------------------------
// trailing update on MFMA:  C[i][j] -= sum_k L[i][k] D[k] L[j][k]
// FP64 fronts use v_mfma_f64_16x16x4f64, FP32 fronts v_mfma_f32_16x16x4f32;
// both take a-lane l -> A[l%16][l/16] and b-lane l -> B[l/16][l%16]; their
// accumulator layouts differ (row of register r: l/16 + 4r for f64,
// 4(l/16) + r for f32).
// ---------------------------------------------------------------------------
typedef float v4f __attribute__((ext_vector_type(4)));
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef v4d acc;
  static __device__ inline acc op(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
  typedef v4f acc;
  static __device__ inline acc op(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// TT: square tile (128, or 64 for launches with too few 128-tiles to fill the
// GPU: four times the workgroups on the same entries, each entry's
// accumulation in the same order).  WN: waves across the tile columns (TT =
// 128: 2 gives four waves of 64 x 64, 4 eight waves of 64 x 32, half the
// accumulator registers per wave and twice the waves to hide latency; TT =
// 64: 4 gives four waves of 64 x 16).  TT / 64 wave rows of 64.
// (the second launch bound is the minimum number of waves per SIMD)
template <typename T, int WN, int TT = TM>
__global__ __launch_bounds__(TT * WN, (WN == 4 || sizeof(T) == 4) ? 4 : 2) void k_trail(
    const TrailTask* __restrict__ tasks, const long* __restrict__ fofs, const int* __restrict__ nrv,
    const int* __restrict__ firstv, const double* __restrict__ Dg, T* F, const int* stats) {
  if (factor_stopped(stats)) return;
  typedef typename Mfma<T>::acc acc_t;
  constexpr int NT = TT * WN;           // threads
  constexpr int NI = TT / 16 / WN;      // 16-column MFMA blocks per wave
  static_assert(NI >= 1, "a wave covers at least one 16-column block");
  constexpr int LDA = TT + 4;
  constexpr int PER = KC * TT / NT;     // staged elements per thread and operand
  __shared__ T As[KC * LDA];
  __shared__ T Bs[KC * LDA];
  __shared__ T Ds[KC];
  const TrailTask Tk = tasks[blockIdx.x];
  const int s = Tk.s, nr = nrv[s], first = firstv[s];
  const long base = fofs[s];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wr = (wv / WN) * 64, wc = (wv % WN) * (16 * NI);  // wave tile origin within the block tile
  // a wave whose 64 x 16NI block holds no stored element (columns past chi,
  // rows past the front, or entirely above the diagonal or below rlo) only
  // helps stage the operands
  const int wi0 = Tk.ti + wr, wj0 = Tk.tj + wc;
  const bool wact = wj0 < Tk.chi && wi0 < nr && wi0 + 63 >= wj0 && wi0 + 63 >= Tk.rlo;
  acc_t acc[4][NI];
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) acc[mi][ni] = (acc_t){0, 0, 0, 0};
  // Staging: a thread loads 16 bytes of consecutive rows (VR of them) of a
  // column at a time.  Rows of the staged operands lie below the staged
  // columns; rows past the front run on into the next columns' storage (the
  // factor buffer is padded past its last front for the tiles at its end)
  // and are zeroed on the way into LDS, as are columns past the K range.
  // Register prefetch of the next chunk, so all loads of a chunk are in
  // flight together while the MFMAs run.  D of the chunk goes to LDS too and
  // scales the B fragments as they are read.
  constexpr int VR = 16 / (int)sizeof(T);  // rows per load
  constexpr int RG = TT / VR;              // row groups of a column
  constexpr int CS = NT / RG;              // column stride of a thread's loads
  constexpr int NQ = KC / CS;              // loads per operand and chunk
  static_assert(NQ * VR == PER, "staging covers the chunk");
  typedef T vrow_t __attribute__((ext_vector_type(VR), aligned(sizeof(T))));
  const int rg0 = (tid % RG) * VR;  // first staged row of the thread
  const int lk = tid / RG;          // first staged column
  const T* Fa = F + base + Tk.ti + rg0;
  const T* Fb = F + base + Tk.tj + rg0;
  vrow_t ra[NQ], rb[NQ];
  T dn = T(0);
  auto load_chunk = [&](int k0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int kq = min(k0 + lk + CS * q, Tk.k1 - 1);
      const long col = colofs(nr, kq) - kq;
      ra[q] = *reinterpret_cast<const vrow_t*>(Fa + col);
      rb[q] = *reinterpret_cast<const vrow_t*>(Fb + col);
    }
    if (tid < KC) dn = (T)Dg[first + min(k0 + tid, Tk.k1 - 1)];
  };
  load_chunk(Tk.k0);
  for (int k0 = Tk.k0; k0 < Tk.k1; k0 += KC) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int kk = lk + CS * q;
      const bool kin = k0 + kk < Tk.k1;
      vrow_t va, vb;
#pragma unroll
      for (int e = 0; e < VR; ++e) {
        va[e] = (kin && Tk.ti + rg0 + e < nr) ? ra[q][e] : T(0);
        vb[e] = (kin && Tk.tj + rg0 + e < nr) ? rb[q][e] : T(0);
      }
      *reinterpret_cast<vrow_t*>(&As[kk * LDA + rg0]) = va;
      *reinterpret_cast<vrow_t*>(&Bs[kk * LDA + rg0]) = vb;
    }
    if (tid < KC) Ds[tid] = dn;
    __syncthreads();
    if (k0 + KC < Tk.k1) load_chunk(k0 + KC);  // in flight while the MFMAs run
    if (wact) {
#pragma unroll
      for (int kk = 0; kk < KC; kk += 4) {
        T af[4], bf[NI];
        const int kr = kk + (lane >> 4);
        const T dk = Ds[kr];
#pragma unroll
        for (int mi = 0; mi < 4; ++mi) af[mi] = As[kr * LDA + wr + mi * 16 + (lane & 15)];
#pragma unroll
        for (int ni = 0; ni < NI; ++ni) bf[ni] = Bs[kr * LDA + wc + ni * 16 + (lane & 15)] * dk;
#pragma unroll
        for (int mi = 0; mi < 4; ++mi)
#pragma unroll
          for (int ni = 0; ni < NI; ++ni) acc[mi][ni] = Mfma<T>::op(af[mi], bf[ni], acc[mi][ni]);
      }
    }
    __syncthreads();
  }
  // write back.  A tile whose entries are all stored (strictly below the
  // diagonal, inside the front and the column range) skips the per-entry
  // checks: one address per row block and column, the lane's four rows at
  // fixed offsets from it; in FP32 they are consecutive rows, read and
  // written as one 16-byte vector.
  const bool interior = Tk.ti + TT <= nr && Tk.tj + TT <= Tk.chi && Tk.ti >= Tk.tj + TT && Tk.ti >= Tk.rlo;
  if (interior) {
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
      T* adr[NI];
#pragma unroll
      for (int ni = 0; ni < NI; ++ni) {
        const int j = Tk.tj + wc + ni * 16 + (lane & 15);
        const int i0 = Tk.ti + wr + mi * 16 + Mfma<T>::row(lane, 0);
        adr[ni] = F + base + fpos(nr, i0, j);
      }
      if constexpr (sizeof(T) == 4) {
        typedef T v4u __attribute__((ext_vector_type(4), aligned(sizeof(T))));
        v4u old[NI];
#pragma unroll
        for (int ni = 0; ni < NI; ++ni) old[ni] = *reinterpret_cast<const v4u*>(adr[ni]);
#pragma unroll
        for (int ni = 0; ni < NI; ++ni) {
          v4u nv;
#pragma unroll
          for (int r = 0; r < 4; ++r) nv[r] = old[ni][r] - acc[mi][ni][r];
          *reinterpret_cast<v4u*>(adr[ni]) = nv;
        }
      } else {
        T old[NI][4];
#pragma unroll
        for (int ni = 0; ni < NI; ++ni)
#pragma unroll
          for (int r = 0; r < 4; ++r) old[ni][r] = adr[ni][Mfma<T>::row(lane, r) - Mfma<T>::row(lane, 0)];
#pragma unroll
        for (int ni = 0; ni < NI; ++ni)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            adr[ni][Mfma<T>::row(lane, r) - Mfma<T>::row(lane, 0)] = old[ni][r] - acc[mi][ni][r];
      }
    }
    return;
  }
  // per row block mi, the old values are loaded together from clamped
  // addresses, then the valid ones are stored
#pragma unroll
  for (int mi = 0; mi < 4; ++mi) {
    T old[NI][4];
    long adr[NI][4];
    bool ok[NI][4];
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) {
      const int j = Tk.tj + wc + ni * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = Tk.ti + wr + mi * 16 + Mfma<T>::row(lane, r);
        ok[ni][r] = j < Tk.chi && i < nr && i >= j && i >= Tk.rlo;
        adr[ni][r] = ok[ni][r] ? base + fpos(nr, i, j) : base;
        old[ni][r] = F[adr[ni][r]];
      }
    }
#pragma unroll
    for (int ni = 0; ni < NI; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (ok[ni][r]) F[adr[ni][r]] = old[ni][r] - acc[mi][ni][r];
  }
}

// ---------------------------------------------------------------------------
// pivot statistics: min |D| (NaN-propagating)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_dstats(const double* __restrict__ D, int N, double* out) {
  __shared__ double red[256];
  double mn = INFINITY;
  bool nan = false;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    double v = fabs(D[i]);
    if (v != v) nan = true;
    mn = fmin(mn, v);
  }
  red[threadIdx.x] = nan ? NAN : mn;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      double a = red[threadIdx.x], b = red[threadIdx.x + w];
      red[threadIdx.x] = (a != a || b != b) ? NAN : fmin(a, b);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = (N > 0) ? red[0] : 1.0;
}

// ---------------------------------------------------------------------------
// solves
// ---------------------------------------------------------------------------
__global__ void k_zero_int(int* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// also empties both hand-off buffers (hand[0, 2N)) for the sweeps
__global__ void k_perm_in(const int* __restrict__ perm, const double* __restrict__ rhs, double* xp,
                          double* hand, int N) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < N) {
    xp[k] = rhs[perm[k]];
    hand[k] = __longlong_as_double(HANDOFF_EMPTY);
    hand[N + k] = __longlong_as_double(HANDOFF_EMPTY);
  }
}
__global__ void k_perm_out(const int* __restrict__ perm, const double* __restrict__ xp, double* x,
                           int N) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < N) x[perm[k]] = xp[k];
}

__device__ inline double bcast(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}

// forward: L y = b on the permuted vector; pivot rows of each front get y,
// non-pivot rows produce the update vector U_s for the parent.  One task per
// 64-row block (lane = row).  Each wave owns the column blocks cb = wv (mod 4):
// it loads the 64x64 block of L into registers *before* polling the block's
// flag, so once y_cb is published only a broadcast-and-FMA pass remains.  The
// diagonal block is staged in LDS up front.
template <typename T>
__global__ __launch_bounds__(256) void k_fwd(
    const SolveTask* __restrict__ tasks, int* ticket, const long* __restrict__ fofs,
    const int* __restrict__ nrv, const int* __restrict__ ncv, const int* __restrict__ firstv,
    const int* __restrict__ rowptr, const int* __restrict__ gptr, const int* __restrict__ gidx,
    const long* __restrict__ uofs, const int* __restrict__ child_ptr, const int* __restrict__ child,
    const int* __restrict__ nblk, const int* __restrict__ flag_ofs, int* flags, int* done, int gen,
    const T* __restrict__ F, double* xp, double* yh, double* U, int* err, SweepCtx C) {
  __shared__ int tk;
  __shared__ double part[4][SB];
  __shared__ double Ld[SB * LDSP];  // Ld[row][col] of the diagonal block
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (tid == 0) tk = atomicAdd(ticket, 1);
  __syncthreads();
  const int ctx = tk % C.nctx;
  const SolveTask Tk = tasks[tk / C.nctx];
  xp += ctx * C.vstride, yh += ctx * C.hstride, U += ctx * C.ustride;
  flags += ctx * C.fstride, done += ctx * C.dstride;
  const int s = Tk.s, rb = Tk.blk, nr = nrv[s], nc = ncv[s], first = firstv[s];
  const long base = fofs[s];
  const int r0 = rb * SB;
  const int ncb = (nc + SB - 1) / SB;
  const int r = r0 + lane;
  const bool rv = r < nr;
  const int rc = rv ? r : nr - 1;  // clamped, always in bounds
  const bool diag = rb < ncb;
  const int kc = diag ? min(SB, nc - r0) : 0;
  // (entries above the diagonal are never used: read the diagonal instead)
  for (int c = wv; c < kc; c += 4) Ld[lane * LDSP + c] = (double)F[base + fpos(nr, max(rc, r0 + c), r0 + c)];
  // right-hand side of this block's pivot rows, read now: only this task
  // writes them, and the read stays off the dependency chain
  const double rhs_r = (wv == 0 && rv && r < nc) ? xp[first + r] : 0.0;
  // Children's update vectors first: they do not depend on this front's own
  // blocks, so gathering them here keeps them off the front's dependency
  // chain.  Every thread waits for a share of the children, every wave sums a
  // share of each row's gather list.
  for (int t = child_ptr[s] + tid; t < child_ptr[s + 1]; t += 256) {
    const int c = child[t];
    wait_count(&done[c * FSTRIDE], nblk[c], err);
  }
  __syncthreads();
  double acc = 0.0;
  if (rv) {
    const int gr = rowptr[s] + r;
    for (int g = gptr[gr] + wv; g < gptr[gr + 1]; g += 4) acc -= load_acq(&U[gidx[g]]);
  }
  const int ncbk = min(rb, ncb);
  // a published block implies every earlier block of the front is published
  // (its task read them all), so a wave only polls past what it already knows
  int ready = -1;
  for (int cb = wv; cb < ncbk; cb += 4) {
    const int cbase = cb * SB;
    const int cend = min(SB, nc - cbase);
    // All 64 columns are read unconditionally (row rc lies below every one of
    // them, so past the pivot columns they are in-bounds entries of the
    // update block) and the unused ones dropped afterwards: every load of the
    // block is in flight together.
    const T* Lr = F + base + fpos(nr, rc, cbase);
    const long step = nr - cbase - 1;  // fpos(rc, j+1) - fpos(rc, j) = nr - j - 1
    T raw[SB];
#pragma unroll
    for (int c = 0; c < SB; ++c) raw[c] = Lr[(long)c * step - (c * (c - 1)) / 2];
    double Lb[SB];
#pragma unroll
    for (int c = 0; c < SB; ++c) Lb[c] = (c < cend) ? (double)raw[c] : 0.0;
    double yv;
    if (cb == ncbk - 1) {
      // the newest block: its values are polled in the hand-off buffer
      yv = wait_value(&yh[first + cbase + lane], lane < cend, err);
    } else {
      if (cb > ready) {
        if (lane == 0) {
          int* fl = &flags[(flag_ofs[s] + ncbk - 1) * FSTRIDE];  // newest block this task needs
          if (__hip_atomic_load(fl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != gen)
            wait_flag(&flags[(flag_ofs[s] + cb) * FSTRIDE], gen, err);
        }
        __builtin_amdgcn_wave_barrier();
        ready = cb;
      }
      yv = (lane < cend) ? load_acq(&xp[first + cbase + lane]) : 0.0;
    }
#pragma unroll
    for (int c = 0; c < SB; ++c) acc += Lb[c] * bcast(yv, c);
  }
  // full block: the lane's row of L11 in registers (zero from the diagonal
  // on), read before the workgroup barrier, off the dependency chain
  double lr[SB];
  if (wv == 0 && kc == SB) {
#pragma unroll
    for (int k = 0; k < SB; ++k) {
      const double v = Ld[lane * LDSP + k];
      lr[k] = (k < lane) ? v : 0.0;
    }
  }
  part[wv][lane] = rv ? acc : 0.0;
  __syncthreads();
  if (wv == 0) {
    double val = 0.0;
    if (rv) val = rhs_r - ((part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]));
    if (kc == SB) {
      // a branch-free chain of broadcast + FMA; a zero entry leaves val
      // unchanged while the broadcast value is finite
#pragma unroll
      for (int k = 0; k < SB; ++k) val -= lr[k] * bcast(val, k);
    } else {
#pragma unroll
      for (int k = 0; k < SB; ++k) {
        if (k < kc) {
          const double yk = bcast(val, k);
          if (lane > k) val -= Ld[lane * LDSP + k] * yk;
        }
      }
    }
    if (rv) {
      if (r < nc) {
        store_shared(&yh[first + r], val);
        store_shared(&xp[first + r], val);
      } else
        store_shared(&U[uofs[s] + (r - nc)], val);
    }
    stores_complete();
    if (lane == 0) {
      if (diag) __hip_atomic_store(&flags[(flag_ofs[s] + rb) * FSTRIDE], gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&done[s * FSTRIDE], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Reduce-scatter of 64 per-lane partials over the wave: afterwards lane c holds
// sum over lanes of a[c].  Six halving steps, 63 shuffles in total.
// Value-level select (a ternary on array elements would be rewritten into a
// per-lane dynamic index, which forces the array into scratch memory).
__device__ inline double bitsel(double x, double y, long long mask) {
  const long long xi = __double_as_longlong(x), yi = __double_as_longlong(y);
  return __longlong_as_double((xi & mask) | (yi & ~mask));
}

__device__ inline double wave_reduce_scatter64(double (&a)[SB], int lane) {
#pragma unroll
  for (int h = 32; h >= 1; h >>= 1) {
    const long long up = -(long long)((lane & h) != 0);
#pragma unroll
    for (int j = 0; j < h; ++j) {
      const double lo = a[j], hi = a[j + h];
      const double keep = bitsel(hi, lo, up);
      const double send = bitsel(lo, hi, up);
      a[j] = keep + __shfl_xor(send, h, 64);
    }
  }
  return a[0];
}

// backward: D^-1 then L^T x = y, one column block of pivots per task.  Rows
// below the block are visited farthest first (non-pivot rows, final once the
// parent front is done, then later pivot blocks in descending order) with
// lane = row: every lane keeps 64 column partials (reduce-scattered once),
// the L block is loaded into registers before the flag is polled.  The two
// 64-row chunks right below the block (blocks cb + 1 and cb + 2 of a dense
// front: the task's last two dependencies) are staged in LDS at the start and
// applied after the far partials are reduced, with lane = column, each from
// the hand-off buffer: the far waves, with their reduce-scatter, then run a
// step ahead of the chain.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd(
    const SolveTask* __restrict__ tasks, int* ticket, const long* __restrict__ fofs,
    const int* __restrict__ nrv, const int* __restrict__ ncv, const int* __restrict__ firstv,
    const int* __restrict__ rowptr, const int* __restrict__ rows, const int* __restrict__ parent,
    const int* __restrict__ nblkb, const int* __restrict__ flag_ofs, int* flags, int* done, int gen,
    const T* __restrict__ F, const double* __restrict__ Dg, double* xp, double* xh, int* err, SweepCtx C) {
  // one block in a fixed order: the broadcast vectors below 64 KB, where
  // their reads take immediate offsets
  struct Lds {
    double xn[2][SB];
    double part[4][SB];
    double Ld[SB * LDSP];     // Ld[row][col] of the diagonal block
    double Ln[2][SB * LDSP];  // Ln[h][row][col] of the two chunks right below it
    int tk;
  };
  __shared__ Lds sm;
  auto& xn = sm.xn;
  auto& part = sm.part;
  auto& Ld = sm.Ld;
  auto& Ln = sm.Ln;
  int& tk = sm.tk;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (tid == 0) tk = atomicAdd(ticket, 1);
  __syncthreads();
  const int ctx = tk % C.nctx;
  const SolveTask Tk = tasks[tk / C.nctx];
  xp += ctx * C.vstride, xh += ctx * C.hstride;
  flags += ctx * C.fstride, done += ctx * C.dstride;
  const int s = Tk.s, cb = Tk.blk, nr = nrv[s], nc = ncv[s], first = firstv[s];
  const long base = fofs[s];
  const int c0 = cb * SB, kc = min(SB, nc - c0);
  const int ncb = (nc + SB - 1) / SB;
  const int rstart = c0 + kc;
  const int nchunks = (nr - rstart + SB - 1) / SB;
  // this block's forward result and pivots, read (and divided) now: only this
  // task writes the block, and both stay off the dependency chain
  double yd = 0.0;
  if (wv == 0 && lane < kc) yd = xp[first + c0 + lane] * (1.0 / Dg[first + c0 + lane]);
  for (int h = 0; h < 2 && h < nchunks; ++h) {
    // lane = row, wave g = columns g, g+4, ...: loads from clamped (valid,
    // lower-triangle) addresses all in flight, then the ones in range stored
    const int r0h = rstart + h * SB;
    const int nnear = min(SB, nr - r0h);
    const int rrow = r0h + min(lane, nnear - 1);
    T v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = F[base + fpos(nr, rrow, c0 + min(wv + 4 * q, kc - 1))];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int c = wv + 4 * q;
      Ln[h][lane * LDSP + c] = (lane < nnear && c < kc) ? (double)v[q] : 0.0;
    }
  }
  for (int c = wv; c < kc; c += 4)
    Ld[lane * LDSP + c] = (lane < kc) ? (double)F[base + fpos(nr, c0 + max(lane, c), c0 + c)] : 0.0;
  if (tid == 0 && parent[s] >= 0) wait_count(&done[parent[s] * FSTRIDE], nblkb[parent[s]], err);
  __syncthreads();
  double acc[SB];
#pragma unroll
  for (int c = 0; c < SB; ++c) acc[c] = 0.0;
  for (int q = nchunks - 1 - wv; q >= 2; q -= 4) {
    const int R = rstart + q * SB;
    const int rr = R + lane;
    const bool rv = rr < nr;
    const int rrc = rv ? rr : nr - 1;
    constexpr int H = SB / 2;
    // row rrc lies below columns c0 .. c0+kc-1; reads past them are clamped to
    // the last one and dropped, so every load of a half is in flight together
    const T* Lr = F + base + fpos(nr, rrc, c0);
    const long step = nr - c0 - 1;  // fpos(rrc, j+1) - fpos(rrc, j) = nr - j - 1
    auto lofs = [&](int c) {
      const int cc = min(c, kc - 1);
      return (long)cc * step - ((long)cc * (cc - 1)) / 2;
    };
    T raw[H];
#pragma unroll
    for (int c = 0; c < H; ++c) raw[c] = Lr[lofs(c)];
    double Lb[H];  // first half of the block, loaded before the flag is polled
#pragma unroll
    for (int c = 0; c < H; ++c) Lb[c] = (c < kc) ? (double)raw[c] : 0.0;
    if (lane == 0 && R < nc) {
      // pivot rows [R, min(R+SB, nc)) belong to blocks b > cb of this front
      const int b0 = R / SB, b1 = (min(R + SB, nc) - 1) / SB;
      for (int b = b1; b >= b0; --b)
        if (b > cb && b < ncb) wait_flag(&flags[(flag_ofs[s] + b) * FSTRIDE], gen, err);
    }
    __builtin_amdgcn_wave_barrier();
    double xv = 0.0;
    if (rv) {
      const int g = (rr < nc) ? first + rr : rows[rowptr[s] + rr];
      xv = load_acq(&xp[g]);
    }
#pragma unroll
    for (int c = 0; c < H; ++c) acc[c] += Lb[c] * xv;
#pragma unroll
    for (int c = 0; c < H; ++c) raw[c] = Lr[lofs(H + c)];
#pragma unroll
    for (int c = 0; c < H; ++c) Lb[c] = (H + c < kc) ? (double)raw[c] : 0.0;
#pragma unroll
    for (int c = 0; c < H; ++c) acc[H + c] += Lb[c] * xv;
  }
  part[wv][lane] = (nchunks - 1 - wv >= 2) ? wave_reduce_scatter64(acc, lane) : 0.0;
  __syncthreads();
  if (wv == 0) {
    // full block: the lane's column of L11 in registers (zero from the
    // diagonal up), read before the waits, off the dependency chain
    double lc[SB];
    if (kc == SB) {
#pragma unroll
      for (int k = 0; k < SB; ++k) {
        const double v = Ld[k * LDSP + lane];
        lc[k] = (lane < k) ? v : 0.0;
      }
    }
    // the second chunk first (its pivot rows were published a step earlier),
    // then the first; pivot rows are polled in the hand-off buffer, rows past
    // the pivots are final since the parent is done
    double pn[2] = {0.0, 0.0};
    for (int h = min(nchunks, 2) - 1; h >= 0; --h) {
      const int rr = rstart + h * SB + lane;
      const bool piv = rr < nc;
      double xv = wait_value(&xh[first + rr], piv, err);
      if (!piv && rr < nr) xv = load_acq(&xp[rows[rowptr[s] + rr]]);
      xn[h][lane] = xv;
      __builtin_amdgcn_wave_barrier();
      double p4[4] = {0.0, 0.0, 0.0, 0.0};  // four partial sums (shorter FMA chain)
#pragma unroll
      for (int r = 0; r < SB; ++r) p4[r & 3] += Ln[h][r * LDSP + lane] * xn[h][r];  // lane = column
      pn[h] = (p4[0] + p4[1]) + (p4[2] + p4[3]);
    }
    double y = 0.0;
    if (lane < kc)
      y = yd - ((((part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane])) + pn[1]) + pn[0]);
    if (kc == SB) {
      // a branch-free chain of broadcast + FMA
#pragma unroll
      for (int k = SB - 1; k >= 0; --k) y -= lc[k] * bcast(y, k);
    } else {
#pragma unroll
      for (int k = SB - 1; k >= 0; --k) {
        if (k < kc) {
          const double xk = bcast(y, k);
          if (lane < k) y -= Ld[k * LDSP + lane] * xk;
        }
      }
    }
    if (lane < kc) {
      store_shared(&xh[first + c0 + lane], y);
      store_shared(&xp[first + c0 + lane], y);
    }
    stores_complete();
    if (lane == 0) {
      __hip_atomic_store(&flags[(flag_ofs[s] + cb) * FSTRIDE], gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&done[s * FSTRIDE], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// y = K x with K symmetric, stored as a full CSR whose values point into the
// lower CSC value array.
__global__ __launch_bounds__(256) void k_spmv_sym(int N, const int* __restrict__ ptr,
                                                  const int* __restrict__ col,
                                                  const int* __restrict__ src,
                                                  const double* __restrict__ Kx,
                                                  const double* __restrict__ x, double* y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double acc = 0.0;
  for (int e = ptr[i]; e < ptr[i + 1]; ++e) acc += Kx[src[e]] * x[col[e]];
  y[i] = acc;
}

}  // namespace

// ===========================================================================
// host side
// ===========================================================================
void DeviceKKT::setup(const LowerPattern& K, int nx, hipStream_t stream) {
  st_ = stream;
  nx_ = nx;
  P_ = plan_supernodes(K);
  const SupernodalPlan& P = P_;
  const int N = P.N, ns = P.ns;
  // front geometry
  std::vector<long> fofs(ns + 1, 0);
  std::vector<int> nr(ns), nc(ns), first(ns);
  for (int s = 0; s < ns; ++s) {
    nr[s] = P.nr(s);
    nc[s] = P.sn_ncol[s];
    first[s] = P.sn_first[s];
    fofs[s + 1] = fofs[s] + colofs(nr[s], nr[s]);
  }
  // K entries grouped by permuted column (= column of the lower permuted entry)
  std::vector<int> kcount(N + 1, 0);
  const int nnzK = K.colptr[N];
  std::vector<int> kpcol(nnzK), kprow(nnzK);
  for (int c = 0; c < N; ++c)
    for (int k = K.colptr[c]; k < K.colptr[c + 1]; ++k) {
      int pr = P.iperm[K.rowind[k]], pc = P.iperm[c];
      int lo = std::min(pr, pc), hi = std::max(pr, pc);
      kpcol[k] = lo;
      kprow[k] = hi;
      kcount[lo + 1]++;
    }
  for (int j = 0; j < N; ++j) kcount[j + 1] += kcount[j];
  std::vector<int> ksrc(nnzK);
  std::vector<long> kdst(nnzK);
  std::vector<signed char> kdiag(nnzK);
  {
    std::vector<int> fill(kcount.begin(), kcount.end() - 1);
    // position of a global row inside a front: binary search in the row list
    for (int c = 0; c < N; ++c)
      for (int k = K.colptr[c]; k < K.colptr[c + 1]; ++k) {
        int lo = kpcol[k], hi = kprow[k];
        int s = P.col_sn[lo];
        const int* rb = P.rows.data() + P.row_ptr[s];
        const int* re = P.rows.data() + P.row_ptr[s + 1];
        const int* it = std::lower_bound(rb, re, hi);
        if (it == re || *it != hi) throw std::runtime_error("kkt plan: row missing from front");
        int rloc = (int)(it - rb), cloc = lo - P.sn_first[s];
        int dst = fill[lo]++;
        ksrc[dst] = k;
        kdst[dst] = fofs[s] + fpos(nr[s], rloc, cloc);
        kdiag[dst] = (K.rowind[k] == c) ? (c < nx ? 1 : -1) : 0;
      }
  }
  // child -> parent relative row maps
  std::vector<int> rmap_ofs(ns + 1, 0), rmap;
  for (int s = 0; s < ns; ++s) rmap_ofs[s + 1] = rmap_ofs[s] + (nr[s] - nc[s]);
  rmap.resize(rmap_ofs[ns]);
  for (int s = 0; s < ns; ++s) {
    int p = P.sn_parent[s];
    if (p < 0) {
      if (nr[s] != nc[s]) throw std::runtime_error("kkt plan: root with update rows");
      continue;
    }
    const int* pb = P.rows.data() + P.row_ptr[p];
    const int* pe = P.rows.data() + P.row_ptr[p + 1];
    for (int i = 0; i < nr[s] - nc[s]; ++i) {
      int g = P.rows[P.row_ptr[s] + nc[s] + i];
      const int* it = std::lower_bound(pb, pe, g);
      if (it == pe || *it != g) throw std::runtime_error("kkt plan: child row not in parent");
      rmap[rmap_ofs[s] + i] = (int)(it - pb);
    }
  }
  std::vector<double> sgn(N);
  for (int k = 0; k < N; ++k) sgn[k] = P.perm[k] < nx ? 1.0 : -1.0;

  fofs_.upload(fofs, st_);
  nr_.upload(nr, st_);
  nc_.upload(nc, st_);
  first_.upload(first, st_);
  rowptr_.upload(P.row_ptr, st_);
  rows_.upload(P.rows, st_);
  child_ptr_.upload(P.child_ptr, st_);
  child_.upload(P.child, st_);
  sgn_.upload(sgn, st_);
  kcolptr_.upload(kcount, st_);
  ksrc_.upload(ksrc, st_);
  kdst_.upload(kdst, st_);
  kdiag_.upload(kdiag, st_);
  rmap_ofs_.upload(rmap_ofs, st_);
  rmap_.upload(rmap, st_);
  F_.alloc(fofs[ns] + TM);  // + TM: the trailing update's staging reads run past the last front
  D_.alloc(N);
  fstat_.alloc(3);
  dstat_.alloc(1);

  build_schedule();

  // --- solve plan ---
  std::vector<int> flag_ofs(ns + 1, 0), nblk(ns), nblkb(ns);
  std::vector<long> uofs(ns + 1, 0);
  for (int s = 0; s < ns; ++s) {
    nblk[s] = ceil_div(nr[s], SB);
    nblkb[s] = ceil_div(nc[s], SB);
    flag_ofs[s + 1] = flag_ofs[s] + nblkb[s];
    uofs[s + 1] = uofs[s] + (nr[s] - nc[s]);
  }
  // gather lists: for each front row, the U entries of children that land on it
  std::vector<int> gcount(P.row_ptr[ns] + 1, 0);
  for (int c = 0; c < ns; ++c) {
    int p = P.sn_parent[c];
    if (p < 0) continue;
    for (int i = 0; i < nr[c] - nc[c]; ++i) gcount[P.row_ptr[p] + rmap[rmap_ofs[c] + i] + 1]++;
  }
  for (size_t i = 0; i + 1 < gcount.size(); ++i) gcount[i + 1] += gcount[i];
  std::vector<int> gidx(gcount.back());
  {
    std::vector<int> fill(gcount.begin(), gcount.end() - 1);
    for (int c = 0; c < ns; ++c) {
      int p = P.sn_parent[c];
      if (p < 0) continue;
      for (int i = 0; i < nr[c] - nc[c]; ++i)
        gidx[fill[P.row_ptr[p] + rmap[rmap_ofs[c] + i]]++] = (int)(uofs[c] + i);
    }
  }
  std::vector<SolveTask> fwd, bwd;
  for (int l = 0; l < P.nlevels; ++l) {
    int maxb = 0;
    for (int t = P.level_ptr[l]; t < P.level_ptr[l + 1]; ++t) maxb = std::max(maxb, nblk[P.level_sn[t]]);
    for (int b = 0; b < maxb; ++b)
      for (int t = P.level_ptr[l]; t < P.level_ptr[l + 1]; ++t) {
        int s = P.level_sn[t];
        if (b < nblk[s]) fwd.push_back({s, b});
      }
  }
  for (int l = P.nlevels - 1; l >= 0; --l) {
    int maxb = 0;
    for (int t = P.level_ptr[l]; t < P.level_ptr[l + 1]; ++t) maxb = std::max(maxb, nblkb[P.level_sn[t]]);
    for (int b = maxb - 1; b >= 0; --b)
      for (int t = P.level_ptr[l]; t < P.level_ptr[l + 1]; ++t) {
        int s = P.level_sn[t];
        if (b < nblkb[s]) bwd.push_back({s, b});
      }
  }
  nfwd_ = (int)fwd.size();
  nbwd_ = (int)bwd.size();
  perm_.upload(P.perm, st_);
  fwd_.upload(fwd, st_);
  bwd_.upload(bwd, st_);
  gptr_.upload(gcount, st_);
  gidx_.upload(gidx, st_);
  uofs_.upload(uofs, st_);
  flag_ofs_.upload(flag_ofs, st_);
  nblk_.upload(nblk, st_);
  nblkb_.upload(nblkb, st_);
  parent_.upload(P.sn_parent, st_);
  // two sweep contexts (solve2_async): flags, done counters, work vectors,
  // hand-off buffers and update vectors twice over
  flags_.alloc(2 * (size_t)std::max(1, flag_ofs[ns]) * FSTRIDE);
  flags_.zero(st_);
  done_.alloc((size_t)(4 * ns + 2) * FSTRIDE);
  serr_.alloc(1);
  serr_.zero(st_);
  xp_.alloc(2 * (size_t)std::max(1, N));
  hand_.alloc(4 * (size_t)std::max(1, N));
  U_.alloc(2 * std::max<long>(1, uofs[ns]));
  gen_ = 0;

  // --- full symmetric CSR of K for the refinement residual ---
  std::vector<int> fptr(N + 1, 0);
  for (int c = 0; c < N; ++c)
    for (int k = K.colptr[c]; k < K.colptr[c + 1]; ++k) {
      int r = K.rowind[k];
      fptr[r + 1]++;
      if (r != c) fptr[c + 1]++;
    }
  for (int i = 0; i < N; ++i) fptr[i + 1] += fptr[i];
  std::vector<int> fcol(fptr[N]), fsrc(fptr[N]);
  {
    std::vector<int> fill(fptr.begin(), fptr.end() - 1);
    for (int c = 0; c < N; ++c)
      for (int k = K.colptr[c]; k < K.colptr[c + 1]; ++k) {
        int r = K.rowind[k];
        fcol[fill[r]] = c, fsrc[fill[r]++] = k;
        if (r != c) fcol[fill[c]] = r, fsrc[fill[c]++] = k;
      }
  }
  kfptr_.upload(fptr, st_);
  kfcol_.upload(fcol, st_);
  kfsrc_.upload(fsrc, st_);
  {
    const char* e = std::getenv("CONICXX_KERNEL_TIMING");
    const int every = e ? std::atoi(e) : 0;
    timing_ = every > 0;
    timing_every_ = std::max(1, every);
  }
  // algorithmic bytes of one triangular sweep: every stored entry of L once
  sweep_bytes_ = 8.0 * P_.nnzL + 3.0 * 8.0 * N;
  CX_HIP(hipStreamSynchronize(st_));
}

void DeviceKKT::build_schedule() {
  const SupernodalPlan& P = P_;
  const int ns = P.ns;
  launches_.clear();
  asm_h_.clear();
  seg_h_.clear();
  small_h_.clear();
  panel_h_.clear();
  dslot_h_.clear();
  trail_h_.clear();
  trail_launch_flops_.clear();
  std::vector<int> rmap_ofs(ns + 1, 0);
  for (int s = 0; s < ns; ++s) rmap_ofs[s + 1] = rmap_ofs[s] + (P.nr(s) - P.sn_ncol[s]);
  // parent-relative maps again (host copy for segment search)
  std::vector<int> rmap(rmap_ofs[ns]);
  for (int s = 0; s < ns; ++s) {
    int p = P.sn_parent[s];
    if (p < 0) continue;
    const int* pb = P.rows.data() + P.row_ptr[p];
    const int* pe = P.rows.data() + P.row_ptr[p + 1];
    for (int i = 0; i < P.nr(s) - P.sn_ncol[s]; ++i)
      rmap[rmap_ofs[s] + i] = (int)(std::lower_bound(pb, pe, P.rows[P.row_ptr[s] + P.sn_ncol[s] + i]) - pb);
  }
  auto push_trail = [&](int s, int ti, int tj, int k0, int k1, int rlo, int chi, int nr, int tt, double& fl) {
    trail_h_.push_back({s, ti, tj, k0, k1, rlo, chi});
    // algorithmic flops: 2 * K * (#updated lower elements in the tile)
    double cnt = 0;
    int jmax = std::min(tj + tt, chi);
    if (ti >= jmax - 1 && ti >= rlo) {  // tile entirely on or below the diagonal
      fl += 2.0 * (k1 - k0) * (double)(jmax - tj) * (double)std::max(0, std::min(ti + tt, nr) - ti);
      return;
    }
    for (int j = tj; j < jmax; ++j) {
      int i0 = std::max({ti, j, rlo});
      int i1 = std::min(ti + tt, nr);
      if (i1 > i0) cnt += i1 - i0;
    }
    fl += 2.0 * (k1 - k0) * cnt;
  };
  // tiles of a rectangle [r0, nr) x [c0, c1) of the lower triangle
  auto count_tiles = [](int r0, int nr, int c0, int c1, int tt) {
    long n = 0;
    for (int ti = r0; ti < nr; ti += tt)
      for (int tj = c0; tj < c1; tj += tt)
        if (ti + tt - 1 >= tj) ++n;
    return n;
  };
  for (int l = 0; l < P.nlevels; ++l) {
    // assembly
    int beg = (int)asm_h_.size();
    std::vector<int> bigs;
    for (int t = P.level_ptr[l]; t < P.level_ptr[l + 1]; ++t) {
      int s = P.level_sn[t];
      int nr = P.nr(s);
      const int cb = std::max(1, std::min(nr, 16384 / std::max(1, nr)));
      const int ntask = (nr + cb - 1) / cb;
      // split every child's row map (sorted parent positions) into the
      // column chunks of the tasks: O(sum of child update sizes + tasks)
      std::vector<std::vector<AsmSeg>> per(ntask);
      for (int q = P.child_ptr[s]; q < P.child_ptr[s + 1]; ++q) {
        const int c = P.child[q];
        const int* rb = rmap.data() + rmap_ofs[c];
        const int* re = rmap.data() + rmap_ofs[c + 1];
        const int nu = (int)(re - rb);
        for (int i = 0; i < nu;) {
          const int t = rb[i] / cb;
          const int iend = (int)(std::lower_bound(rb + i, re, (t + 1) * cb) - rb);
          per[t].push_back({c, i, iend});
          i = iend;
        }
      }
      for (int t = 0; t < ntask; ++t) {
        AsmTask T{s, t * cb, std::min(nr, (t + 1) * cb), (int)seg_h_.size(), 0};
        seg_h_.insert(seg_h_.end(), per[t].begin(), per[t].end());
        T.seg_end = (int)seg_h_.size();
        asm_h_.push_back(T);
      }
      if (nr <= SMALL_NR)
        small_h_.push_back(s);
      else
        bigs.push_back(s);
    }
    launches_.push_back({L_ASSEMBLE, beg, (int)asm_h_.size()});
    // small fronts of this level
    int sb = 0;
    for (int t = P.level_ptr[l]; t < P.level_ptr[l + 1]; ++t)
      if (P.nr(P.level_sn[t]) <= SMALL_NR) ++sb;
    if (sb) launches_.push_back({L_SMALL, (int)small_h_.size() - sb, (int)small_h_.size()});
    // big fronts: blocked right-looking in sub-panels of NB columns; each front
    // takes the trailing update of its outer panel (panel_width columns) once
    // the panel's last sub-panel is done.  One panel launch and one trailing
    // launch per sub-panel step: the in-panel updates of fronts inside a panel
    // and the outer updates of fronts whose panel just ended share it.
    int maxnc = 0;
    for (int s : bigs) maxnc = std::max(maxnc, P.sn_ncol[s]);
    for (int a0 = 0; a0 < maxnc; a0 += NB) {
      int tbeg = (int)trail_h_.size();
      double fl = 0;
      const int rbeg = (int)panel_h_.size();
      // the launch's tile: 128 unless its 128-tiles would leave most of the
      // GPU idle (fewer than one per CU)
      long n128 = 0;
      for (int s : bigs) {
        int nc = P.sn_ncol[s], nr = P.nr(s);
        if (a0 >= nc) continue;
        const int pw = panel_width(nr);
        const int P0 = (a0 / pw) * pw;
        const int pend = std::min(P0 + pw, nc), b = a0 + std::min(NB, pend - a0);
        if (b < pend)
          n128 += count_tiles(b, nr, b, pend, TM);
        else if (pend < nr)
          n128 += count_tiles(pend, nr, pend, nr, TM);
      }
      const int tt = n128 < SMALL_TRAIL_TILES ? TM / 2 : TM;
      for (int s : bigs) {
        int nc = P.sn_ncol[s], nr = P.nr(s);
        if (a0 >= nc) continue;
        const int pw = panel_width(nr);
        const int P0 = (a0 / pw) * pw;
        int pend = std::min(P0 + pw, nc);
        int w = std::min(NB, pend - a0);
        // the first task of a sub-panel (r0 == a0 + w) stages the diagonal block
        const int blk = (int)dslot_h_.size();
        dslot_h_.push_back({s, a0, w, 0, 0, blk});
        if (a0 + w >= nr) panel_h_.push_back({s, a0, w, a0 + w, a0 + w, blk});
        for (int r0 = a0 + w; r0 < nr; r0 += PANEL_ROWS)
          panel_h_.push_back({s, a0, w, r0, std::min(nr, r0 + PANEL_ROWS), blk});
        int b = a0 + w;
        if (b < pend) {
          for (int ti = b; ti < nr; ti += tt)
            for (int tj = b; tj < pend; tj += tt)
              if (ti + tt - 1 >= tj) push_trail(s, ti, tj, a0, b, b, pend, nr, tt, fl);
        } else if (pend < nr) {  // panel [P0, pend) complete: update everything right of it
          for (int ti = pend; ti < nr; ti += tt)
            for (int tj = pend; tj < nr; tj += tt)
              if (ti + tt - 1 >= tj) push_trail(s, ti, tj, P0, pend, pend, nr, nr, tt, fl);
        }
      }
      if ((int)panel_h_.size() > rbeg) launches_.push_back({L_PANEL, rbeg, (int)panel_h_.size()});
      if ((int)trail_h_.size() > tbeg) {
        launches_.push_back({tt == TM ? L_TRAIL : L_TRAIL64, tbeg, (int)trail_h_.size()});
        trail_launch_flops_.push_back(fl);
      }
    }
  }
  asm_.upload(asm_h_, st_);
  seg_.upload(seg_h_.empty() ? std::vector<AsmSeg>{{0, 0, 0}} : seg_h_, st_);
  small_.upload(small_h_.empty() ? std::vector<int>{0} : small_h_, st_);
  panel_.upload(panel_h_.empty() ? std::vector<PanelTask>{{0, 0, 0, 0, 0, 0}} : panel_h_, st_);
  dslot_.upload(dslot_h_.empty() ? std::vector<PanelTask>{{0, 0, 0, 0, 0, 0}} : dslot_h_, st_);
  Fd_.alloc(std::max<size_t>(1, dslot_h_.size()) * NB * NB);
  trail_.upload(trail_h_.empty() ? std::vector<TrailTask>{{0, 0, 0, 0, 0, 0, 0}} : trail_h_, st_);
}

void DeviceKKT::tbegin() {
  if (!sample_) return;
  if (evused_ + 2 > evpool_.size()) {
    size_t add = std::max<size_t>(256, evpool_.size());
    for (size_t i = 0; i < add; ++i) {
      hipEvent_t e;
      CX_HIP(hipEventCreate(&e));
      evpool_.push_back(e);
    }
  }
  CX_HIP(hipEventRecord(evpool_[evused_], st_));
}

void DeviceKKT::tend(int kind, double work) {
  if (!timing_) return;
  KindRaw& r = in_factor_ ? fpend_[kind] : kraw_[kind];
  r.all += 1;
  r.work += work;
  if (!sample_) return;
  CX_HIP(hipEventRecord(evpool_[evused_ + 1], st_));
  evpending_.push_back(kind);
  evused_ += 2;
}

DeviceKKT::KindStats DeviceKKT::kind_stats(int k) const {
  const KindRaw& r = kraw_[k];
  KindStats t;
  t.ms = r.timed ? r.ms_timed * (double)r.all / (double)r.timed : 0.0;
  t.work = r.work;
  t.launches = r.all;
  return t;
}

void DeviceKKT::collect_timing() {
  if (!timing_ || evpending_.empty()) return;
  CX_HIP(hipStreamSynchronize(st_));
  for (size_t i = 0; i < evpending_.size(); ++i) {
    float ms = 0;
    CX_HIP(hipEventElapsedTime(&ms, evpool_[2 * i], evpool_[2 * i + 1]));
    KindRaw& k = kraw_[evpending_[i]];
    k.ms_timed += ms;
    k.timed += 1;
  }
  evpending_.clear();
  evused_ = 0;
}

void DeviceKKT::reset_timing() {
  collect_timing();
  for (auto& k : kraw_) k = KindRaw{};
  fact_seq_ = sweep_seq_ = 0;
}

template <typename T>
void DeviceKKT::launch_factor(T* F, const double* Kx, double extra_p, double extra_a, int mode, double eps,
                              double delta) {
  size_t tl = 0;
  sample_ = timing_ && fact_seq_++ % timing_every_ == 0;
  in_factor_ = true;
  fact_ev0_ = evpending_.size();
  for (auto& k : fpend_) k = KindRaw{};
  for (const Launch& L : launches_) {
    const int n = L.end - L.beg;
    if (n <= 0) continue;
    switch (L.kind) {
      case L_ASSEMBLE:
        tbegin();
        hipLaunchKernelGGL(k_assemble<T>, dim3(n), dim3(256), 0, st_, asm_.p + L.beg, seg_.p, fofs_.p,
                           nr_.p, nc_.p, first_.p, kcolptr_.p, ksrc_.p, kdst_.p, kdiag_.p, Kx,
                           extra_p, extra_a, rmap_ofs_.p, rmap_.p, F, fstat_.p);
        tend(T_ASM, 0.0);
        break;
      case L_SMALL:
        tbegin();
        hipLaunchKernelGGL(k_factor_small<T>, dim3(n), dim3(256), 0, st_, small_.p + L.beg, fofs_.p,
                           nr_.p, nc_.p, first_.p, sgn_.p, F, D_.p, mode, eps, delta, fstat_.p);
        tend(T_SMALL, 0.0);
        break;
      case L_PANEL:
        tbegin();
        hipLaunchKernelGGL(k_panel<T>, dim3(n), dim3(256), 0, st_, panel_.p + L.beg, fofs_.p, nr_.p, first_.p,
                           sgn_.p, F, reinterpret_cast<T*>(Fd_.p), D_.p, mode, eps, delta, fstat_.p);
        tend(T_PANEL, 0.0);
        break;
      case L_TRAIL:
        tbegin();
        if constexpr (sizeof(T) == 4)
          hipLaunchKernelGGL((k_trail<T, 2>), dim3(n), dim3(256), 0, st_, trail_.p + L.beg, fofs_.p, nr_.p,
                             first_.p, D_.p, F, fstat_.p);
        else
          hipLaunchKernelGGL((k_trail<T, 4>), dim3(n), dim3(512), 0, st_, trail_.p + L.beg, fofs_.p, nr_.p,
                             first_.p, D_.p, F, fstat_.p);
        tend(T_TRAIL, trail_launch_flops_[tl]);
        ++tl;
        break;
      case L_TRAIL64:
        tbegin();
        hipLaunchKernelGGL((k_trail<T, 4, TM / 2>), dim3(n), dim3(TM / 2 * 4), 0, st_, trail_.p + L.beg, fofs_.p,
                           nr_.p, first_.p, D_.p, F, fstat_.p);
        tend(T_TRAIL, trail_launch_flops_[tl]);
        ++tl;
        break;
    }
    CX_LAUNCH_CHECK();
  }
  if (!dslot_h_.empty()) {
    tbegin();
    hipLaunchKernelGGL(k_diag_store<T>, dim3((unsigned)dslot_h_.size()), dim3(256), 0, st_, dslot_.p, fofs_.p, nr_.p,
                       reinterpret_cast<const T*>(Fd_.p), F, fstat_.p);
    tend(T_PANEL, 0.0);
    CX_LAUNCH_CHECK();
  }
  sample_ = false;
  in_factor_ = false;
}

void DeviceKKT::factorize_async(const double* Kx, double extra_p, double extra_a, int mode,
                                double eps, double delta, int prec) {
  CX_HIP(hipMemsetAsync(fstat_.p, 0, 3 * sizeof(int), st_));
  if (prec == 1) {
    launch_factor<float>(reinterpret_cast<float*>(F_.p), Kx, extra_p, extra_a, mode, eps, delta);
  } else {
    launch_factor<double>(F_.p, Kx, extra_p, extra_a, mode, eps, delta);
  }
  prec_ = prec;
  hipLaunchKernelGGL(k_dstats, dim3(1), dim3(256), 0, st_, D_.p, P_.N, dstat_.p);
  CX_LAUNCH_CHECK();
}

FactorStats DeviceKKT::factor_stats() {
  FactorStats fs;
  int st[3] = {0, 0, 0};
  double mn = 0;
  fstat_.download(st, 3, st_);
  dstat_.download(&mn, 1, st_);
  CX_HIP(hipStreamSynchronize(st_));
  if (timing_) {
    if (st[2]) {
      // stopped at a failing pivot: drop its timed launches
      evpending_.resize(std::min(evpending_.size(), fact_ev0_));
      evused_ = 2 * evpending_.size();
    } else {
      for (int k = 0; k < T_NKINDS; ++k) kraw_[k].all += fpend_[k].all, kraw_[k].work += fpend_[k].work;
    }
    for (auto& k : fpend_) k = KindRaw{};
  }
  collect_timing();
  fs.zero_pivot = st[0];
  fs.nreg = st[1];
  fs.stopped = st[2];
  fs.min_abs_d = mn;
  return fs;
}

void DeviceKKT::solve_async(const double* rhs, double* x) {
  const double* r[1] = {rhs};
  double* xo[1] = {x};
  sweeps(1, r, xo);
}

void DeviceKKT::solve2_async(const double* rhs0, double* x0, const double* rhs1, double* x1) {
  const double* r[2] = {rhs0, rhs1};
  double* xo[2] = {x0, x1};
  sweeps(2, r, xo);
}

// Forward and backward sweeps over nctx right-hand sides.  Counter layout:
// forward done counters of context c at [c ns, (c+1) ns), backward ones at
// [(2 + c) ns, (3 + c) ns), then the two tickets (FSTRIDE ints each).
void DeviceKKT::sweeps(int nctx, const double* const* rhs, double* const* x) {
  const int N = P_.N, ns = P_.ns;
  if (N == 0) return;
  const SweepCtx C{nctx, (long)std::max(1, N), 2L * std::max(1, N), (long)(U_.n / 2), (long)(flags_.n / 2),
                   (long)ns * FSTRIDE};
  const int nctr = (4 * ns + 2) * FSTRIDE;
  hipLaunchKernelGGL(k_zero_int, dim3(std::min(1024, ceil_div(nctr, 256))), dim3(256), 0, st_, done_.p, nctr);
  for (int c = 0; c < nctx; ++c)
    hipLaunchKernelGGL(k_perm_in, dim3(ceil_div(N, 256)), dim3(256), 0, st_, perm_.p, rhs[c], xp_.p + c * C.vstride,
                       hand_.p + c * C.hstride, N);
  int* tick = done_.p + (size_t)4 * ns * FSTRIDE;
  int* done_b = done_.p + (size_t)2 * ns * FSTRIDE;
  ++gen_;
  sample_ = timing_ && sweep_seq_++ % timing_every_ == 0;
  const double bytes = sweep_bytes_ + (nctx - 1) * 24.0 * N;  // L once, the vectors per context
  tbegin();
  if (prec_ == 1)
    hipLaunchKernelGGL(k_fwd<float>, dim3(nfwd_ * nctx), dim3(256), 0, st_, fwd_.p, tick, fofs_.p, nr_.p, nc_.p,
                       first_.p, rowptr_.p, gptr_.p, gidx_.p, uofs_.p, child_ptr_.p, child_.p, nblk_.p,
                       flag_ofs_.p, flags_.p, done_.p, gen_, reinterpret_cast<const float*>(F_.p), xp_.p, hand_.p,
                       U_.p, serr_.p, C);
  else
    hipLaunchKernelGGL(k_fwd<double>, dim3(nfwd_ * nctx), dim3(256), 0, st_, fwd_.p, tick, fofs_.p, nr_.p, nc_.p,
                       first_.p, rowptr_.p, gptr_.p, gidx_.p, uofs_.p, child_ptr_.p, child_.p, nblk_.p,
                       flag_ofs_.p, flags_.p, done_.p, gen_, (const double*)F_.p, xp_.p, hand_.p, U_.p, serr_.p, C);
  tend(T_FWD, bytes);
  ++gen_;
  tbegin();
  if (prec_ == 1)
    hipLaunchKernelGGL(k_bwd<float>, dim3(nbwd_ * nctx), dim3(256), 0, st_, bwd_.p, tick + FSTRIDE, fofs_.p, nr_.p,
                       nc_.p, first_.p, rowptr_.p, rows_.p, parent_.p, nblkb_.p, flag_ofs_.p, flags_.p, done_b,
                       gen_, reinterpret_cast<const float*>(F_.p), D_.p, xp_.p, hand_.p + N, serr_.p, C);
  else
    hipLaunchKernelGGL(k_bwd<double>, dim3(nbwd_ * nctx), dim3(256), 0, st_, bwd_.p, tick + FSTRIDE, fofs_.p, nr_.p,
                       nc_.p, first_.p, rowptr_.p, rows_.p, parent_.p, nblkb_.p, flag_ofs_.p, flags_.p, done_b,
                       gen_, (const double*)F_.p, D_.p, xp_.p, hand_.p + N, serr_.p, C);
  tend(T_BWD, bytes);
  sample_ = false;
  for (int c = 0; c < nctx; ++c)
    hipLaunchKernelGGL(k_perm_out, dim3(ceil_div(N, 256)), dim3(256), 0, st_, perm_.p, xp_.p + c * C.vstride, x[c],
                       N);
  CX_LAUNCH_CHECK();
}

void DeviceKKT::spmv_async(const double* Kx, const double* x, double* y) {
  const int N = P_.N;
  if (N == 0) return;
  hipLaunchKernelGGL(k_spmv_sym, dim3(ceil_div(N, 256)), dim3(256), 0, st_, N, kfptr_.p, kfcol_.p,
                     kfsrc_.p, Kx, x, y);
  CX_LAUNCH_CHECK();
}

int DeviceKKT::take_solve_error() {
  int e = 0;
  serr_.download(&e, 1, st_);
  CX_HIP(hipStreamSynchronize(st_));
  collect_timing();
  if (e) serr_.zero(st_);
  return e;
}

}  // namespace cx
