// shesolve.hip -- MI355X (gfx950) SHE linear solver behind the C ABI in include/shesolve.h.
//
// Replaces the ViennaCL CPU path of the reference solver call
// (src/solvers/viennacl/serial_linear_solver.hpp:99-156: CSR copy, ilu0_precond,
// bicgstab; src/solvers/viennacl/parallel_linear_solver.hpp:50-127: block ILU0).
//
// Kernels (all HBM-bound sparse / vector work; no dense contraction, so no MFMA):
//   - CSR SpMV: LPR lanes per row (a power of two dividing the 64-wide wavefront),
//     coalesced streaming of values and column indices, shuffle row reduction,
//     fixed persistent grid so the fused dot products reduce to a fixed partial count;
//   - ILU(0) factorisation and forward/backward sweeps, level-scheduled: rows of a
//     level are independent, one launch per level; the factorisation keeps the
//     working row in LDS;
//   - fused BiCGStab vector updates with per-block partial dot products reduced by a
//     one-block finalise kernel that also computes the recurrence scalars on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <queue>
#include <thread>
#include <atomic>
#include <mutex>
#include <sys/mman.h>
#include <utility>
#include <vector>

#include "shesolve.h"

namespace {

thread_local std::string g_err;
void set_err(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
}

#define HIPCHK(expr)                                                                 \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) {                                                          \
      set_err("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
      return -10;                                                                    \
    }                                                                                \
  } while (0)

constexpr int kBlock = 256;      // threads per block for streaming kernels
constexpr int kGrid = 2048;      // persistent grid (8 workgroups per CU on 256 CUs)
constexpr int kPartStride = 4 * kGrid;  // per-workgroup partial sums: two dot products, up to 4 kGrid workgroups
constexpr int kMaxRowLDS = 512;  // longest row the LDS factorisation accepts

// ---------------------------------------------------------------- device scalars
struct Scalars {
  double rho, alpha, omega, beta;
  double r0v, ts, tt, rr, r0r, ss;
  double bnorm, res;
  // halt: set by the iteration's last finalize on convergence (res <= tol),
  // breakdown or a non-finite residual; the update kernels and finalizes of a
  // later iteration then leave x, r, p and the scalars alone, so that the host
  // may enqueue an iteration before it has read the previous one's result
  int breakdown, halt;
  double tol;
};

// Sum over aligned groups of LPR lanes (1, 2, 4, 8 or 16), every lane of a
// group receiving it, by DPP moves inside rows of 16 lanes (no LDS round trip,
// as a lane shuffle would take): quad_perm [1,0,3,2] and [2,3,0,1], then
// row_half_mirror and row_mirror.  Every lane of a group must be active.
template <int CTRL>
__device__ inline float dpp_mov(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ inline double dpp_mov(double x) {
  const long long bits = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_mov_dpp((int)bits, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(bits >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int LPR, typename T>
__device__ inline T group_sum(T a) {
  static_assert(LPR == 1 || LPR == 2 || LPR == 4 || LPR == 8 || LPR == 16, "groups of 1 to 16 lanes");
  if constexpr (LPR >= 2) a += dpp_mov<0xB1>(a);
  if constexpr (LPR >= 4) a += dpp_mov<0x4E>(a);
  if constexpr (LPR >= 8) a += dpp_mov<0x141>(a);
  if constexpr (LPR >= 16) a += dpp_mov<0x140>(a);
  return a;
}

// ---------------------------------------------------------------- reductions
__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
// Block-wide sum of NV values; result valid in thread 0.
template <int NV>
__device__ inline void block_sum(double (&v)[NV]) {
  __shared__ double red[NV][kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double s = wave_sum(v[q]);
    if (lane == 0) red[q][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double s = 0;
      for (int w = 0; w < kBlock / 64; ++w) s += red[q][w];
      v[q] = s;
    }
  }
}

// ---------------------------------------------------------------- SpMV
// y = A x            (mode 0)
// y = b - A x        (mode 1)
// plus optional fused dots of y with d0 and d1 (per-block partials).
template <int LPR, typename TV, typename T, typename TA>
__global__ __launch_bounds__(kBlock) void k_spmv(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                                 const TV *__restrict__ v, const T *__restrict__ x,
                                                 T *__restrict__ y, const T *__restrict__ b,
                                                 const T *__restrict__ d0, const T *__restrict__ d1,
                                                 double *__restrict__ part) {
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = gtid & (LPR - 1);
  const int groups = gridDim.x * blockDim.x / LPR;
  double acc0 = 0.0, acc1 = 0.0;
  // all lanes of an LPR group share `row`, so groups leave the loop together
  for (int row = gtid / LPR; row < n; row += groups) {
    TA s = 0;
    const int beg = rp[row], end = rp[row + 1];
    for (int k = beg + lane; k < end; k += LPR) s += (TA)v[k] * (TA)x[ci[k]];
    s = group_sum<LPR>(s);
    if (lane == 0) {
      T out = b ? (T)(b[row] - s) : (T)s;
      y[row] = out;
      if (d0) acc0 += (double)d0[row] * (double)out;
      if (d1) acc1 += (double)d1[row] * (double)out;
    }
  }
  if (part) {
    double vv[2] = {acc0, acc1};
    block_sum<2>(vv);
    if (threadIdx.x == 0) {
      part[blockIdx.x] = vv[0];
      part[kPartStride + blockIdx.x] = vv[1];
    }
  }
}

// Per-system state of a batched BiCGStab (see the batched kernels below).
struct BSys {
  double rho, alpha, omega, beta;
  double bnorm, res, tol, div_limit;
  int active;     // 1 while the system iterates
  int status;     // 0 converged, 1 iteration cap, 2 breakdown / divergence
  int iters;      // iterations of the current solve
  int max_iters;
  int bd, best_it;
  double best;    // best residual so far (stall test of FP32 inner solves)
  int stall;      // iterations without a 1 % gain before giving up (0 = off)
  int pad;
};

// ---------------------------------------------------------------- CSR-stream kernels
// A row block is a run of consecutive rows holding at most kStreamNnz entries
// (a single longer row forms a block of its own).  The block's entries are
// streamed with fully coalesced loads into LDS as products v[k] * x[col[k]],
// then each thread reduces whole rows from LDS.  Used for the SpMV and for the
// sweeps of every contiguous level of the triangular factors.
constexpr int kStreamThreads = 256;
// The pipelined SpMV loads column indices and values four entries at a time
// from the block's first entry rounded down to a multiple of four: the
// matrix arrays carry kVecPad entries past the end, and a block holds at most
// spmv_cap(items) entries so that the rounded range fits items * 256.
constexpr int kVecPad = 4;
constexpr int spmv_cap(int items) { return items * kStreamThreads - 3; }
constexpr int kStreamRows = 256;  // rows per block (rows without entries must not pile up)
// Row blocks of the colour sweeps: entries per thread and rows per block
// (8 and 256 measured fastest on the 2D L=3 device of 16.1 M rows against
// 4-16 entries and 256-1024 rows).
#ifndef SHESOLVE_SWEEP_ITEMS
#define SHESOLVE_SWEEP_ITEMS 8
#endif
#ifndef SHESOLVE_SWEEP_ROWS
#define SHESOLVE_SWEEP_ROWS 256
#endif
// Segment sweeps: entries per lane held in registers (J) and the largest
// lane group per row.
#ifndef SHESOLVE_SEG_J
#define SHESOLVE_SEG_J 2
#endif
#ifndef SHESOLVE_SEG_LPR_MAX
#define SHESOLVE_SEG_LPR_MAX 8
#endif
constexpr int kSegJ = SHESOLVE_SEG_J;
constexpr int kSegLprMax = SHESOLVE_SEG_LPR_MAX;
constexpr double kSegLevelUs = 1.4;      // segment sweep of a 1024-thread workgroup: per level ...
constexpr double kSegLevelUs256 = 1.0;   // ... of a 256-thread workgroup
constexpr double kSegLevelUs64 = 0.8;    // ... of a 64-thread workgroup
constexpr double kSegEntryUs = 6e-4;     // ... plus per entry (and row) of the level
constexpr double kLevelLaunchUs = 5.0;   // one launch per level
constexpr int kSweepItems = SHESOLVE_SWEEP_ITEMS;
constexpr int kSweepBlockRows = SHESOLVE_SWEEP_ROWS;
static_assert(kSweepItems == 4 || kSweepItems == 8 || kSweepItems == 16, "sweep items: 4, 8 or 16");
static_assert(kSweepBlockRows >= 1, "sweep rows per block");
constexpr int kXcds = 8;           // MI355X: 8 XCDs, workgroups dispatched round-robin over them
// Block multicolour ordering: blocks of kBmcRows consecutive natural rows are
// coloured, and a sweep gives each block half a wavefront (one lane per row)
constexpr int kBmcRows = 32;
static_assert(kBmcRows == 32, "one block per half wavefront, two DPP rows");

enum { MODE_SPMV = 0, MODE_LSWEEP = 1, MODE_USWEEP = 2 };

template <typename T>
__device__ inline T block_sum_t(T v) {
  __shared__ T red[kStreamThreads / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  T t = 0;
  for (int w = 0; w < kStreamThreads / 64; ++w) t += red[w];
  __syncthreads();
  return t;
}

// MODE_SPMV:   y[row] = (b ? b[row] - s : s), dots of y with d0, d1 (per persistent block)
// MODE_LSWEEP: y[row] = b[perm ? perm[row] : row] - s           (y == x: the factor-space vector)
// MODE_USWEEP: y[row] = dinv[row] * ((row < c0 ? b[row] : y[row]) - s), out2[perm[row]] = y[row] when perm
// c0: the first level of the forward sweep, stored as the contiguous range
// [0, c0), has no entries, so its output would be b itself; that launch is
// skipped, and the forward sweep's gathers and the backward sweep's own rows
// below c0 read b instead.
// xcd: 0 = row blocks strided over the whole grid; 1 = each XCD (blockIdx % 8)
// works through its own contiguous eighth of the row blocks, so the gathered
// vector entries of neighbouring rows are reused in that XCD's L2.
template <typename TV, typename T, int MODE, int ITEMS, bool SKIP = false>
__global__ __launch_bounds__(kStreamThreads) void k_stream(int nblk, int xcd, const int *__restrict__ blocks,
                                                           const int *__restrict__ rp, const int *__restrict__ ci,
                                                           const TV *__restrict__ v, const T *x, T *y,
                                                           const T *__restrict__ b, const int *__restrict__ perm,
                                                           const T *__restrict__ dinv, T *__restrict__ out2,
                                                           const T *__restrict__ d0, const T *__restrict__ d1,
                                                           double *__restrict__ part, const int *__restrict__ pairs,
                                                           int c0, const int *__restrict__ blk_sys,
                                                           const BSys *__restrict__ skip) {
  constexpr int kStreamNnz = ITEMS * kStreamThreads;
  __shared__ T prod[kStreamNnz];
  double a0 = 0.0, a1 = 0.0;
  int bk0 = blockIdx.x, bk1 = nblk, bstep = gridDim.x;
  if (xcd) {
    const int x = blockIdx.x % kXcds, per = gridDim.x / kXcds;
    bk0 = (int)((int64_t)nblk * x / kXcds) + (int)blockIdx.x / kXcds;
    bk1 = (int)((int64_t)nblk * (x + 1) / kXcds);
    bstep = per;
  }
  for (int bk = bk0; bk < bk1; bk += bstep) {
    // batched iterations: row blocks of converged systems are left alone (a
    // separate instantiation: the check in the loop of every SpMV measured
    // 416 -> 598 us on the fused-dot launch of the 161x81 L=3 device)
    if constexpr (SKIP) {
      if (!skip[blk_sys[bk]].active) continue;
    }
    // consecutive row blocks, or an explicit (first, end) pair per block
    const int rs = pairs ? pairs[2 * bk] : blocks[bk], re = pairs ? pairs[2 * bk + 1] : blocks[bk + 1];
    const int k0 = rp[rs], k1 = rp[re];
    if (k1 - k0 <= kStreamNnz) {
      // all column indices and values first, then the gathers: 8 independent
      // loads in flight per thread
      constexpr int kItems = ITEMS;
      const int cnt = k1 - k0;
      int cc[kItems];
      TV vv[kItems];
#pragma unroll
      for (int u = 0; u < kItems; ++u) {
        const int k = (int)threadIdx.x + u * kStreamThreads;
        if (k < cnt) {
          cc[u] = __builtin_nontemporal_load(ci + k0 + k);
          vv[u] = __builtin_nontemporal_load(v + k0 + k);
        }
      }
#pragma unroll
      for (int u = 0; u < kItems; ++u) {
        const int k = (int)threadIdx.x + u * kStreamThreads;
        if (k < cnt) prod[k] = (T)vv[u] * ((MODE == MODE_LSWEEP && cc[u] < c0) ? b[cc[u]] : x[cc[u]]);
      }
      __syncthreads();
      for (int row = rs + (int)threadIdx.x; row < re; row += kStreamThreads) {
        T acc = 0;
        const int e = rp[row + 1] - k0;
        for (int k = rp[row] - k0; k < e; ++k) acc += prod[k];
        if (MODE == MODE_SPMV) {
          const T out = b ? (T)(b[row] - acc) : acc;
          y[row] = out;
          if (d0) a0 += (double)d0[row] * (double)out;
          if (d1) a1 += (double)d1[row] * (double)out;
        } else if (MODE == MODE_LSWEEP) {
          y[row] = (perm ? b[perm[row]] : b[row]) - acc;
        } else {
          const T val = dinv[row] * ((row < c0 ? b[row] : y[row]) - acc);
          y[row] = val;
          if (perm) out2[perm[row]] = val;
        }
      }
      __syncthreads();
    } else {
      // one long row
      T acc = 0;
      for (int k = k0 + (int)threadIdx.x; k < k1; k += kStreamThreads) {
        const int cc = ci[k];
        acc += (T)v[k] * ((MODE == MODE_LSWEEP && cc < c0) ? b[cc] : x[cc]);
      }
      acc = block_sum_t<T>(acc);
      if (threadIdx.x == 0) {
        const int row = rs;
        if (MODE == MODE_SPMV) {
          const T out = b ? (T)(b[row] - acc) : acc;
          y[row] = out;
          if (d0) a0 += (double)d0[row] * (double)out;
          if (d1) a1 += (double)d1[row] * (double)out;
        } else if (MODE == MODE_LSWEEP) {
          y[row] = (perm ? b[perm[row]] : b[row]) - acc;
        } else {
          const T val = dinv[row] * ((row < c0 ? b[row] : y[row]) - acc);
          y[row] = val;
          if (perm) out2[perm[row]] = val;
        }
      }
      __syncthreads();
    }
  }
  if (part) {
    double vv[2] = {a0, a1};
    block_sum<2>(vv);
    if (threadIdx.x == 0) {
      part[blockIdx.x] = vv[0];
      part[kPartStride + blockIdx.x] = vv[1];
    }
  }
}

// Row block descriptors of the pipelined SpMV: {first row, end row, first
// entry, end entry} per block, from consecutive boundaries or (first, end)
// pairs.
__global__ __launch_bounds__(kBlock) void k_block_desc(int nblk, const int *__restrict__ blocks,
                                                      const int *__restrict__ pairs, const int *__restrict__ rp,
                                                      int4 *__restrict__ desc) {
  const int bk = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (bk >= nblk) return;
  const int rs = pairs ? pairs[2 * bk] : blocks[bk], re = pairs ? pairs[2 * bk + 1] : blocks[bk + 1];
  desc[bk] = make_int4(rs, re, rp[rs], rp[re]);
}

// CSR-stream SpMV, y = (b ? b - A x : A x), with the dot products of y with d0
// and d1 per workgroup (as k_stream MODE_SPMV), software-pipelined over the
// workgroup's row blocks.  Each iteration issues, in this order, the
// descriptor of block t+3, the row and entry loads of block t+2 and the x
// gathers of block t+1 (each from data that arrived during the previous
// iteration), then reduces block t; the in-order memory counter then lets
// every wait leave the younger stages in flight.  With FP32 values, column
// indices and values are loaded 16 bytes per lane (four entries, from the
// block's first entry rounded down to a multiple of four); loads use clamped
// indices and fallback
// pointers instead of branches.  Every block spans at most ITEMS * 256 entries
// from that rounded start (the host launches this kernel only then).
typedef float vfloat4 __attribute__((ext_vector_type(4)));
typedef int vint4 __attribute__((ext_vector_type(4)));
__device__ inline void store4(float *p, const float (&v)[4]) {
  *reinterpret_cast<vfloat4 *>(p) = vfloat4{v[0], v[1], v[2], v[3]};
}

// Raw buffer access (MUBUF): a resource of `bytes` from `p` (a null pointer
// gives an empty resource); loads past the end return zero and stores past it
// are dropped, so the pipelined kernels need neither index clamps nor branches.
constexpr int kBufferDword3 = 0x00020000;  // gfx9-family buffer resource word 3
typedef unsigned vuint2 __attribute__((ext_vector_type(2)));
__device__ inline __amdgpu_buffer_rsrc_t buf_rsrc(const void *p, int64_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, p ? (int)bytes : 0, kBufferDword3);
}
template <typename T>
__device__ inline T buf_load(__amdgpu_buffer_rsrc_t r, int off) {
  if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <typename T>
__device__ inline void buf_store(T v, __amdgpu_buffer_rsrc_t r, int off) {
  if constexpr (sizeof(T) == 8)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(vuint2, v), r, off, 0, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
}
constexpr int kBufferDrop = 0x7ffffff0;  // an offset past every resource: the store is dropped

// CODED: the column of an entry is not read from ci but decoded from a 16-bit
// code, table[block's table start + (code >> 5)] + (code & 31): the table
// lists the first solve-space row of each 32-row block of the block
// multicolour ordering that the row block's columns fall in (at most
// kCodeSlots), and the low five bits are the column's row inside that block.
// The descriptor then has eight ints per block, {first row, end row, first
// entry, end entry, table start, table end, 0, 0}, and every wavefront holds
// the row block's table in two registers per lane, read with ds_bpermute.
constexpr int kCodeSlots = 128;
template <typename TV, typename T, int ITEMS, bool CODED = false>
__global__ __launch_bounds__(kStreamThreads) void k_spmv_pipe(int nblk, int xcd, int n, int64_t nnz_alloc,
                                                              const int *__restrict__ desc,
                                                              const int *__restrict__ rp, const int *__restrict__ ci,
                                                              const TV *__restrict__ v, const T *x, T *y,
                                                              const T *__restrict__ b, const T *d0, const T *d1,
                                                              double *__restrict__ part,
                                                              const unsigned short *__restrict__ code = nullptr,
                                                              const int *__restrict__ tab = nullptr, int64_t ntab = 0) {
  // W consecutive entries per lane and load: four with FP32 values (16-byte
  // loads), one with FP64 values (measured faster there: 563 against 594 us
  // on the 161x81 L=3 device)
  constexpr int NT = kStreamThreads, W = sizeof(TV) == 4 ? 4 : 1, NV = ITEMS / W;
  static_assert(ITEMS % 4 == 0, "four entries per load");
  __shared__ __align__(16) T prod[ITEMS * NT];
  double a0 = 0.0, a1 = 0.0;
  int bk0 = (int)blockIdx.x, bk1 = nblk, bstep = (int)gridDim.x;
  if (xcd) {
    const int xx = (int)blockIdx.x % kXcds;
    bk0 = (int)((int64_t)nblk * xx / kXcds) + (int)blockIdx.x / kXcds;
    bk1 = (int)((int64_t)nblk * (xx + 1) / kXcds);
    bstep = (int)gridDim.x / kXcds;
  }
  const int count = bk0 < bk1 ? (bk1 - bk0 + bstep - 1) / bstep : 0;
  const int tid = (int)threadIdx.x;
  // a dot product with y itself takes the new value, not a prefetched one
  const bool d0y = (const void *)d0 == (const void *)y, d1y = (const void *)d1 == (const void *)y;
  constexpr int DW = CODED ? 8 : 4;  // descriptor ints per block
  const auto rdesc = buf_rsrc(desc, (int64_t)nblk * DW * 4), rrp = buf_rsrc(rp, (int64_t)(n + 1) * 4);
  const auto rci = buf_rsrc(CODED ? nullptr : ci, nnz_alloc * 4), rv = buf_rsrc(v, nnz_alloc * (int64_t)sizeof(TV));
  const auto rcode = buf_rsrc(CODED ? code : nullptr, nnz_alloc * 2), rtab = buf_rsrc(CODED ? tab : nullptr, ntab * 4);
  const int lane = tid & 63;
  const auto rx = buf_rsrc(x, (int64_t)n * sizeof(T)), ry = buf_rsrc(y, (int64_t)n * sizeof(T));
  const auto rb = buf_rsrc(b, (int64_t)n * sizeof(T));
  const auto r0 = buf_rsrc(d0 && !d0y ? d0 : nullptr, (int64_t)n * sizeof(T));
  const auto r1 = buf_rsrc(d1 && !d1y ? d1 : nullptr, (int64_t)n * sizeof(T));
  if (count > 0) {
    const int last = bk0 + (count - 1) * bstep;
    struct StB {
      int rs, re, k0, k1, r0, r1;
      T bb, e0, e1;
      int cc[ITEMS];  // columns, or (CODED) codes
      TV vv[ITEMS];
      int tb0, tb1, wide;  // CODED: the block's table (slots lane and 64 + lane), more than 64 slots
    };
    // lane l holds component l % DW of the block's descriptor
    auto stage_a = [&](int t) {
      return __builtin_amdgcn_raw_buffer_load_b32(rdesc, (DW * min(bk0 + t * bstep, last) + (tid & (DW - 1))) * 4,
                                                 0, 0);
    };
    // entry j of vector q of this lane: base + W (tid + q NT) + j, base = k0
    // rounded down to a multiple of four; a load that starts past the block's
    // last entry is not made (its offset lies past the buffer, which reads
    // zero), and entries outside the block are never summed
    auto stage_b = [&](int dv, StB &s) {
      s.rs = __builtin_amdgcn_readlane(dv, 0);
      s.re = __builtin_amdgcn_readlane(dv, 1);
      s.k0 = __builtin_amdgcn_readlane(dv, 2);
      s.k1 = __builtin_amdgcn_readlane(dv, 3);
      if constexpr (CODED) {
        const int t0 = __builtin_amdgcn_readlane(dv, 4), t1 = __builtin_amdgcn_readlane(dv, 5);
        s.tb0 = (int)__builtin_amdgcn_raw_buffer_load_b32(rtab, t0 + lane < t1 ? (t0 + lane) * 4 : kBufferDrop, 0, 0);
        s.tb1 = (int)__builtin_amdgcn_raw_buffer_load_b32(rtab, t0 + 64 + lane < t1 ? (t0 + 64 + lane) * 4 : kBufferDrop,
                                                         0, 0);
        s.wide = t1 - t0 > 64;
      }
      const int rc = s.rs + tid;
      s.r0 = (int)__builtin_amdgcn_raw_buffer_load_b32(rrp, rc * 4, 0, 0);
      s.r1 = (int)__builtin_amdgcn_raw_buffer_load_b32(rrp, rc * 4 + 4, 0, 0);
      s.bb = buf_load<T>(rb, rc * (int)sizeof(T));
      s.e0 = buf_load<T>(r0, rc * (int)sizeof(T));
      s.e1 = buf_load<T>(r1, rc * (int)sizeof(T));
      const int base = s.k0 & ~3;
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        const int e = base + W * (tid + q * NT);
        const int eo = e < s.k1 ? e * 4 : kBufferDrop;
        if constexpr (W == 4 && CODED) {
          // four 16-bit codes in one 8-byte load, split when the gathers are issued
          const vuint2 c2 = __builtin_amdgcn_raw_buffer_load_b64(rcode, e < s.k1 ? e * 2 : kBufferDrop, 0, 0);
          const vfloat4 v4 = __builtin_bit_cast(vfloat4, __builtin_amdgcn_raw_buffer_load_b128(rv, eo, 0, 0));
#pragma unroll
          for (int j = 0; j < 4; ++j) s.vv[4 * q + j] = v4[j];
          s.cc[4 * q] = (int)c2[0];
          s.cc[4 * q + 1] = (int)c2[1];
        } else if constexpr (W == 4) {
          // the whole 16 bytes cast at once (a cast of one element of the
          // builtin's result compiled to a one-dword load reused for all four)
          const vint4 c4 = __builtin_bit_cast(vint4, __builtin_amdgcn_raw_buffer_load_b128(rci, eo, 0, 0));
          const vfloat4 v4 = __builtin_bit_cast(vfloat4, __builtin_amdgcn_raw_buffer_load_b128(rv, eo, 0, 0));
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            s.cc[4 * q + j] = c4[j];
            s.vv[4 * q + j] = v4[j];
          }
        } else {
          if constexpr (CODED)
            s.cc[q] = (int)__builtin_amdgcn_raw_buffer_load_b16(rcode, e < s.k1 ? e * 2 : kBufferDrop, 0, 0) & 0xffff;
          else
            s.cc[q] = (int)__builtin_amdgcn_raw_buffer_load_b32(rci, eo, 0, 0);
          s.vv[q] = buf_load<TV>(rv, e < s.k1 ? e * (int)sizeof(TV) : kBufferDrop);
        }
      }
    };
    // column of a code: the table slot through ds_bpermute (slot j < 64 from
    // lane j of the first half, 64 + j from lane j of the second)
    auto decode = [&](const StB &s, int cd) {
      const int sl = cd >> 5, lo = __builtin_amdgcn_ds_bpermute((sl & 63) * 4, s.tb0);
      const int hi = s.wide ? __builtin_amdgcn_ds_bpermute((sl & 63) * 4, s.tb1) : 0;
      return (sl < 64 ? lo : hi) + (cd & 31);
    };
    auto stage_c = [&](const StB &s, T *xg) {
      if constexpr (CODED && W == 4) {
#pragma unroll
        for (int u = 0; u < ITEMS / 2; ++u) {
          const unsigned pk = (unsigned)s.cc[(u / 2) * 4 + (u % 2)];
          xg[2 * u] = buf_load<T>(rx, decode(s, (int)(pk & 0xffff)) * (int)sizeof(T));
          xg[2 * u + 1] = buf_load<T>(rx, decode(s, (int)(pk >> 16)) * (int)sizeof(T));
        }
      } else if constexpr (CODED) {
#pragma unroll
        for (int u = 0; u < ITEMS; ++u) xg[u] = buf_load<T>(rx, decode(s, s.cc[u]) * (int)sizeof(T));
      } else {
#pragma unroll
        for (int u = 0; u < ITEMS; ++u) xg[u] = buf_load<T>(rx, s.cc[u] * (int)sizeof(T));
      }
    };
    // three sets of row/entry loads and two of gathers and of descriptors, their
    // roles rotated by unrolling (copying a register whose load is in flight
    // would wait for it)
    StB s0, s1, s2;
    T x0[ITEMS], x1[ITEMS];
    int dA, dB;
    dA = stage_a(0);
    stage_b(dA, s0);
    stage_c(s0, x0);
    dA = stage_a(1);
    stage_b(dA, s1);
    dA = stage_a(2);
    // one wait here, so that the loop's waits count only the loop's own operations
    __builtin_amdgcn_s_waitcnt(0);
    // iteration t: descriptor of t+3 into dn, rows and entries of t+2 (from dc)
    // into nn, gathers of t+1 (from nx) into xn, then block t (cur, xc)
    auto body = [&](int t, int &dc, int &dn, StB &cur, StB &nx, StB &nn, T(&xc)[ITEMS], T(&xn)[ITEMS]) {
      // the stages are issued in this order (the scheduler may not interleave them)
      dn = stage_a(t + 3);
      __builtin_amdgcn_sched_barrier(0);
      stage_b(dc, nn);
      __builtin_amdgcn_sched_barrier(0);
      stage_c(nx, xn);
      __builtin_amdgcn_sched_barrier(0);
      const int base = cur.k0 & ~3;
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        T pw[W];
#pragma unroll
        for (int j = 0; j < W; ++j) pw[j] = (T)cur.vv[W * q + j] * xc[W * q + j];
        if constexpr (W == 4) store4(prod + 4 * (tid + q * NT), pw);
        else prod[tid + q * NT] = pw[0];
      }
      __syncthreads();
      const bool own = cur.rs + tid < cur.re;
      T acc = 0;
      {
        const int lo = own ? cur.r0 : base, hi = own ? cur.r1 : base;
        for (int k = lo; k < hi; ++k) acc += prod[k - base];
      }
      const T out = b ? (T)(cur.bb - acc) : acc;
      buf_store<T>(out, ry, own ? (cur.rs + tid) * (int)sizeof(T) : kBufferDrop);
      if (own) {
        a0 += d0 ? (double)(d0y ? out : cur.e0) * (double)out : 0.0;
        a1 += d1 ? (double)(d1y ? out : cur.e1) * (double)out : 0.0;
      }
      __syncthreads();
    };
    for (int t = 0; t < count; t += 6) {
      body(t, dA, dB, s0, s1, s2, x0, x1);
      if (t + 1 >= count) break;
      body(t + 1, dB, dA, s1, s2, s0, x1, x0);
      if (t + 2 >= count) break;
      body(t + 2, dA, dB, s2, s0, s1, x0, x1);
      if (t + 3 >= count) break;
      body(t + 3, dB, dA, s0, s1, s2, x1, x0);
      if (t + 4 >= count) break;
      body(t + 4, dA, dB, s1, s2, s0, x0, x1);
      if (t + 5 >= count) break;
      body(t + 5, dB, dA, s2, s0, s1, x1, x0);
    }
  }
  if (part) {
    double vv[2] = {a0, a1};
    block_sum<2>(vv);
    if (threadIdx.x == 0) {
      part[blockIdx.x] = vv[0];
      part[kPartStride + blockIdx.x] = vv[1];
    }
  }
}

// Compile-time unrolled loop: f(std::integral_constant<int, J>) for J = 0..N-1
// (the DPP lane selects below must be constants).
template <typename F, int... J>
__device__ inline void unroll_seq(F &&f, std::integer_sequence<int, J...>) {
  (f(std::integral_constant<int, J>{}), ...);
}
template <int N, typename F>
__device__ inline void unroll(F &&f) {
  unroll_seq(f, std::make_integer_sequence<int, N>{});
}

// Value of `x` in lane L of this lane's row of 16 lanes (DPP row_newbcast).
template <int L>
__device__ inline float row_bcast(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x150 + L, 0xF, 0xF, true));
}
template <int L>
__device__ inline double row_bcast(double x) {
  const long long bits = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_mov_dpp((int)bits, 0x150 + L, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(bits >> 32), 0x150 + L, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
// acc += c * (src of lane L of this lane's row of 16).  FP32: one
// v_fmac_f32_dpp (the broadcast applied to the multiplicand), after the two
// wait states a DPP read of a VGPR written by the previous VALU op needs.
template <int L>
__device__ inline void fma_bcast(float &acc, float src, float c) {
  asm("s_nop 1\n\tv_fmac_f32_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
      : "+v"(acc) : "v"(src), "v"(c), "n"(L));
}
template <int L>
__device__ inline void fma_bcast_self(float &acc, float c) {
  asm("s_nop 1\n\tv_fmac_f32_dpp %0, %0, %1 row_newbcast:%2 row_mask:0xf bank_mask:0xf" : "+v"(acc) : "v"(c), "n"(L));
}
template <int L>
__device__ inline void fma_bcast(double &acc, double src, double c) {
  acc = fma(c, row_bcast<L>(src), acc);
}
template <int L>
__device__ inline void fma_bcast_self(double &acc, double c) {
  acc = fma(c, row_bcast<L>(acc), acc);
}
// Rows of 16 lanes exchanged in pairs (v_permlane16_swap): lanes 16-31 (and
// 48-63) receive the value of lanes 0-15 (32-47) when `upper` is false, and
// lanes 0-15 (32-47) that of lanes 16-31 (48-63) when it is true.
template <bool UPPER>
__device__ inline float row_swap(float x) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
  return __uint_as_float(UPPER ? r[1] : r[0]);
}
template <bool UPPER>
__device__ inline double row_swap(double x) {
  const long long bits = __double_as_longlong(x);
  const unsigned lo = (unsigned)bits, hi = (unsigned)(bits >> 32);
  const auto a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
  const auto b = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
  const unsigned rl = UPPER ? a[1] : a[0], rh = UPPER ? b[1] : b[0];
  return __longlong_as_double((long long)(((unsigned long long)rh << 32) | rl));
}

// One colour of a block multicolour sweep.  Blocks [b0, b1) of the solve-space
// block starts bstart (at most kBmcRows rows each; blocks of one colour are
// independent); each triangle is split at set-up into the entries that couple
// to other blocks (columns of other colours, already final), a CSR by row
// (orp, oci, ov; bK[blk] = orp at the block's first row), and the entries
// inside each block as (LDS slot, value) lists per block (iptr, icode, iv).
//   forward (unit L):  z[row] = in[row] - sum_k L[k] z[ci[k]]
//   backward:          z[row] = dinv[row] * (z[row] - sum_k U[k] z[ci[k]])
// A workgroup takes kBmcGroupOf<T> consecutive blocks of the colour (groups
// counted from the colour's first block; with xcd set each XCD, workgroup id
// mod 8, takes a contiguous eighth of them), half a wavefront per block and
// one lane per row.  It streams the group's coupling entries as one contiguous
// range (consecutive lanes, consecutive entries; with FP32 values four entries
// per lane in one 16-byte load per array), multiplies each by its
// gathered z into LDS and sums each row's run, while the in-block lists fill
// each block's coefficient tile.  Then each block is eliminated in registers:
// in the elimination order (forward: rows ascending; backward: descending,
// written m = 31 - row below) the block's first 16 rows form one DPP row of the
// half wavefront and the last 16 the other.  The first row's lanes eliminate
// among themselves, each step broadcasting the finished value of lane j to its
// row (row_bcast) and adding coefficient times value; the finished values are
// then handed to the second row's lanes (row_swap), which add their couplings
// to all 16 and eliminate among themselves the same way.  The tile holds, for
// lane m and step j (both in elimination order), the negated coefficient:
// -L(row, col) forward, -U(row, col) / diag(col) backward (so that the value
// broadcast is the unscaled one); first-row lanes at j*16 + m, second-row lanes
// at kBmcSecond + j*16 + (m - 16), each step's reads being 16 consecutive words
// per row of lanes, the two rows in disjoint banks.
constexpr int kBmcHalf = 16;                                // lanes of a DPP row
constexpr int kBmcSecond = 272;                             // second-row region (16 mod 32 words)
constexpr int kBmcTile = kBmcSecond + kBmcRows * kBmcHalf;  // words per block
constexpr int kBmcGroupMax = 8;                             // blocks per workgroup (FP32)
constexpr int kBmcItems = 4;                                // coupling entries per thread held per group
constexpr int kBmcIn = 2;                                   // in-block entries per thread held per group
static_assert(kBmcGroupMax * kBmcTile < 65536, "in-block slot codes are 16-bit");
static_assert(kBmcSecond >= kBmcHalf * kBmcHalf && kBmcSecond % 32 == 16, "tile regions");
template <typename T>
constexpr int kBmcGroupOf = sizeof(T) == 8 ? 4 : 8;  // FP64: 4 blocks, so that the tile fits beside others
// slot of the in-block coefficient of (row r, column j), local to the block;
// j < r forward, j > r backward
__host__ __device__ constexpr int bmc_slot(bool backward, int j, int r) {
  const int m = backward ? kBmcRows - 1 - r : r, jm = backward ? kBmcRows - 1 - j : j;
  return m < kBmcHalf ? jm * kBmcHalf + m : kBmcSecond + jm * kBmcHalf + (m - kBmcHalf);
}

template <bool BACKWARD, typename TV, typename T>
__global__ __launch_bounds__(kBmcGroupOf<T> * kBmcRows) void k_bmc_group(
    int b0, int b1, int n, int nb, int64_t no, int64_t ni, const int *__restrict__ bstart, const int *__restrict__ bK,
    const int *__restrict__ orp, const int *__restrict__ oci, const TV *__restrict__ ov, const int *__restrict__ iptr,
    const unsigned short *__restrict__ icode, const TV *__restrict__ iv, const TV *__restrict__ dinv,
    const T *__restrict__ in, T *z) {
  constexpr int G = kBmcGroupOf<T>, NT = G * kBmcRows, CAP = kBmcItems * NT, ICAP = kBmcIn * NT;
  __shared__ __align__(16) T tile[G * kBmcTile];
  __shared__ T prod[CAP];
  // persistent: the groups of XCD x (workgroup id mod 8) are a contiguous
  // range, and the P workgroups of that XCD take every P-th group of it
  const int ngroups = (b1 - b0 + G - 1) / G;
  const int P = (int)gridDim.x / kXcds, x = (int)blockIdx.x % kXcds, wi = (int)blockIdx.x / kXcds;
  const int per = (ngroups + kXcds - 1) / kXcds;
  const int xe = min(ngroups, (x + 1) * per), first = x * per + wi;
  if (first >= xe) return;  // whole workgroup
  const int count = (xe - first + P - 1) / P, last = first + (count - 1) * P;
  const int tid = (int)threadIdx.x, sub = tid / kBmcRows, hl = tid % kBmcRows;
  // buffer resources: loads past an array read zero, stores past it are dropped
  const auto rbst = buf_rsrc(bstart, (int64_t)(nb + 1) * 4), rbK = buf_rsrc(bK, (int64_t)(nb + 1) * 4);
  const auto rip = buf_rsrc(iptr, (int64_t)(nb + 1) * 4), rorp = buf_rsrc(orp, (int64_t)(n + 1) * 4);
  // the coupling arrays carry kVecPad entries past the end (16-byte loads of the last group)
  const auto roci = buf_rsrc(oci, (no + kVecPad) * 4), rov = buf_rsrc(ov, (no + (sizeof(TV) == 4 ? kVecPad : 0)) * (int64_t)sizeof(TV));
  const auto ricode = buf_rsrc(icode, ni * 2), riv = buf_rsrc(iv, ni * (int64_t)sizeof(TV));
  const auto rdinv = buf_rsrc(BACKWARD ? (const void *)dinv : nullptr, (int64_t)n * sizeof(TV));
  const auto rin = buf_rsrc(BACKWARD ? (const void *)z : (const void *)in, (int64_t)n * sizeof(T));
  const auto rz = buf_rsrc(z, (int64_t)n * sizeof(T));

  // Three load stages per group, issued in consecutive iterations so that each
  // stage's latency overlaps the elimination of an earlier group.  Past the
  // last group the stages load the last group again.
  struct StA {  // the lane's block start and end; the group's coupling and in-block bounds (lane parity picks which)
    int g0, gn, rlo, rhi, kb, ip;
  };
  struct StB {  // the lane's row, the group's entries (up to CAP / ICAP)
    int row, nr, K0, K1, I0, I1, ks, ke;
    T acc0, d;
    int cc[kBmcItems];
    TV vv[kBmcItems];
    int ic[kBmcIn];
    TV ivv[kBmcIn];
  };
  auto stage_a = [&](int t, StA &a) {
    const int g = min(first + t * P, last);
    a.g0 = b0 + g * G;
    a.gn = min(G, b1 - a.g0);
    const int sb = a.g0 + min(sub, a.gn - 1), gb = (tid & 1) ? a.g0 + a.gn : a.g0;
    a.rlo = (int)__builtin_amdgcn_raw_buffer_load_b32(rbst, sb * 4, 0, 0);
    a.rhi = (int)__builtin_amdgcn_raw_buffer_load_b32(rbst, sb * 4 + 4, 0, 0);
    a.kb = (int)__builtin_amdgcn_raw_buffer_load_b32(rbK, gb * 4, 0, 0);
    a.ip = (int)__builtin_amdgcn_raw_buffer_load_b32(rip, gb * 4, 0, 0);
  };
  auto stage_b = [&](const StA &a, StB &b) {
    // lane 0 of every wavefront holds the group's first bound, lane 1 its last
    b.K0 = __builtin_amdgcn_readlane(a.kb, 0);
    b.K1 = __builtin_amdgcn_readlane(a.kb, 1);
    b.I0 = __builtin_amdgcn_readlane(a.ip, 0);
    b.I1 = __builtin_amdgcn_readlane(a.ip, 1);
    b.nr = sub < a.gn ? a.rhi - a.rlo : 0;
    b.row = a.rlo + hl;
    b.acc0 = buf_load<T>(rin, b.row * (int)sizeof(T));
    b.d = BACKWARD ? (T)buf_load<TV>(rdinv, b.row * (int)sizeof(TV)) : (T)1;
    b.ks = (int)__builtin_amdgcn_raw_buffer_load_b32(rorp, b.row * 4, 0, 0);
    b.ke = (int)__builtin_amdgcn_raw_buffer_load_b32(rorp, b.row * 4 + 4, 0, 0);
    if constexpr (sizeof(TV) == 4 && kBmcItems == 4) {
      // entries (K0 rounded down to four) + 4 tid + u in one 16-byte load per
      // array; a load starting past K1 is not made
      const int e = (b.K0 & ~3) + 4 * tid;
      const int eo = e < b.K1 ? e * 4 : kBufferDrop;
      const vint4 c4 = __builtin_bit_cast(vint4, __builtin_amdgcn_raw_buffer_load_b128(roci, eo, 0, 0));
      const vfloat4 v4 = __builtin_bit_cast(vfloat4, __builtin_amdgcn_raw_buffer_load_b128(rov, eo, 0, 0));
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        b.cc[u] = c4[u];
        b.vv[u] = v4[u];
      }
    } else {
#pragma unroll
      for (int u = 0; u < kBmcItems; ++u) {
        const int k = b.K0 + tid + u * NT;  // past K1: not loaded (zero), never used
        b.cc[u] = (int)__builtin_amdgcn_raw_buffer_load_b32(roci, k < b.K1 ? k * 4 : kBufferDrop, 0, 0);
        b.vv[u] = buf_load<TV>(rov, k < b.K1 ? k * (int)sizeof(TV) : kBufferDrop);
      }
    }
#pragma unroll
    for (int u = 0; u < kBmcIn; ++u) {
      const int k = b.I0 + tid + u * NT;
      b.ic[u] = (int)__builtin_amdgcn_raw_buffer_load_b16(ricode, k < b.I1 ? k * 2 : kBufferDrop, 0, 0) & 0xffff;
      b.ivv[u] = buf_load<TV>(riv, k < b.I1 ? k * (int)sizeof(TV) : kBufferDrop);
    }
  };
  // the gathered values of the coupling entries (entries past the group's
  // range gather some value that is never summed)
  auto stage_c = [&](const StB &b, T *zz) {
#pragma unroll
    for (int u = 0; u < kBmcItems; ++u) zz[u] = buf_load<T>(rz, b.cc[u] * (int)sizeof(T));
  };

  StA a;
  StB bc, bn, bnn;
  T zc[kBmcItems], zn[kBmcItems];
  stage_a(0, a);
  stage_b(a, bc);
  stage_c(bc, zc);
  stage_a(1, a);
  stage_b(a, bn);
  stage_a(2, a);
  // one wait here, so that the loop's waits count only the loop's own operations
  __builtin_amdgcn_s_waitcnt(0);
  for (int t = 0; t < count; ++t) {
    stage_c(bn, zn);
    stage_b(a, bnn);
    stage_a(t + 3, a);
    // eliminate the group of bc / zc
    {
      float4 *t4 = reinterpret_cast<float4 *>(tile);
      static_assert((G * kBmcTile * sizeof(T)) % 16 == 0, "tile cleared 16 bytes at a time");
      for (int i = tid; i < (int)(G * kBmcTile * sizeof(T) / 16); i += NT) t4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    __syncthreads();  // tiles cleared
    constexpr bool V4 = sizeof(TV) == 4 && kBmcItems == 4;
    // products are held from kb (K0, rounded down to four with 16-byte loads)
    const int kb = V4 ? (bc.K0 & ~3) : bc.K0;
    const int icnt = bc.I1 - bc.I0, ccnt = min(bc.K1, kb + CAP) - bc.K0;
#pragma unroll
    for (int u = 0; u < kBmcIn; ++u)
      if (tid + u * NT < icnt) {
        const int code = bc.ic[u];
        tile[((code / kBmcTile) % G) * kBmcTile + code % kBmcTile] = (T)bc.ivv[u];
      }
    for (int k = bc.I0 + ICAP + tid; k < bc.I1; k += NT) {  // groups beyond the held entries
      const int code = icode[k];
      tile[((code / kBmcTile) % G) * kBmcTile + code % kBmcTile] = (T)iv[k];
    }
#pragma unroll
    for (int u = 0; u < kBmcItems; ++u) prod[V4 ? 4 * tid + u : tid + u * NT] = (T)bc.vv[u] * zc[u];  // outside [K0, K0 + ccnt): never summed
    __syncthreads();  // in-block values and products in place
    const bool own = hl < bc.nr;
    T s = 0;
    {
      const int lo = own ? bc.ks : bc.K0, hi = own ? min(bc.ke, bc.K0 + ccnt) : bc.K0;
      for (int k = lo; k < hi; ++k) s += prod[k - kb];
    }
    // coupling entries beyond the held ones, a chunk at a time (uniform)
    for (int kc = kb + CAP; kc < bc.K1; kc += CAP) {
      __syncthreads();
      const int cnt = min(CAP, bc.K1 - kc);
      for (int k = tid; k < cnt; k += NT) prod[k] = (T)ov[kc + k] * z[oci[kc + k]];
      __syncthreads();
      if (own)
        for (int k = max(bc.ks, kc); k < min(bc.ke, kc + cnt); ++k) s += prod[k - kc];
    }
    T acc = bc.acc0 - s;
    // elimination order: m = row (forward) or 31 - row (backward); source lane of
    // step j inside its row of 16: j (forward) or 15 - j (backward), mod 16
    const int m = BACKWARD ? kBmcRows - 1 - hl : hl;
    const bool second = m >= kBmcHalf;
    const T *tl = tile + sub * kBmcTile + (second ? kBmcSecond + (m - kBmcHalf) : m);
    T cf[kBmcRows];
#pragma unroll
    for (int j = 0; j < kBmcRows; ++j) cf[j] = (second || j < kBmcHalf) ? tl[j * kBmcHalf] : (T)0;
    __syncthreads();  // tiles and products free for the next group
    if (!second) {
      unroll<kBmcHalf>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        fma_bcast_self<BACKWARD ? kBmcHalf - 1 - j : j>(acc, cf[j]);
      });
    }
    const T tv = row_swap<BACKWARD>(acc);  // second-row lanes: the first row's finished values
    if (second) {
      unroll<kBmcHalf>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        fma_bcast<BACKWARD ? kBmcHalf - 1 - j : j>(acc, tv, cf[j]);
      });
      unroll<kBmcHalf>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        fma_bcast_self<BACKWARD ? kBmcHalf - 1 - j : j>(acc, cf[kBmcHalf + j]);
      });
    }
    buf_store<T>(BACKWARD ? acc * bc.d : acc, rz, own ? bc.row * (int)sizeof(T) : kBufferDrop);
    bc = bn;
    bn = bnn;
#pragma unroll
    for (int u = 0; u < kBmcItems; ++u) zc[u] = zn[u];
  }
}

// Row-list sweeps for levels that are not contiguous ranges (natural-order ILU0).
template <int LPR, typename TV, typename T>
__global__ __launch_bounds__(kBlock) void k_lsweep_rows(int cnt, const int *__restrict__ rows, const int *__restrict__ rp,
                                                        const int *__restrict__ ci, const TV *__restrict__ v,
                                                        const T *__restrict__ rhs, T *out, const int *__restrict__ perm) {
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = gtid / LPR, lane = gtid & (LPR - 1);
  if (g >= cnt) return;
  const int row = rows[g];
  T acc = 0;
  for (int k = rp[row] + lane; k < rp[row + 1]; k += LPR) acc += (T)v[k] * out[ci[k]];
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LPR);
  if (lane == 0) out[row] = (perm ? rhs[perm[row]] : rhs[row]) - acc;
}
template <int LPR, typename TV, typename T>
__global__ __launch_bounds__(kBlock) void k_usweep_rows(int cnt, const int *__restrict__ rows, const int *__restrict__ rp,
                                                        const int *__restrict__ ci, const TV *__restrict__ v,
                                                        const T *__restrict__ dinv, T *out, const int *__restrict__ perm,
                                                        T *__restrict__ out2) {
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = gtid / LPR, lane = gtid & (LPR - 1);
  if (g >= cnt) return;
  const int row = rows[g];
  T acc = 0;
  for (int k = rp[row] + lane; k < rp[row + 1]; k += LPR) acc += (T)v[k] * out[ci[k]];
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LPR);
  if (lane == 0) {
    const T val = dinv[row] * (out[row] - acc);
    out[row] = val;
    if (perm) out2[perm[row]] = val;
  }
}

// Split of the factor into strict lower / strict upper parts and 1/diagonal.
__global__ __launch_bounds__(kBlock) void k_split(int n, const int *__restrict__ frp, const int *__restrict__ fdg,
                                                  const double *__restrict__ f, const int *__restrict__ Lrp,
                                                  const int *__restrict__ Urp, double *__restrict__ L,
                                                  double *__restrict__ U, double *__restrict__ dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = frp[i], d = fdg[i], e = frp[i + 1];
  for (int k = b; k < d; ++k) L[Lrp[i] + (k - b)] = f[k];
  for (int k = d + 1; k < e; ++k) U[Urp[i] + (k - d - 1)] = f[k];
  dinv[i] = 1.0 / f[d];
}

// ---------------------------------------------------------------- vector kernels
template <typename T>
__global__ __launch_bounds__(kBlock) void k_copy_dot(int n, const T *__restrict__ a, T *__restrict__ out,
                                                     double *__restrict__ part) {
  // out = a; partial (a, a)
  double s = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    T v = a[i];
    out[i] = v;
    s += (double)v * (double)v;
  }
  double vv[1] = {s};
  block_sum<1>(vv);
  if (threadIdx.x == 0) part[blockIdx.x] = vv[0];
}

// per-workgroup partial sums of (x - xr, x - xr) and (xr, xr)
__global__ __launch_bounds__(kBlock) void k_err_partials(int n, const double *__restrict__ x,
                                                         const double *__restrict__ xr, double *__restrict__ part) {
  double a0 = 0, a1 = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double d = x[i] - xr[i];
    a0 += d * d;
    a1 += xr[i] * xr[i];
  }
  double vv[2] = {a0, a1};
  block_sum<2>(vv);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = vv[0];
    part[kPartStride + blockIdx.x] = vv[1];
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_update_p(int n, T *__restrict__ p, const T *__restrict__ r,
                                                     const T *__restrict__ v, const Scalars *__restrict__ sc) {
  if (sc->halt) return;
  const T beta = (T)sc->beta, omega = (T)sc->omega;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = r[i] + beta * (p[i] - omega * v[i]);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_update_s(int n, T *__restrict__ s, const T *__restrict__ r,
                                                     const T *__restrict__ v, const Scalars *__restrict__ sc) {
  if (sc->halt) return;
  const T alpha = (T)sc->alpha;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s[i] = r[i] - alpha * v[i];
}

// x += alpha*ph + omega*sh ; r = s - omega*t ; partials (r,r), (r0,r)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_update_xr(int n, T *__restrict__ x, T *__restrict__ r,
                                                      const T *__restrict__ ph, const T *__restrict__ sh,
                                                      const T *__restrict__ s, const T *__restrict__ t,
                                                      const T *__restrict__ r0, const Scalars *__restrict__ sc,
                                                      double *__restrict__ part) {
  if (sc->halt) return;
  const T alpha = (T)sc->alpha, omega = (T)sc->omega;
  double a0 = 0, a1 = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    x[i] += alpha * ph[i] + omega * sh[i];
    T ri = s[i] - omega * t[i];
    r[i] = ri;
    a0 += (double)ri * (double)ri;
    a1 += (double)r0[i] * (double)ri;
  }
  double vv[2] = {a0, a1};
  block_sum<2>(vv);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = vv[0];
    part[kPartStride + blockIdx.x] = vv[1];
  }
}

__global__ __launch_bounds__(kBlock) void k_gather(int64_t n, const int *__restrict__ pos, const double *__restrict__ src,
                                                   double *__restrict__ dst) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = pos[i] >= 0 ? src[pos[i]] : 0.0;  // fill positions of ILU(k) start at zero
}

// dir 0: out[i] = in[perm[i]] (to the permuted ordering); 1: out[perm[i]] = in[i] (back)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_permute(int n, const int *__restrict__ perm, const T *__restrict__ in,
                                                    T *__restrict__ out, int dir) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (dir == 0) out[i] = in[perm[i]];
  else out[perm[i]] = in[i];
}

// FP64 iterative-refinement helpers
__global__ __launch_bounds__(kBlock) void k_cast_scale(int n, const double *__restrict__ a, float *__restrict__ out,
                                                       const double *__restrict__ nrm) {
  const double s = nrm[0] > 0 ? 1.0 / nrm[0] : 1.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = (float)(a[i] * s);
}
__global__ __launch_bounds__(kBlock) void k_axpy_f32(int n, double *__restrict__ x, const float *__restrict__ d,
                                                     const double *__restrict__ nrm) {
  const double s = nrm[0];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] += (double)d[i] * s;
}
template <typename TI, typename TO>
__global__ __launch_bounds__(kBlock) void k_cast(int64_t n, const TI *__restrict__ a, TO *__restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (TO)a[i];
}
// In-block coefficients of the block multicolour sweep: -src[pos], divided by
// the column's diagonal when dinv is given (backward triangle), in FP64 and FP32.
__global__ __launch_bounds__(kBlock) void k_bmc_inblock(int64_t n, const int *__restrict__ pos,
                                                       const int *__restrict__ col, const double *__restrict__ src,
                                                       const double *__restrict__ dinv, double *__restrict__ o64,
                                                       float *__restrict__ o32) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = -src[pos[i]];
  if (dinv) v *= dinv[col[i]];
  o64[i] = v;
  o32[i] = (float)v;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(int64_t n, T *__restrict__ a, T v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) a[i] = v;
}

// ---------------------------------------------------------------- finalise kernels (1 block)
__device__ inline double reduce_part(const double *part, int count) {
  __shared__ double red[kBlock / 64];
  double s = 0;
  for (int i = threadIdx.x; i < count; i += blockDim.x) s += part[i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  double t = 0;
  for (int w = 0; w < kBlock / 64; ++w) t += red[w];
  __syncthreads();
  return t;
}

// op: 0 = init (rr -> bnorm/res, rho=alpha=omega=1, r0r=rr)
//     1 = after v = A ph: alpha = rho / r0v
//     2 = after t = A sh: omega = ts/tt
//     3 = after x/r update: res, beta for next iteration
//     4 = store norm of partial sum (sqrt) into slot
__global__ __launch_bounds__(kBlock) void k_finalize(int op, int cnt, const double *__restrict__ part, Scalars *sc,
                                                     double *__restrict__ out_norm) {
  if (op >= 1 && op <= 3 && sc->halt) return;  // an iteration enqueued behind a halt
  if (op == 0) {
    double rr = reduce_part(part, cnt);
    if (threadIdx.x == 0) {
      sc->rr = rr;
      sc->bnorm = sqrt(rr);
      sc->res = 1.0;
      sc->rho = sc->alpha = sc->omega = 1.0;
      sc->r0r = rr;
      sc->beta = 0.0;
      sc->breakdown = 0;
      sc->halt = 0;
      // first direction: p = r (beta = 0 with p = v = 0)
      sc->beta = (sc->r0r / sc->rho) * (sc->alpha / sc->omega);
      sc->rho = sc->r0r;
    }
  } else if (op == 1) {
    double r0v = reduce_part(part, cnt);
    if (threadIdx.x == 0) {
      sc->r0v = r0v;
      if (r0v == 0.0 || !isfinite(r0v)) { sc->breakdown = 1; sc->alpha = 0.0; }
      else sc->alpha = sc->rho / r0v;
    }
  } else if (op == 2) {
    double ts = reduce_part(part, cnt);
    double tt = reduce_part(part + kPartStride, cnt);
    if (threadIdx.x == 0) {
      sc->ts = ts;
      sc->tt = tt;
      sc->omega = tt > 0.0 ? ts / tt : 0.0;
    }
  } else if (op == 3) {
    double rr = reduce_part(part, cnt);
    double r0r = reduce_part(part + kPartStride, cnt);
    if (threadIdx.x == 0) {
      sc->rr = rr;
      sc->r0r = r0r;
      sc->res = sc->bnorm > 0 ? sqrt(rr) / sc->bnorm : 0.0;
      if (sc->omega == 0.0 || r0r == 0.0 || !isfinite(rr)) {
        sc->breakdown = 1;
      } else {
        sc->beta = (r0r / sc->rho) * (sc->alpha / sc->omega);
        sc->rho = r0r;
      }
      if (sc->breakdown || !(sc->res > sc->tol)) sc->halt = 1;  // also a NaN residual
    }
  } else if (op == 4) {
    double s = reduce_part(part, cnt);
    if (threadIdx.x == 0) out_norm[0] = sqrt(s);
  }
}

// ---------------------------------------------------------------- level-scheduled ILU(0)
// ILU(0) of the rows of one level; one 64-lane block per row, row kept in LDS.
// Row-wise IKJ order (Saad, Iterative Methods, 2nd ed., Alg. 10.4).
// drop > 0: threshold rule of ILUT inside the fixed pattern -- multipliers with
// |l| < drop * ||a_row||_2 are set to zero and not used, upper entries below the
// same bound are zeroed once the row is done.
__global__ __launch_bounds__(64) void k_ilu0_level(int cnt, const int *__restrict__ rows, const int *__restrict__ rp,
                                                   const int *__restrict__ ci, const int *__restrict__ dg,
                                                   double *__restrict__ fv, double drop) {
  __shared__ int cols[kMaxRowLDS];
  __shared__ double vals[kMaxRowLDS];
  if ((int)blockIdx.x >= cnt) return;
  const int row = rows[blockIdx.x];
  const int beg = rp[row], len = rp[row + 1] - beg, nl = dg[row] - beg;
  double sq = 0.0;
  for (int t = threadIdx.x; t < len; t += 64) {
    cols[t] = ci[beg + t];
    vals[t] = fv[beg + t];
    sq += vals[t] * vals[t];
  }
  const double tau = drop > 0.0 ? drop * sqrt(wave_sum(sq)) : 0.0;
  __syncthreads();
  for (int t = 0; t < nl; ++t) {
    const int kc = cols[t];
    double l = vals[t] / fv[dg[kc]];
    if (drop > 0.0 && fabs(l) < tau) {
      if (threadIdx.x == 0) vals[t] = 0.0;
      __syncthreads();
      continue;
    }
    const int ub = dg[kc] + 1, ue = rp[kc + 1];
    for (int q = ub + threadIdx.x; q < ue; q += 64) {
      const int j = ci[q];
      int lo = t + 1, hi = len;
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (cols[mid] < j) lo = mid + 1; else hi = mid;
      }
      if (lo < len && cols[lo] == j) vals[lo] -= l * fv[q];
    }
    if (threadIdx.x == 0) vals[t] = l;
    __syncthreads();
  }
  for (int t = threadIdx.x; t < len; t += 64) {
    double v = vals[t];
    if (drop > 0.0 && t > nl && fabs(v) < tau) v = 0.0;
    fv[beg + t] = v;
  }
}

// The same factorisation for short rows, 8 lanes per row and 32 rows per
// workgroup, everything in registers: lane k of a row's group holds the
// row's positions k, k + 8, k + 16, k + 24, and for each L entry the U part of
// that entry's row is loaded across the group and applied one entry at a time
// (broadcast column, the lane owning that column updates it).  The updates
// run in the same order as k_ilu0_level's, so the factor is the same.  Rows
// and the U parts they read must have at most kShortRow entries.
constexpr int kShortRow = 32;
constexpr int kShortLanes = 8;
__global__ __launch_bounds__(256) void k_ilu0_short(int cnt, const int *__restrict__ rows, const int *__restrict__ rp,
                                                    const int *__restrict__ ci, const int *__restrict__ dg,
                                                    double *__restrict__ fv, double drop) {
  constexpr int E = kShortRow / kShortLanes;
  const int gid = (int)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / kShortLanes);
  const int lane = threadIdx.x & (kShortLanes - 1);
  const bool ok = gid < cnt;
  const int row = ok ? rows[gid] : 0;
  const int beg = ok ? rp[row] : 0, len = ok ? rp[row + 1] - beg : 0, nl = ok ? dg[row] - beg : 0;
  int col[E];
  double val[E];
  double sq = 0.0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int t = lane + kShortLanes * e;
    col[e] = t < len ? ci[beg + t] : -1;
    val[e] = t < len ? fv[beg + t] : 0.0;
    sq += val[e] * val[e];
  }
  for (int off = kShortLanes / 2; off > 0; off >>= 1) sq += __shfl_xor(sq, off, kShortLanes);
  const double tau = drop > 0.0 ? drop * sqrt(sq) : 0.0;
  // groups run their own trip counts; every shuffle below stays inside the group
  for (int t = 0; t < nl; ++t) {
    const int owner = t % kShortLanes, et = t / kShortLanes;
    int kc = 0;
    double vt = 0.0;
#pragma unroll
    for (int e = 0; e < E; ++e)
      if (e == et) { kc = col[e]; vt = val[e]; }
    kc = __shfl(kc, owner, kShortLanes);
    vt = __shfl(vt, owner, kShortLanes);
    const double l = vt / fv[dg[kc]];
    const bool dropped = drop > 0.0 && fabs(l) < tau;
    if (lane == owner) {
#pragma unroll
      for (int e = 0; e < E; ++e)
        if (e == et) val[e] = dropped ? 0.0 : l;
    }
    if (dropped) continue;
    const int ub = dg[kc] + 1, m = rp[kc + 1] - ub;
    int uj[E];
    double uv[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int q = lane + kShortLanes * e;
      uj[e] = q < m ? ci[ub + q] : -1;
      uv[e] = q < m ? fv[ub + q] : 0.0;
    }
    for (int q = 0; q < m; ++q) {
      const int qo = q % kShortLanes, qe = q / kShortLanes;
      int j = 0;
      double u = 0.0;
#pragma unroll
      for (int e = 0; e < E; ++e)
        if (e == qe) { j = uj[e]; u = uv[e]; }
      j = __shfl(j, qo, kShortLanes);
      u = __shfl(u, qo, kShortLanes);
#pragma unroll
      for (int e = 0; e < E; ++e)
        if (col[e] == j) val[e] -= l * u;
    }
  }
  if (!ok) return;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int t = lane + kShortLanes * e;
    if (t < len) {
      double v = val[e];
      if (drop > 0.0 && t > nl && fabs(v) < tau) v = 0.0;
      fv[beg + t] = v;
    }
  }
}

// ---------------------------------------------------------------- batched BiCGStab
// Systems are concatenated block-diagonally; system s owns rows
// [sys_ptr[s], sys_ptr[s+1]).  Every system runs its own recurrence: scalars,
// dot products, stopping test and a convergence mask per system, all on the
// device, so that a sequence of iterations needs no host round trip.
constexpr int kBatchParts = 32;  // partial sums per system and dot product
constexpr int kStallWindow = 100; // FP32 inner solves: iterations without a 1 % gain before giving up
constexpr int kBatchStallWindow = 40;  // the same in batches, where a stalled system holds up the batch
constexpr double kStallKeep = 0.1;     // batches: a stalled FP32 correction that cut the residual 10x keeps FP32


// NV dot products per system into part[(q * S + s) * kBatchParts + j]
// (systems with mask[s] == 0, or with bs[s].active == 0 when mask is NULL, are skipped)
template <typename T, int NV>
__global__ __launch_bounds__(kBlock) void k_bdot(int S, const int *__restrict__ sys_ptr, const BSys *__restrict__ bs,
                                                 const int *__restrict__ mask, const T *a0, const T *b0, const T *a1,
                                                 const T *b1, double *__restrict__ part) {
  const int s = blockIdx.y, j = blockIdx.x;
  if (mask ? !mask[s] : !bs[s].active) return;
  double vv[2] = {0.0, 0.0};
  const int r0 = sys_ptr[s], len = sys_ptr[s + 1] - r0;
  const int lo = r0 + (int)((int64_t)len * j / kBatchParts), hi = r0 + (int)((int64_t)len * (j + 1) / kBatchParts);
  for (int i = lo + (int)threadIdx.x; i < hi; i += kBlock) {
    vv[0] += (double)a0[i] * (double)b0[i];
    if (NV > 1) vv[1] += (double)a1[i] * (double)b1[i];
  }
  block_sum<2>(vv);
  if (threadIdx.x == 0) {
    part[(0 * S + s) * kBatchParts + j] = vv[0];
    if (NV > 1) part[(1 * S + s) * kBatchParts + j] = vv[1];
  }
}

__device__ inline double bsum(const double *part, int S, int q, int s) {
  double t = 0.0;
  for (int j = 0; j < kBatchParts; ++j) t += part[(q * S + s) * kBatchParts + j];
  return t;
}

// Per-system finalise, one thread per system:
//   0 init from (r,r), (b,b)      1 alpha = rho / (r0,v)      2 omega = (t,s)/(t,t)
//   3 stopping test from (r,r), (r0,r); beta, rho for the next iteration
//   5 nrm[s] = sqrt(dot 0) for every system (no mask)
__global__ void k_bfinalize(int op, int S, const double *__restrict__ part, BSys *bs, double *__restrict__ nrm) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  if (op == 5) {
    nrm[s] = sqrt(bsum(part, S, 0, s));
    return;
  }
  BSys &c = bs[s];
  if (!c.active) return;
  if (op == 0) {
    const double rr = bsum(part, S, 0, s), bb = bsum(part, S, 1, s);
    c.bnorm = sqrt(bb);
    c.res = c.bnorm > 0 ? sqrt(rr) / c.bnorm : 0.0;
    c.rho = rr;
    c.alpha = c.omega = 1.0;
    c.beta = 0.0;
    c.bd = 0;
    c.iters = 0;
    c.best = c.res;
    c.best_it = 0;
    c.status = 1;
    if (c.bnorm == 0.0 || c.res <= c.tol || rr == 0.0) {
      c.status = 0;
      c.active = 0;
    }
  } else if (op == 1) {
    const double r0v = bsum(part, S, 0, s);
    if (r0v == 0.0 || !isfinite(r0v)) { c.bd = 1; c.alpha = 0.0; }
    else c.alpha = c.rho / r0v;
  } else if (op == 2) {
    const double ts = bsum(part, S, 0, s), tt = bsum(part, S, 1, s);
    c.omega = tt > 0.0 ? ts / tt : 0.0;
  } else if (op == 3) {
    const double rr = bsum(part, S, 0, s), r0r = bsum(part, S, 1, s);
    c.res = c.bnorm > 0 ? sqrt(rr) / c.bnorm : 0.0;
    c.iters += 1;
    if (c.res < 0.99 * c.best) {
      c.best = c.res;
      c.best_it = c.iters;
    }
    if (!isfinite(c.res) || c.res > c.div_limit || (c.stall > 0 && c.iters - c.best_it >= c.stall)) {
      c.status = 2;
      c.active = 0;
    } else if (c.res <= c.tol) {
      c.status = 0;
      c.active = 0;
    } else if (c.bd || c.omega == 0.0 || r0r == 0.0) {
      c.status = 2;
      c.active = 0;
    } else if (c.iters >= c.max_iters) {
      c.status = 1;
      c.active = 0;
    } else {
      c.beta = (r0r / c.rho) * (c.alpha / c.omega);
      c.rho = r0r;
    }
  }
}

// Masked vector updates (rows of inactive systems are left alone):
//   0 p = r + beta (p - omega v)          a = p, c1 = r, c2 = v
//   1 s = r - alpha v                     a = s, c1 = r, c2 = v
//   2 x += alpha ph + omega sh; r = s - omega t    a = x, bb = r, c1 = ph, c2 = sh, c3 = s, c4 = t
//   3 restart: r0 = r, p = 0, v = 0       a = r0, bb = p, cw = v, c1 = r
// (the rows are taken kBlock at a time; rsys[b], the system of rows
// [b kBlock, (b + 1) kBlock) or -1 where they span two, lets a workgroup skip
// the blocks of converged systems without reading their rows' systems)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bupdate(int op, int n, const int *__restrict__ row_sys,
                                                    const int *__restrict__ rsys, const BSys *__restrict__ bs, T *a,
                                                    T *bb, T *cw, const T *c1, const T *c2, const T *c3,
                                                    const T *c4) {
  const int nb = (n + kBlock - 1) / kBlock;
  for (int blk = (int)blockIdx.x; blk < nb; blk += (int)gridDim.x) {
    const int sb = rsys[blk];
    if (sb >= 0 && !bs[sb].active) continue;
    const int i = blk * kBlock + (int)threadIdx.x;
    if (i >= n) continue;
    const BSys &c = bs[sb >= 0 ? sb : row_sys[i]];
    if (!c.active) continue;
    if (op == 0) {
      a[i] = c1[i] + (T)c.beta * (a[i] - (T)c.omega * c2[i]);
    } else if (op == 1) {
      a[i] = c1[i] - (T)c.alpha * c2[i];
    } else if (op == 2) {
      a[i] += (T)c.alpha * c1[i] + (T)c.omega * c2[i];
      bb[i] = c3[i] - (T)c.omega * c4[i];
    } else {
      a[i] = c1[i];
      bb[i] = 0;
      cw[i] = 0;
    }
  }
}

// Masked FP64 refinement helpers (mask = int per system):
//   0 r32 = r64 / nrm[s]      1 x += nrm[s] * d      2 dst = src (FP64 copy)     3 x = 0     4 f = 0
__global__ __launch_bounds__(kBlock) void k_bref(int op, int n, const int *__restrict__ row_sys,
                                                 const int *__restrict__ rsys, const int *__restrict__ mask,
                                                 double *__restrict__ x, const double *__restrict__ src,
                                                 float *__restrict__ f, const double *__restrict__ nrm) {
  const int nb = (n + kBlock - 1) / kBlock;
  for (int blk = (int)blockIdx.x; blk < nb; blk += (int)gridDim.x) {
    const int sb = rsys[blk];
    if (sb >= 0 && !mask[sb]) continue;
    const int i = blk * kBlock + (int)threadIdx.x;
    if (i >= n) continue;
    const int s = sb >= 0 ? sb : row_sys[i];
    if (!mask[s]) continue;
    if (op == 0) f[i] = (float)(src[i] * (nrm[s] > 0 ? 1.0 / nrm[s] : 1.0));
    else if (op == 1) x[i] += (double)f[i] * nrm[s];
    else if (op == 2) x[i] = src[i];
    else if (op == 3) x[i] = 0.0;
    else f[i] = 0.0f;
  }
}

// Sets the per-system state before a batched solve: active from mask, tolerance,
// iteration cap and divergence limit.
__global__ void k_binit(int S, BSys *bs, const int *__restrict__ mask, double tol, int max_iters, double div_limit,
                        int stall, const int *__restrict__ caps) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  BSys &c = bs[s];
  c.active = mask[s] ? 1 : 0;
  c.status = mask[s] ? 1 : 0;
  c.tol = tol;
  c.max_iters = caps ? min(caps[s], max_iters) : max_iters;
  c.div_limit = div_limit;
  c.stall = stall;
  c.iters = 0;
  c.bd = 0;
}

// Triangular sweep of segment-local level schedules, one workgroup per
// segment (one system of a batch): the workgroup walks its segment's levels in
// order, so all communication stays inside one CU.
//
// Everything the sweep touches is stored in level order -- position q of the
// schedule: its entries lrp[q]..lrp[q+1] (each naming its column by the
// column's position cq), right-hand side rhs[q], 1/diagonal dq[q], output
// zq[q] -- so one CU streams contiguous memory level after level; the
// permutations to and from natural order run as separate full-GPU passes.
// The outputs of the most recent positions are also kept in an LDS window of
// W slots (slot = position mod W): a dependency at position qc is read from
// LDS when qc >= q1 - W (q1 = end of the current level, so no write of this
// level can hit its slot) and from zq[qc] otherwise.
//
// Between levels only LDS has to be in order, so the levels are separated by
// an LDS-only barrier (wait for this wave's LDS operations, then s_barrier);
// the next level's prefetched row data and this level's stores keep flying.
// A full barrier, which also drains global memory operations, is taken
// whenever the next level could read from zq a position written after the
// last full barrier (q_safe tracks what a full barrier has made visible), and
// when a new chunk of level pointers is staged in LDS.
//   lower: zq[q] = rhs[q] - sum L[k] x(cq[k])
//   upper: zq[q] = dq[q] (rhs[q] - sum U[k] x(cq[k]))
constexpr int kLvlChunk = 256;
__device__ inline void lds_barrier() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt and expcnt left at their maximum (no wait)
  __builtin_amdgcn_s_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

template <int NT, int LPR, int R, int J, typename TV, typename T, bool UPPER, bool WIN = false>
__global__ __launch_bounds__(NT) void k_wgsweep(const int *__restrict__ seg_lvl, const int *__restrict__ lvl_ptr,
                                                const int *__restrict__ lrp, const int *__restrict__ cq,
                                                const TV *__restrict__ v, const T *__restrict__ rhs, T *zq,
                                                const TV *__restrict__ dq, const BSys *__restrict__ bs,
                                                const int *__restrict__ seg_sys, int W) {
  // WIN: every column the sweep reads lies within W positions of the end of
  // the level that reads it (the host checks the whole factor), so x comes
  // from the LDS window only and levels are separated by LDS-only barriers.
  extern __shared__ unsigned char lds_raw[];
  T *win = reinterpret_cast<T *>(lds_raw);
  __shared__ int lp[kLvlChunk + 1];
  const int s = blockIdx.x;
  if (bs && !bs[seg_sys ? seg_sys[s] : s].active) return;
  constexpr int G = NT / LPR;
  const int Wm = W - 1;
  const int l0 = seg_lvl[s], l1 = seg_lvl[s + 1];
  if (l0 >= l1) return;
  const int g = threadIdx.x / LPR, lane = threadIdx.x & (LPR - 1);
  // Software pipeline over levels: the row data (entry range, right-hand side,
  // 1/diagonal) of level l + 2 and the entries of level l + 1 are loaded while
  // level l is computed; each lane group holds R rows, each lane J entries of a row.
  struct Meta { int kb[R], ke[R]; T rh[R], d[R]; };
  struct Ent { int c[R][J]; TV v[R][J]; };
  auto meta = [&](Meta &M, int q0, int q1) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int q = q0 + g + r * G;
      const bool ok = q < q1;
      M.kb[r] = ok ? lrp[q] : 0;
      M.ke[r] = ok ? lrp[q + 1] : 0;
      M.rh[r] = ok ? rhs[q] : (T)0;
      M.d[r] = (UPPER && ok) ? (T)dq[q] : (T)1;
    }
  };
  auto ents = [&](Ent &E, const Meta &M) {
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const int k = M.kb[r] + lane + j * LPR;
        const bool kk = k < M.ke[r];
        E.c[r][j] = kk ? cq[k] : 0;
        E.v[r][j] = kk ? v[k] : (TV)0;
      }
  };
  // level pointers of the segment, staged in LDS kLvlChunk levels at a time;
  // the chunk always holds levels l .. l + 3 while level l is swept (so they
  // are read from LDS: a pointer that may be global or LDS would compile to
  // flat loads, whose waits drain the prefetched rows of the next levels)
  int lc = l0;  // first level of the staged chunk
  for (int t = threadIdx.x; t <= kLvlChunk && lc + t <= l1; t += NT) lp[t] = lvl_ptr[lc + t];
  __syncthreads();
  auto qp = [&](int l) { return lp[l - lc]; };
  // x of position qc: the window for positions >= lo, zq otherwise.  zq is
  // read through a buffer resource over the segment's positions (the host
  // keeps a segment under 2 GiB): an LDS-or-global choice of address would
  // compile to a flat load, whose wait drains the LDS and the global counter
  // together, prefetched rows included.
  const int qs0 = lp[0];
  const __amdgpu_buffer_rsrc_t rz = buf_rsrc(zq + qs0, (int64_t)(lvl_ptr[l1] - qs0) * (int64_t)sizeof(T));
  auto xval = [&](int qc, int lo) -> T {
    if constexpr (WIN) return win[qc & Wm];
    else return qc >= lo ? win[qc & Wm] : buf_load<T>(rz, (qc - qs0) * (int)sizeof(T));
  };
  int q_safe = lp[0];  // positions below q_safe are visible in zq to every wave
  // level l with the row data of l (cur) and l + 1 (nxt) and the entries of l
  // (ecur) in registers; loads ahead (l + 2) and enxt (l + 1).  The register
  // sets rotate by unrolling (a copy of a register whose load is in flight
  // would wait for it).  Returns whether level l + 1 exists.
  auto level = [&](int l, const Meta &cur, const Meta &nxt, Meta &ahead, const Ent &ecur, Ent &enxt) -> bool {
    const int q0 = qp(l), q1 = qp(l + 1);
    const bool more = l + 1 < l1;
    if (l + 2 < l1) meta(ahead, qp(l + 2), qp(l + 3));
    if (more) ents(enxt, nxt);
    const int lo = q1 - W;  // oldest position the window may serve during this level
    for (int base = q0 + g; base < q1; base += G * R) {
      Meta m = cur;
      Ent e = ecur;
      if (base != q0 + g) {
        // later chunks of a large level: load them now
        meta(m, base - g, q1);
        ents(e, m);
      }
      T acc[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        acc[r] = 0;
#pragma unroll
        for (int j = 0; j < J; ++j) {
          const int k = m.kb[r] + lane + j * LPR;
          if (k < m.ke[r]) acc[r] += (T)e.v[r][j] * xval(e.c[r][j], lo);
        }
        // entries beyond the prefetched J per lane
        for (int k = m.kb[r] + lane + J * LPR; k < m.ke[r]; k += LPR) acc[r] += (T)v[k] * xval(cq[k], lo);
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        T a = acc[r];
        a = group_sum<LPR>(a);
        const int q = base + r * G;
        if (lane == 0 && q < q1) {
          const T out = UPPER ? m.d[r] * (m.rh[r] - a) : m.rh[r] - a;
          zq[q] = out;
          if (WIN || q >= lo) win[q & Wm] = out;
        }
      }
    }
    if (!more) return false;
    const int nq1 = qp(l + 2);
    if (l + 4 - lc > kLvlChunk) {
      // stage the next chunk of level pointers (levels l + 1 ..)
      __syncthreads();
      lc = l + 1;
      for (int t = threadIdx.x; t <= kLvlChunk && lc + t <= l1; t += NT) lp[t] = lvl_ptr[lc + t];
      __syncthreads();
      q_safe = q1;
    } else if (!WIN && nq1 - W > q_safe) {
      // the next level may read positions >= q_safe from zq
      __syncthreads();
      q_safe = q1;
    } else {
      lds_barrier();
    }
    return true;
  };
  Meta m0, m1, m2;
  Ent e0, e1;
  meta(m0, qp(l0), qp(l0 + 1));
  ents(e0, m0);
  if (l0 + 1 < l1) meta(m1, qp(l0 + 1), qp(l0 + 2));
  for (int l = l0;; l += 6) {
    if (!level(l, m0, m1, m2, e0, e1)) break;
    if (!level(l + 1, m1, m2, m0, e1, e0)) break;
    if (!level(l + 2, m2, m0, m1, e0, e1)) break;
    if (!level(l + 3, m0, m1, m2, e1, e0)) break;
    if (!level(l + 4, m1, m2, m0, e0, e1)) break;
    if (!level(l + 5, m2, m0, m1, e1, e0)) break;
  }
}

// Segment sweep for narrow level schedules (the energy slabs of 1D devices are
// chains: a level holds the few rows of one cell), one wavefront per segment.
// At set-up each segment's levels are cut into chunks of at most kChainRows
// rows and kChainEnts entries; a chunk's level pointers, rows (entry range,
// right-hand side, 1/diagonal) and entries are copied into LDS while the
// previous chunk is swept, so the walk from level to level reads LDS only: the
// row data, the entries, and x from the window of the last W outputs (every
// column lies within W positions of the end of the level that reads it).  One
// global round trip per chunk instead of one per level.  chunk[8 c ...] =
// {first level, end level, first position, end position, first entry, end
// entry}; chunks of segment s are [cptr[s], cptr[s + 1]).
//   lower: zq[q] = rhs[q] - sum L[k] x(cq[k])
//   upper: zq[q] = dq[q] (rhs[q] - sum U[k] x(cq[k]))
constexpr int kChainRows = 32;
constexpr int kChainEnts = 256;
template <int LPR, typename TV, typename T, bool UPPER>
__global__ __launch_bounds__(64) void k_chainsweep(const int *__restrict__ cptr, const int *__restrict__ chunk,
                                                   const int *__restrict__ lvl_ptr, const int *__restrict__ lrp,
                                                   const int *__restrict__ cq, const TV *__restrict__ v,
                                                   const T *__restrict__ rhs, T *zq, const TV *__restrict__ dq,
                                                   const BSys *__restrict__ bs, const int *__restrict__ seg_sys,
                                                   int W) {
  constexpr int NE = kChainEnts / 64, G = 64 / LPR;
  struct Buf {
    int lv[kChainRows + 1], kr[kChainRows + 1];
    T rh[kChainRows], d[kChainRows];
    int c[kChainEnts];
    TV v[kChainEnts];
  };
  __shared__ Buf buf[2];
  extern __shared__ unsigned char lds_raw[];
  T *win = reinterpret_cast<T *>(lds_raw);
  const int s = blockIdx.x;
  if (bs && !bs[seg_sys ? seg_sys[s] : s].active) return;
  const int c0 = cptr[s], c1 = cptr[s + 1];
  if (c0 >= c1) return;
  const int t = (int)threadIdx.x, g = t / LPR, lane = t & (LPR - 1), Wm = W - 1;
  // a chunk in flight: this lane's level pointer, row data and entries
  struct Fly {
    int la, lb, qa, qb, ka, kb;
    int lv, kr, c[NE];
    T rh, d;
    TV v[NE];
  };
  auto issue = [&](int ci, Fly &f) {
    const int *h = chunk + 8 * ci;
    f.la = h[0]; f.lb = h[1]; f.qa = h[2]; f.qb = h[3]; f.ka = h[4]; f.kb = h[5];
    f.lv = t <= f.lb - f.la ? lvl_ptr[f.la + t] : 0;
    const bool row = t <= f.qb - f.qa;
    f.kr = row ? lrp[f.qa + t] : 0;
    f.rh = (row && t < f.qb - f.qa) ? rhs[f.qa + t] : (T)0;
    f.d = (UPPER && row && t < f.qb - f.qa) ? (T)dq[f.qa + t] : (T)1;
#pragma unroll
    for (int u = 0; u < NE; ++u) {
      const int k = f.ka + t + 64 * u;
      const bool ok = k < f.kb;
      f.c[u] = ok ? cq[k] : 0;
      f.v[u] = ok ? v[k] : (TV)0;
    }
  };
  auto store = [&](const Fly &f, Buf &B) {
    if (t <= f.lb - f.la) B.lv[t] = f.lv;
    if (t <= f.qb - f.qa) B.kr[t] = f.kr - f.ka;
    if (t < f.qb - f.qa) {
      B.rh[t] = f.rh;
      B.d[t] = f.d;
    }
#pragma unroll
    for (int u = 0; u < NE; ++u) {
      B.c[t + 64 * u] = f.c[u];
      B.v[t + 64 * u] = f.v[u];
    }
  };
  auto fence = [] {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  // sweep the levels of the chunk in B (positions from qa)
  auto walk = [&](const Buf &B, int la, int lb, int qa) {
    for (int l = la; l < lb; ++l) {
      const int q0 = B.lv[l - la], q1 = B.lv[l - la + 1];
      for (int base = q0 + g; base < q1; base += G) {
        const int i = base - qa;
        T acc = 0;
        for (int k = B.kr[i] + lane; k < B.kr[i + 1]; k += LPR) acc += (T)B.v[k] * win[B.c[k] & Wm];
        acc = group_sum<LPR>(acc);
        if (lane == 0) {
          const T out = UPPER ? B.d[i] * (B.rh[i] - acc) : B.rh[i] - acc;
          win[base & Wm] = out;
          zq[base] = out;
        }
      }
      fence();  // this level's outputs in the window before the next level reads them
    }
  };
  Fly f0, f1;
  issue(c0, f0);
  __builtin_amdgcn_s_waitcnt(0);
  store(f0, buf[0]);
  fence();
  // chunk ci is in buf[ci & 1]; the loads of chunk ci + 1 fly while it is swept
  // (static rotation of the two register sets: a copy would wait for its loads)
  for (int ci = c0; ci < c1; ci += 2) {
    if (ci + 1 < c1) issue(ci + 1, f1);
    walk(buf[(ci - c0) & 1], f0.la, f0.lb, f0.qa);
    if (ci + 1 >= c1) break;
    store(f1, buf[(ci + 1 - c0) & 1]);
    fence();
    if (ci + 2 < c1) issue(ci + 2, f0);
    walk(buf[(ci + 1 - c0) & 1], f1.la, f1.lb, f1.qa);
    if (ci + 2 >= c1) break;
    store(f0, buf[(ci + 2 - c0) & 1]);
    fence();
  }
}

// level-order permutations of the segment sweeps: dst[i] = src[map[i]] / dst[map[i]] = src[i]
// (batches: a block of kBlock positions whose system has converged is
// skipped; qsys[b] is the system of positions [b kBlock, (b + 1) kBlock), -1
// where a block spans two systems)
template <typename TI, typename TO>
__global__ __launch_bounds__(kBlock) void k_qgather(int n, const int *__restrict__ map, const TI *__restrict__ src,
                                                    TO *__restrict__ dst, const BSys *__restrict__ bs = nullptr,
                                                    const int *__restrict__ qsys = nullptr) {
  if (bs && qsys[blockIdx.x] >= 0 && !bs[qsys[blockIdx.x]].active) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (TO)src[map[i]];
}
template <typename T>
__global__ __launch_bounds__(kBlock) void k_qscatter(int n, const int *__restrict__ map, const T *__restrict__ src,
                                                     T *__restrict__ dst, const BSys *__restrict__ bs = nullptr,
                                                     const int *__restrict__ qsys = nullptr) {
  if (bs && qsys[blockIdx.x] >= 0 && !bs[qsys[blockIdx.x]].active) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[map[i]] = src[i];
}

// Host arrays of the analysis that are written in full after they are sized:
// no value initialisation, so resizing costs no serial zero fill and the pages
// are first touched by the (parallel) writers.
// Large ones ask for transparent huge pages, so that first touches fault a
// 2 MiB page at a time (where the kernel has them enabled on request).
template <typename T>
struct NoInitAlloc : std::allocator<T> {
  template <typename U>
  struct rebind {
    using other = NoInitAlloc<U>;
  };
  NoInitAlloc() = default;
  template <typename U>
  NoInitAlloc(const NoInitAlloc<U> &) {}
  T *allocate(size_t m) {
    T *p = std::allocator<T>::allocate(m);
    const size_t bytes = m * sizeof(T);
    if (bytes >= ((size_t)64 << 20)) {
      const uintptr_t a = ((uintptr_t)p + 4095) & ~(uintptr_t)4095, e = ((uintptr_t)p + bytes) & ~(uintptr_t)4095;
      if (e > a) madvise((void *)a, e - a, MADV_HUGEPAGE);
    }
    return p;
  }
  template <typename U, typename... A>
  void construct(U *p, A &&...a) {
    if constexpr (sizeof...(A) == 0) ::new ((void *)p) U;
    else ::new ((void *)p) U(std::forward<A>(a)...);
  }
};
template <typename T>
using hvec = std::vector<T, NoInitAlloc<T>>;

// ---------------------------------------------------------------- host side
inline int grid_for(int64_t items, int per_block) {
  int64_t g = (items + per_block - 1) / per_block;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, 1 << 30));
}

struct Levels {
  std::vector<int> ptr;   // level boundaries into rows
  int *d_rows = nullptr;  // rows ordered by level (ascending within a level)
  int count() const { return (int)ptr.size() - 1; }
};

}  // namespace

struct shesolve_ctx {
  int device = 0;
  shesolve_params prm{};
  hipStream_t stream = nullptr;
  // matrix
  int n = 0;
  int64_t nnz = 0;
  int *d_rp = nullptr, *d_ci = nullptr;
  double *d_v64 = nullptr;
  float *d_v32 = nullptr;
  std::vector<int> h_rp;  // pattern copy for analysis
  hvec<int> h_ci;
  // factor (pattern = A or its block-diagonal part)
  bool factor_is_A = true;
  int64_t fnnz = 0;
  int *d_frp = nullptr, *d_fci = nullptr, *d_fdg = nullptr, *d_fpos = nullptr;
  double *d_f64 = nullptr;
  int fmaxlen = 0;  // longest factor row
  std::vector<int> h_frp, h_fdg;
  hvec<int> h_fci;
  std::vector<int> blk_of_row;
  int *d_perm = nullptr;            // multicolour ordering: new -> old row
  // solve space: the matrix the Krylov iteration multiplies by -- the stored
  // matrix, or for MC_ILU0 that matrix permuted into the colour ordering (rows
  // and columns), so that neither the SpMV nor the sweeps go through perm;
  // b and x are permuted once per solve
  bool solve_perm = false;
  const int *s_rp = nullptr, *s_ci = nullptr, *s_blk = nullptr, *s_pairs = nullptr;
  int s_nblk = 0;
  double *d_sv64 = nullptr;
  float *d_sv32 = nullptr;
  int *d_sblk = nullptr;
  double *pb64 = nullptr, *px64 = nullptr;
  // reference solution of the refinement history: natural order, its
  // solve-space copy for a permuted matrix, and the one the running solve uses
  double *d_xref = nullptr, *d_xref_s = nullptr;
  const double *xref_cur = nullptr;
  std::vector<shesolve_ir_step> ir_hist;  // records of the last solve
  double *z64 = nullptr;            // permuted-space work vectors of the sweeps
  float *z32 = nullptr;
  int ncolors = 0;
  std::vector<int> h_perm, h_colour;
  // block multicolour factor: solve-space block starts, and per colour the
  // range [bmc_cptr[q], bmc_cptr[q+1]) of its blocks
  int *d_bmc_start = nullptr;
  std::vector<int> bmc_cptr;
  // grouped sweep layout of each triangle: coupling entries (CSR by row, with
  // their positions in the split factor) and in-block (slot, value) lists
  struct BmcPart {
    int *d_orp = nullptr, *d_oci = nullptr, *d_opos = nullptr, *d_iptr = nullptr, *d_ipos = nullptr;
    int *d_bK = nullptr, *d_icol = nullptr;
    unsigned short *d_icode = nullptr;
    double *d_ov64 = nullptr, *d_iv64 = nullptr;
    float *d_ov32 = nullptr, *d_iv32 = nullptr;
    int64_t no = 0, ni = 0;
  } bmcL, bmcU;
  int bmc_cap[4] = {0, 0, 0, 0};  // persistent grid of the sweep (FP32 forward, backward, FP64 forward, backward), from the occupancy at first use
  // CSR-stream row blocks of A and the tuned launch shape; in a batch no block
  // straddles two systems, and d_blk_sys holds each block's system
  int *d_ablk = nullptr;
  int nablk = 0;
  int *d_blk_sys = nullptr;
  const BSys *spmv_skip = nullptr;  // set during a batched iteration: SpMV skips converged systems
  int items = 8;   // entries per thread of a row block (row block = 256 * items entries)
  int xcd = 1;     // XCD-contiguous block mapping
  int spmv_grid = kGrid;  // workgroups of the last SpMV with fused dot products
  double tune_us[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  // the last tuned launch shapes (natural matrix: items, mapping, vector-CSR
  // lanes; solve space: items, mapping) and the row count and preconditioner
  // they were tuned for: a later matrix of that preconditioner within 10 % of
  // that size takes them over instead of timing the candidates again (the
  // successive systems of a Gummel sequence on one handle)
  int tuned_n = 0, tuned_precond = -1, tuned_items = 8, tuned_xcd = 0, tuned_lpr = 0;
  int tuned_sitems = 8, tuned_sxcd = 0;
  bool tune_reuse = false;
  int spmv_lpr = 0;  // 0: CSR-stream SpMV; else lanes per row of the vector-CSR SpMV
  // descriptors of the pipelined SpMV, for the row blocks at desc_src (blocks
  // or pairs) of the solve-space matrix
  int4 *d_desc = nullptr;
  // coded columns of the pipelined SpMV (block multicolour solve space): 16-bit
  // codes per entry, the row blocks' tables and eight-int descriptors, valid
  // for the row blocks at code_src
  unsigned short *d_code = nullptr;
  int *d_tab = nullptr, *d_cdesc = nullptr;
  int64_t ntab = 0;
  const int *code_src = nullptr;
  int code_nblk = 0;
  const int *desc_src = nullptr;
  int desc_nblk = 0;
  int desc_maxrow = 0;  // longest row of that matrix (a block holds more entries only as a single row)
  // split factor: strict lower / strict upper parts and 1/diagonal (factor space)
  int *d_Lrp = nullptr, *d_Lci = nullptr, *d_Urp = nullptr, *d_Uci = nullptr, *d_Lsrc = nullptr;
  double *d_L64 = nullptr, *d_U64 = nullptr, *d_dinv64 = nullptr;
  float *d_L32 = nullptr, *d_U32 = nullptr, *d_dinv32 = nullptr;
  int64_t nL = 0, nU = 0;
  struct LevelRun { int r0, r1, blk_off, blk_cnt; };  // contiguous level: rows [r0,r1), row blocks
  std::vector<LevelRun> runL, runU;
  int c0 = 0;  // first forward level [0, c0) skipped by the sweeps (its output is the right-hand side)
  int *d_lblk = nullptr, *d_ublk = nullptr;
  Levels lvL, lvU;
  int lpr = 8, flpr = 8;
  // segment-local level schedules (systems of a batch): segment boundaries in
  // rows, and per segment its levels (seg -> level range, level -> row range
  // of the level-ordered row list) for the one-workgroup-per-segment sweeps
  std::vector<int> seg_ptr;
  std::vector<int> seg_fill;  // ILU(k) level per system of a batch (empty = prm.fill_level)
  // segments of the factor the sweeps walk: the systems of a batch, or for the
  // block preconditioners the per-energy slabs (of every system); fseg_sys[s]
  // is the system of segment s (its convergence mask)
  std::vector<int> fseg_ptr, fseg_sys;
  int *d_fseg_sys = nullptr;
  int seg_nt = 1024;  // threads of a segment sweep's workgroup
  int nseg = 0;
  int *d_segL = nullptr, *d_lvlL = nullptr, *d_rowsL = nullptr;
  int *d_segU = nullptr, *d_lvlU = nullptr, *d_rowsU = nullptr;
  int *d_loLrp = nullptr, *d_loUrp = nullptr;  // level-ordered row pointers of the split factor
  int seg_lpr = 4;                             // lanes per row of the segment sweeps
  // farthest back a segment sweep reads (level end - column position, lower
  // and upper triangle): within the LDS window the sweep reads LDS only
  int seg_reach[2] = {0, 0};
  std::vector<int> h_lvlq[2], h_segq[2];  // level boundaries and per-segment levels of the schedules (set-up only)
  // chain sweeps: per triangle the chunks of kChainRows rows / kChainEnts
  // entries and each segment's first chunk (null when some level is wider)
  int *d_chunkL = nullptr, *d_chunkU = nullptr, *d_cptrL = nullptr, *d_cptrU = nullptr;
  int *d_qsys = nullptr;  // batches: system of each kBlock positions of the schedules (-1: two systems)
  int *d_mapLU = nullptr;                      // U position -> L position of the same row
  double *d_dq64 = nullptr, *d_tq64 = nullptr, *d_zq64 = nullptr;  // level-ordered 1/diag, rhs, output
  float *d_dq32 = nullptr, *d_tq32 = nullptr, *d_zq32 = nullptr;
  std::vector<int> h_segrowsL, h_segrowsU;     // level-ordered rows (host copy during analysis)
  const BSys *seg_mask = nullptr;  // device per-segment state; inactive segments are skipped
  int seg_levels_max = 0;
  // vectors (FP64 set and FP32 set)
  double *v64[9] = {};
  float *v32[9] = {};
  double *d_part = nullptr;
  Scalars *d_sc = nullptr;
  Scalars *h_ring = nullptr;         // pinned: the scalars of the last two enqueued iterations
  hipEvent_t iter_ev[2] = {nullptr, nullptr};
  double *d_norm = nullptr;
  int vec_n = 0;
  double last_setup_ms = 0;
  // profiling
  int prof = 0;  // 0 off, 1 every timed launch waited for, 2 SpMV launches timed live
  double ktime[5] = {0, 0, 0, 0, 0};
  int64_t kcount[5] = {0, 0, 0, 0, 0};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // pinned chunks of the staged device -> host copies (kept until destroy)
  void *stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  // live SpMV timing (profiling mode 2): event pairs recorded around SpMV
  // launches without waiting, resolved when the statistics are read
  std::vector<hipEvent_t> evpool;
  std::vector<int> evkind;  // kind per recorded pair
  int evused = 0;
};

namespace {

template <typename T>
void dfree(T *&p) {
  if (p) hipFree(p);
  p = nullptr;
}

// solve-space values (FP64 / FP32)
inline const double *sv64(const shesolve_ctx *c) { return c->solve_perm ? c->d_sv64 : c->d_v64; }
inline const float *sv32(const shesolve_ctx *c) { return c->solve_perm ? c->d_sv32 : c->d_v32; }

int pick_lpr(double avg) {
  if (avg <= 3) return 2;
  if (avg <= 6) return 4;
  if (avg <= 14) return 8;
  if (avg <= 28) return 16;
  if (avg <= 56) return 32;
  return 64;
}

// timing helper: brackets a launch with events when profiling is on
void resolve_events(shesolve_ctx *c) {
  if (!c->evused) return;
  hipEventSynchronize(c->evpool[2 * c->evused - 1]);
  for (int i = 0; i < c->evused; ++i) {
    float ms = 0;
    hipEventElapsedTime(&ms, c->evpool[2 * i], c->evpool[2 * i + 1]);
    c->ktime[c->evkind[i]] += ms * 1e3;
    c->kcount[c->evkind[i]] += 1;
  }
  c->evused = 0;
}

struct KTimer {
  shesolve_ctx *c;
  int kind;
  int slot = -1;
  KTimer(shesolve_ctx *c_, int k) : c(c_), kind(k) {
    if (c->prof == 2) {
      if (kind != 0 && kind != 4) return;
      if (2 * (c->evused + 1) > (int)c->evpool.size()) resolve_events(c);
      slot = c->evused++;
      c->evkind[slot] = kind;
      hipEventRecord(c->evpool[2 * slot], c->stream);
    } else if (c->prof) {
      hipEventRecord(c->ev0, c->stream);
    }
  }
  ~KTimer() {
    if (c->prof == 2) {
      if (slot >= 0) hipEventRecord(c->evpool[2 * slot + 1], c->stream);
    } else if (c->prof) {
      hipEventRecord(c->ev1, c->stream);
      hipEventSynchronize(c->ev1);
      float ms = 0;
      hipEventElapsedTime(&ms, c->ev0, c->ev1);
      c->ktime[kind] += ms * 1e3;
      c->kcount[kind] += 1;
    }
  }
};

// grid of a stream launch: a multiple of the XCD count, at most 4 kGrid
inline int stream_grid(int nblk, bool fixed) {
  if (fixed) return kGrid;
  int g = std::min(nblk, 4 * kGrid);
  g = (g + kXcds - 1) / kXcds * kXcds;
  return std::max(g, kXcds);
}

template <typename TV, typename T, int MODE>
void launch_stream(shesolve_ctx *c, int g, int nblk, const int *blocks, const int *rp, const int *ci, const TV *v,
                   const T *x, T *y, const T *b, const int *perm, const T *dinv, T *out2, const T *d0, const T *d1,
                   double *part, const int *pairs = nullptr, int c0 = 0) {
  // the sweeps' row blocks are built for kSweepItems entries per thread
  const int items = MODE == MODE_SPMV ? c->items : kSweepItems;
#define ST_CASE(IT, SK)                                                                                            \
  case IT:                                                                                                         \
    hipLaunchKernelGGL((k_stream<TV, T, MODE, IT, SK>), dim3(g), dim3(kStreamThreads), 0, c->stream, nblk, c->xcd, \
                       blocks, rp, ci, v, x, y, b, perm, dinv, out2, d0, d1, part, pairs, c0, c->d_blk_sys,        \
                       c->spmv_skip);                                                                              \
    break;
  if (MODE == MODE_SPMV && c->spmv_skip) {
    switch (items) { ST_CASE(4, true) ST_CASE(16, true) default: ST_CASE(8, true) }
  } else {
    switch (items) { ST_CASE(4, false) ST_CASE(16, false) default: ST_CASE(8, false) }
  }
#undef ST_CASE
}

template <typename TV, typename T, typename TA>
void launch_spmv(shesolve_ctx *c, int /*lpr*/, const TV *v, const T *x, T *y, const T *b, const T *d0, const T *d1,
                 double *part) {
  KTimer kt(c, sizeof(TV) == 8 ? 4 : 0);
  const int *rp = c->s_rp ? c->s_rp : c->d_rp, *ci = c->s_rp ? c->s_ci : c->d_ci;
  if (c->spmv_lpr) {
    // vector CSR: LPR lanes per row, grid-stride over rows; fused dots need the
    // fixed grid (one partial per workgroup)
    const int g = part ? kGrid : std::min(grid_for((int64_t)c->n * c->spmv_lpr, kBlock), 4 * kGrid);
    if (part) c->spmv_grid = g;
#define VS_CASE(LL)                                                                                              \
  case LL:                                                                                                       \
    hipLaunchKernelGGL((k_spmv<LL, TV, T, T>), dim3(g), dim3(kBlock), 0, c->stream, c->n, rp, ci, v, x, y, b, d0, d1, \
                       part);                                                                                    \
    break;
    switch (c->spmv_lpr) { VS_CASE(4) VS_CASE(16) default: VS_CASE(8) }
#undef VS_CASE
    return;
  }
  const int *blk = c->s_rp ? c->s_blk : c->d_ablk;
  const int nblk = c->s_rp ? c->s_nblk : c->nablk;
  // the fused dots leave one partial per workgroup of the persistent grid
  // (a grid of up to 4 kGrid measured no faster)
  const int g = stream_grid(nblk, part != nullptr);
  if (part) c->spmv_grid = g;
  const int *pairs = c->s_rp ? c->s_pairs : nullptr;
  if (!c->spmv_skip && c->d_desc && c->desc_src == (pairs ? pairs : blk) && c->desc_nblk == nblk &&
      c->desc_maxrow <= spmv_cap(c->items) &&
      ((int64_t)(c->solve_perm ? c->fnnz : c->nnz) + kVecPad) * (int64_t)sizeof(TV) < (int64_t)1 << 31) {
    const int64_t alloc = (int64_t)(c->solve_perm ? c->fnnz : c->nnz) + kVecPad;
    if (c->d_cdesc && pairs && c->code_src == pairs && c->code_nblk == nblk) {
#define SP_CASE(IT)                                                                                               \
  case IT:                                                                                                        \
    hipLaunchKernelGGL((k_spmv_pipe<TV, T, IT, true>), dim3(g), dim3(kStreamThreads), 0, c->stream, nblk, c->xcd, \
                       c->n, alloc, (const int *)c->d_cdesc, rp, ci, v, x, y, b, d0, d1, part, c->d_code, c->d_tab, \
                       c->ntab);                                                                                  \
    break;
      switch (c->items) { SP_CASE(4) SP_CASE(16) default: SP_CASE(8) }
#undef SP_CASE
      return;
    }
#define SP_CASE(IT)                                                                                               \
  case IT:                                                                                                        \
    hipLaunchKernelGGL((k_spmv_pipe<TV, T, IT>), dim3(g), dim3(kStreamThreads), 0, c->stream, nblk, c->xcd, c->n,  \
                       alloc, (const int *)c->d_desc, rp, ci, v, x, y, b, d0, d1, part);                         \
    break;
    switch (c->items) { SP_CASE(4) SP_CASE(16) default: SP_CASE(8) }
#undef SP_CASE
    return;
  }
  launch_stream<TV, T, MODE_SPMV>(c, g, nblk, blk, rp, ci, v, x, y, b,
                                  (const int *)nullptr, (const T *)nullptr, (T *)nullptr, d0, d1, part,
                                  c->s_rp ? c->s_pairs : nullptr);
}

// rows in flight per lane group of the forward segment sweep (the pipeline's
// registers must fit 128 VGPRs at 1024 threads); the backward sweep uses 2
template <typename T>
constexpr int kSweepRowsL = sizeof(T) == 8 ? 2 : 4;

// LDS window of the segment sweeps, a power of two number of entries: 32 KiB
// beside the level-pointer chunk for a 1024-thread workgroup, less for the
// smaller ones so that several fit a CU
template <typename T>
int sweep_window(int nt) {
  const int bytes = nt >= 1024 ? 32768 : (nt >= 256 ? 16384 : 4096);
  return (int)(bytes / sizeof(T));
}

template <typename TV, typename T>
void launch_precond(shesolve_ctx *c, const TV * /*unused*/, const T *in, T *out) {
  const int n = c->n;
  if (c->prm.precond == SHESOLVE_PRECOND_NONE || !c->d_frp) {
    hipMemcpyAsync(out, in, sizeof(T) * n, hipMemcpyDeviceToDevice, c->stream);
    return;
  }
  const TV *Lv, *Uv, *dinv;
  if constexpr (sizeof(TV) == 8) {
    Lv = (const TV *)c->d_L64; Uv = (const TV *)c->d_U64; dinv = (const TV *)c->d_dinv64;
  } else {
    Lv = (const TV *)c->d_L32; Uv = (const TV *)c->d_U32; dinv = (const TV *)c->d_dinv32;
  }
  const int L = c->flpr;
  const int *perm = c->solve_perm ? nullptr : c->d_perm;
  if (!c->bmc_cptr.empty() && !perm) {
    // block multicolour factor in the solve space: one launch per colour and
    // direction, forward in ascending and backward in descending colours
    const int nq = (int)c->bmc_cptr.size() - 1, nbk = c->bmc_cptr.back();
    constexpr int G = kBmcGroupOf<T>;
    // persistent grid: what the device holds at once (a multiple of the XCD
    // count), or fewer when the colour has fewer groups
    int *cap = c->bmc_cap + (sizeof(T) == 8 ? 2 : 0);
    if (!cap[0]) {
      int cus = 0;
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device);
      const void *fn[2] = {(const void *)k_bmc_group<false, TV, T>, (const void *)k_bmc_group<true, TV, T>};
      for (int d = 0; d < 2; ++d) {
        int per_cu = 0;
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn[d], G * kBmcRows, 0);
        cap[d] = std::max(kXcds, std::max(per_cu, 1) * std::max(cus, 1) / kXcds * kXcds);
      }
    }
    auto grid = [&](int b0, int b1, bool backward) {
      const int ng = (b1 - b0 + G - 1) / G;
      return std::min(cap[backward ? 1 : 0], (ng + kXcds - 1) / kXcds * kXcds);
    };
    const auto &PL = c->bmcL, &PU = c->bmcU;
    const TV *oL, *iL, *oU, *iU;
    if constexpr (sizeof(TV) == 8) {
      oL = (const TV *)PL.d_ov64; iL = (const TV *)PL.d_iv64; oU = (const TV *)PU.d_ov64; iU = (const TV *)PU.d_iv64;
    } else {
      oL = (const TV *)PL.d_ov32; iL = (const TV *)PL.d_iv32; oU = (const TV *)PU.d_ov32; iU = (const TV *)PU.d_iv32;
    }
    for (int q = 0; q < nq; ++q) {
      KTimer kt(c, 1);
      const int b0 = c->bmc_cptr[q], b1 = c->bmc_cptr[q + 1];
      hipLaunchKernelGGL((k_bmc_group<false, TV, T>), dim3(grid(b0, b1, false)), dim3(G * kBmcRows), 0, c->stream, b0, b1,
                         n, nbk, PL.no, PL.ni, c->d_bmc_start, PL.d_bK, PL.d_orp, PL.d_oci, oL, PL.d_iptr, PL.d_icode,
                         iL, dinv, in, out);
    }
    for (int q = nq - 1; q >= 0; --q) {
      KTimer kt(c, 2);
      const int b0 = c->bmc_cptr[q], b1 = c->bmc_cptr[q + 1];
      hipLaunchKernelGGL((k_bmc_group<true, TV, T>), dim3(grid(b0, b1, true)), dim3(G * kBmcRows), 0, c->stream, b0, b1,
                         n, nbk, PU.no, PU.ni, c->d_bmc_start, PU.d_bK, PU.d_orp, PU.d_oci, oU, PU.d_iptr, PU.d_icode,
                         iU, dinv, in, out);
    }
    return;
  }
  if (c->nseg > 0 && !perm) {
    KTimer kt(c, 1);
    const int W = sweep_window<T>(c->seg_nt);
    const size_t lds = sizeof(T) * (size_t)W;
    const int n = c->n, g = grid_for(n, kBlock);
    T *tq, *zq;
    const TV *dq;
    if constexpr (sizeof(T) == 8) { tq = (T *)c->d_tq64; zq = (T *)c->d_zq64; } else { tq = (T *)c->d_tq32; zq = (T *)c->d_zq32; }
    if constexpr (sizeof(TV) == 8) dq = (const TV *)c->d_dq64; else dq = (const TV *)c->d_dq32;
    // in (natural) -> forward schedule order; forward sweep; -> backward order;
    // backward sweep; -> natural order
    // a batch skips the positions of converged systems
    const BSys *qbs = c->d_qsys ? c->seg_mask : nullptr;
    hipLaunchKernelGGL((k_qgather<T, T>), dim3(g), dim3(kBlock), 0, c->stream, n, c->d_rowsL, in, tq, qbs, c->d_qsys);
    // the window-only form where every read of both triangles lies within the window
    const bool win = c->seg_reach[0] <= W && c->seg_reach[1] <= W;
    if (win && c->seg_nt == 64 && c->d_chunkL && c->d_chunkU) {
      // chain sweeps: a window of the next power of two above the reach
      int Wc = 64;
      while (Wc <= std::max(c->seg_reach[0], c->seg_reach[1])) Wc *= 2;
      const size_t ldsc = sizeof(T) * (size_t)Wc;
#define CH_CASE(LL)                                                                                                \
  case LL:                                                                                                         \
    hipLaunchKernelGGL((k_chainsweep<LL, TV, T, false>), dim3(c->nseg), dim3(64), ldsc, c->stream, c->d_cptrL,      \
                       c->d_chunkL, c->d_lvlL, c->d_loLrp, c->d_Lci, Lv, (const T *)tq, zq, (const TV *)nullptr,   \
                       c->seg_mask, c->d_fseg_sys, Wc);                                                            \
    hipLaunchKernelGGL((k_qgather<T, T>), dim3(g), dim3(kBlock), 0, c->stream, n, c->d_mapLU, (const T *)zq, tq,    \
                       qbs, c->d_qsys);                                                                            \
    hipLaunchKernelGGL((k_chainsweep<LL, TV, T, true>), dim3(c->nseg), dim3(64), ldsc, c->stream, c->d_cptrU,      \
                       c->d_chunkU, c->d_lvlU, c->d_loUrp, c->d_Uci, Uv, (const T *)tq, zq, dq, c->seg_mask,       \
                       c->d_fseg_sys, Wc);                                                                         \
    break;
      switch (c->seg_lpr) { CH_CASE(2) CH_CASE(4) default: CH_CASE(8) }
#undef CH_CASE
      hipLaunchKernelGGL((k_qscatter<T>), dim3(g), dim3(kBlock), 0, c->stream, n, c->d_rowsU, (const T *)zq, out, qbs,
                         c->d_qsys);
      return;
    }
#define WG_LAUNCH(NT, LL, WN)                                                                                      \
  hipLaunchKernelGGL((k_wgsweep<NT, LL, kSweepRowsL<T>, kSegJ, TV, T, false, WN>), dim3(c->nseg), dim3(NT), lds,    \
                     c->stream, c->d_segL, c->d_lvlL, c->d_loLrp, c->d_Lci, Lv, (const T *)tq, zq,                 \
                     (const TV *)nullptr, c->seg_mask, c->d_fseg_sys, W);                                          \
  hipLaunchKernelGGL((k_qgather<T, T>), dim3(g), dim3(kBlock), 0, c->stream, n, c->d_mapLU, (const T *)zq, tq, qbs, \
                     c->d_qsys);                                                                                   \
  hipLaunchKernelGGL((k_wgsweep<NT, LL, 2, kSegJ, TV, T, true, WN>), dim3(c->nseg), dim3(NT), lds, c->stream,      \
                     c->d_segU, c->d_lvlU, c->d_loUrp, c->d_Uci, Uv, (const T *)tq, zq, dq, c->seg_mask,           \
                     c->d_fseg_sys, W);
#define WG_CASE(NT, LL)                                                                                            \
  case LL:                                                                                                         \
    if (NT == 64 && win) {                                                                                         \
      WG_LAUNCH(64, LL, true)                                                                                      \
    } else {                                                                                                       \
      WG_LAUNCH(NT, LL, false)                                                                                     \
    }                                                                                                              \
    break;
    if (c->seg_nt >= 1024) {
      switch (c->seg_lpr) { WG_CASE(1024, 2) WG_CASE(1024, 4) default: WG_CASE(1024, 8) }
    } else if (c->seg_nt >= 256) {
      switch (c->seg_lpr) { WG_CASE(256, 2) WG_CASE(256, 4) default: WG_CASE(256, 8) }
    } else {
      switch (c->seg_lpr) { WG_CASE(64, 2) WG_CASE(64, 4) default: WG_CASE(64, 8) }
    }
#undef WG_CASE
#undef WG_LAUNCH
    hipLaunchKernelGGL((k_qscatter<T>), dim3(g), dim3(kBlock), 0, c->stream, n, c->d_rowsU, (const T *)zq, out, qbs,
                       c->d_qsys);
    return;
  }
  T *z = out;
  if (perm) {
    if constexpr (sizeof(T) == 8) z = (T *)c->z64; else z = (T *)c->z32;
  }
  // the first forward level, [0, c0), would copy the right-hand side: the
  // later launches read it from `in` directly
  const int c0 = perm ? 0 : c->c0;
  for (int l = c0 > 0 ? 1 : 0; l < c->lvL.count(); ++l) {
    KTimer kt(c, 1);
    const auto &run = c->runL[l];
    if (run.blk_cnt > 0) {
      launch_stream<TV, T, MODE_LSWEEP>(c, stream_grid(run.blk_cnt, false), run.blk_cnt, c->d_lblk + run.blk_off,
                                        c->d_Lrp, c->d_Lci, Lv, z, z, in, perm, (const T *)nullptr, (T *)nullptr,
                                        (const T *)nullptr, (const T *)nullptr, (double *)nullptr, nullptr, c0);
      continue;
    }
    int cnt = c->lvL.ptr[l + 1] - c->lvL.ptr[l];
    const int *rows = c->lvL.d_rows + c->lvL.ptr[l];
    int g = grid_for((int64_t)cnt * L, kBlock);
#define LS_CASE(LL)                                                                                          \
  case LL:                                                                                                   \
    hipLaunchKernelGGL((k_lsweep_rows<LL, TV, T>), dim3(g), dim3(kBlock), 0, c->stream, cnt, rows, c->d_Lrp, \
                       c->d_Lci, Lv, in, z, perm);                                                           \
    break;
    switch (L) { LS_CASE(2) LS_CASE(4) LS_CASE(8) LS_CASE(16) LS_CASE(32) default: LS_CASE(64) }
#undef LS_CASE
  }
  for (int l = 0; l < c->lvU.count(); ++l) {
    KTimer kt(c, 2);
    const auto &run = c->runU[l];
    if (run.blk_cnt > 0) {
      launch_stream<TV, T, MODE_USWEEP>(c, stream_grid(run.blk_cnt, false), run.blk_cnt, c->d_ublk + run.blk_off,
                                        c->d_Urp, c->d_Uci, Uv, z, z, c0 > 0 ? in : (const T *)nullptr, perm,
                                        (const T *)dinv, out, (const T *)nullptr, (const T *)nullptr,
                                        (double *)nullptr, nullptr, c0);
      continue;
    }
    int cnt = c->lvU.ptr[l + 1] - c->lvU.ptr[l];
    const int *rows = c->lvU.d_rows + c->lvU.ptr[l];
    int g = grid_for((int64_t)cnt * L, kBlock);
#define US_CASE(LL)                                                                                          \
  case LL:                                                                                                   \
    hipLaunchKernelGGL((k_usweep_rows<LL, TV, T>), dim3(g), dim3(kBlock), 0, c->stream, cnt, rows, c->d_Urp, \
                       c->d_Uci, Uv, (const T *)dinv, z, perm, out);                                         \
    break;
    switch (L) { US_CASE(2) US_CASE(4) US_CASE(8) US_CASE(16) US_CASE(32) default: US_CASE(64) }
#undef US_CASE
  }
}

int ensure_vectors(shesolve_ctx *c) {
  if (c->vec_n == c->n && c->v64[0]) return 0;
  for (auto &p : c->v64) dfree(p);
  for (auto &p : c->v32) dfree(p);
  size_t n = (size_t)std::max(c->n, 1);
  for (auto &p : c->v64) HIPCHK(hipMalloc(&p, n * sizeof(double)));
  for (auto &p : c->v32) HIPCHK(hipMalloc(&p, n * sizeof(float)));
  if (!c->d_part) HIPCHK(hipMalloc(&c->d_part, sizeof(double) * 2 * kPartStride));
  if (!c->d_sc) HIPCHK(hipMalloc(&c->d_sc, sizeof(Scalars)));
  if (!c->h_ring) HIPCHK(hipHostMalloc(&c->h_ring, 2 * sizeof(Scalars), hipHostMallocDefault));
  for (auto &e : c->iter_ev)
    if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  if (!c->d_norm) HIPCHK(hipMalloc(&c->d_norm, sizeof(double) * 4));
  c->vec_n = c->n;
  return 0;
}

void free_codes(shesolve_ctx *c) {
  dfree(c->d_code); dfree(c->d_tab); dfree(c->d_cdesc);
  c->ntab = 0;
  c->code_src = nullptr;
  c->code_nblk = 0;
}

void free_matrix(shesolve_ctx *c) {
  dfree(c->d_rp); dfree(c->d_ci); dfree(c->d_v64); dfree(c->d_v32);
  if (!c->factor_is_A) { dfree(c->d_frp); dfree(c->d_fci); }
  c->d_frp = c->d_fci = nullptr;
  dfree(c->d_fdg); dfree(c->d_fpos); dfree(c->d_f64);
  dfree(c->d_perm); dfree(c->z64); dfree(c->z32);
  dfree(c->d_sv64); dfree(c->d_sv32); dfree(c->d_sblk); dfree(c->pb64); dfree(c->px64);
  dfree(c->d_xref); dfree(c->d_xref_s);
  c->xref_cur = nullptr;
  dfree(c->d_bmc_start);
  c->bmc_cptr.clear();
  for (auto *P : {&c->bmcL, &c->bmcU}) {
    dfree(P->d_orp); dfree(P->d_oci); dfree(P->d_opos); dfree(P->d_iptr); dfree(P->d_ipos); dfree(P->d_icode);
    dfree(P->d_bK); dfree(P->d_icol);
    dfree(P->d_ov64); dfree(P->d_iv64); dfree(P->d_ov32); dfree(P->d_iv32);
    P->no = P->ni = 0;
  }
  c->solve_perm = false;
  c->s_rp = c->s_ci = c->s_blk = c->s_pairs = nullptr;
  c->s_nblk = 0;
  dfree(c->d_ablk); c->nablk = 0;
  dfree(c->d_desc); c->desc_src = nullptr; c->desc_nblk = 0;
  free_codes(c);
  dfree(c->d_blk_sys);
  c->spmv_skip = nullptr;
  dfree(c->d_Lrp); dfree(c->d_Lci); dfree(c->d_Urp); dfree(c->d_Uci); dfree(c->d_Lsrc);
  dfree(c->d_L64); dfree(c->d_U64); dfree(c->d_dinv64); dfree(c->d_L32); dfree(c->d_U32); dfree(c->d_dinv32);
  dfree(c->d_lblk); dfree(c->d_ublk);
  c->runL.clear(); c->runU.clear();
  dfree(c->lvL.d_rows); dfree(c->lvU.d_rows);
  c->lvL.ptr.clear(); c->lvU.ptr.clear();
  dfree(c->d_segL); dfree(c->d_lvlL); dfree(c->d_rowsL);
  dfree(c->d_chunkL); dfree(c->d_chunkU); dfree(c->d_cptrL); dfree(c->d_cptrU); dfree(c->d_qsys);
  dfree(c->d_segU); dfree(c->d_lvlU); dfree(c->d_rowsU);
  dfree(c->d_loLrp); dfree(c->d_loUrp); dfree(c->d_mapLU);
  dfree(c->d_dq64); dfree(c->d_tq64); dfree(c->d_zq64); dfree(c->d_dq32); dfree(c->d_tq32); dfree(c->d_zq32);
  dfree(c->d_fseg_sys);
  c->fseg_ptr.clear();
  c->fseg_sys.clear();
  c->nseg = 0;
}

// Row blocks over rows [r0, r1) of a CSR (host row pointer): consecutive rows
// holding at most `cap` entries and `rowcap` rows; a longer row forms a block
// of its own.
void make_blocks(const int *rp, int r0, int r1, int cap, std::vector<int> &out, int rowcap = kStreamRows) {
  out.push_back(r0);
  int start = r0;
  for (int r = r0; r < r1; ++r) {
    if (r > start && (rp[r + 1] - rp[start] > cap || r - start >= rowcap)) {
      out.push_back(r);
      start = r;
    }
  }
  if (out.back() != r1) out.push_back(r1);
}

// Level sets of the lower (forward) and upper (backward) triangles of the factor pattern.
// Runs f(i0, i1) over [0, n) cut into contiguous ranges, one per host thread
// (at most 16, and none shorter than min_per_thread items).
template <typename F>
void host_parallel(int64_t n, F &&f, int64_t min_per_thread = 1 << 15) {
  const int hw = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(hw, n / std::max<int64_t>(min_per_thread, 1)));
  if (nt <= 1) {
    if (n > 0) f((int64_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) th.emplace_back([&f, n, nt, t] { f(n * t / nt, n * (t + 1) / nt); });
  f((int64_t)0, n / nt);
  for (auto &x : th) x.join();
}

// SHESOLVE_VERBOSE: durations of the set-up phases on stderr
struct PhaseClock {
  bool on = getenv("SHESOLVE_VERBOSE") != nullptr;
  hipStream_t stream;
  std::chrono::steady_clock::time_point tp = std::chrono::steady_clock::now();
  explicit PhaseClock(hipStream_t s) : stream(s) {}
  void operator()(const char *what) {
    if (!on) return;
    hipStreamSynchronize(stream);
    const auto t = std::chrono::steady_clock::now();
    fprintf(stderr, "[shesolve] set-up %s: %.1f ms\n", what, std::chrono::duration<double, std::milli>(t - tp).count());
    tp = t;
  }
};

// Rows bucketed by level: ptr[l] .. ptr[l + 1] in rows, ascending row within a level.
// (on host threads over row slices: per-slice counts, then each slice places
// its rows after the earlier slices' rows of the same level)
void bucket_levels(const std::vector<int> &lev, int maxl, int n, std::vector<int> &ptr, std::vector<int> &rows) {
  const int L = maxl + 1;
  constexpr int kSlices = 16;
  const int ns = n >= (1 << 20) ? kSlices : 1;
  std::vector<std::vector<int>> cnt(ns, std::vector<int>(L, 0));
  auto slices = [&](auto fn) {
    std::vector<std::thread> th;
    for (int t = 1; t < ns; ++t) th.emplace_back([&, t] { fn(t); });
    fn(0);
    for (auto &x : th) x.join();
  };
  auto lo = [&](int t) { return (int)((int64_t)n * t / ns); };
  slices([&](int t) {
    for (int i = lo(t); i < lo(t + 1); ++i) cnt[t][lev[i]]++;
  });
  ptr.assign(L + 1, 0);
  for (int l = 0; l < L; ++l) {
    int acc = ptr[l];
    for (int t = 0; t < ns; ++t) {
      const int c = cnt[t][l];
      cnt[t][l] = acc;  // this slice's first place in level l
      acc += c;
    }
    ptr[l + 1] = acc;
  }
  rows.resize(n);
  slices([&](int t) {
    std::vector<int> &pos = cnt[t];
    for (int i = lo(t); i < lo(t + 1); ++i) rows[pos[lev[i]]++] = i;
  });
}

void compute_levels(const std::vector<int> &rp, const hvec<int> &ci, const std::vector<int> &dg, int n,
                    std::vector<int> &Lptr, std::vector<int> &Lrows, std::vector<int> &Uptr, std::vector<int> &Urows) {
  std::vector<int> lev(n);
  int maxl = 0;
  for (int i = 0; i < n; ++i) {
    int l = 0;
    for (int k = rp[i]; k < dg[i]; ++k) l = std::max(l, lev[ci[k]] + 1);
    lev[i] = l;
    maxl = std::max(maxl, l);
  }
  bucket_levels(lev, maxl, n, Lptr, Lrows);
  maxl = 0;
  for (int i = n - 1; i >= 0; --i) {
    int l = 0;
    for (int k = dg[i] + 1; k < rp[i + 1]; ++k) l = std::max(l, lev[ci[k]] + 1);
    lev[i] = l;
    maxl = std::max(maxl, l);
  }
  bucket_levels(lev, maxl, n, Uptr, Urows);
}

// The same levels when the factor is block diagonal over segments (row ranges
// sp[s] .. sp[s + 1] that no entry leaves): each segment on its own, the
// segments on host threads.  Returns false (nothing computed) when an entry
// leaves its segment.
bool compute_levels_segments(const std::vector<int> &rp, const hvec<int> &ci, const std::vector<int> &dg, int n,
                             const std::vector<int> &sp, std::vector<int> &Lptr, std::vector<int> &Lrows,
                             std::vector<int> &Uptr, std::vector<int> &Urows) {
  const int S = (int)sp.size() - 1;
  std::vector<int> lev(n);
  std::atomic<bool> ok{true};
  std::mutex mu;
  for (int upper = 0; upper < 2; ++upper) {
    int maxl = 0;
    host_parallel(S, [&](int64_t s0, int64_t s1) {
      int ml = 0;
      for (int s = (int)s0; s < (int)s1; ++s) {
        const int r0 = sp[s], r1 = sp[s + 1];
        for (int ii = r0; ii < r1; ++ii) {
          const int i = upper ? r1 - 1 - (ii - r0) : ii;
          const int k0 = upper ? dg[i] + 1 : rp[i], k1 = upper ? rp[i + 1] : dg[i];
          int l = 0;
          for (int k = k0; k < k1; ++k) {
            const int j = ci[k];
            if (j < r0 || j >= r1) ok = false;
            else l = std::max(l, lev[j] + 1);
          }
          lev[i] = l;
          ml = std::max(ml, l);
        }
      }
      std::lock_guard<std::mutex> g(mu);
      maxl = std::max(maxl, ml);
    }, 1);
    if (!ok) return false;
    if (!upper) bucket_levels(lev, maxl, n, Lptr, Lrows);
    else bucket_levels(lev, maxl, n, Uptr, Urows);
  }
  return true;
}

// The same levels for a block multicolour factor, computed on host threads:
// blocks (rows bst[b] .. bst[b + 1]) of one colour (blocks cptr[q] .. cptr[q + 1])
// share no entry, so a colour's blocks are independent once the earlier
// colours (forward) or the later ones (backward) are done.
void compute_levels_blocked(const std::vector<int> &rp, const hvec<int> &ci, const std::vector<int> &dg, int n,
                            const std::vector<int> &bst, const std::vector<int> &cptr, std::vector<int> &Lptr,
                            std::vector<int> &Lrows, std::vector<int> &Uptr, std::vector<int> &Urows) {
  std::vector<int> lev(n);
  const int nq = (int)cptr.size() - 1;
  std::mutex mu;
  for (int upper = 0; upper < 2; ++upper) {
    int maxl = 0;
    for (int qq = 0; qq < nq; ++qq) {
      const int q = upper ? nq - 1 - qq : qq;
      host_parallel(cptr[q + 1] - cptr[q], [&](int64_t b0, int64_t b1) {
        int ml = 0;
        for (int b = cptr[q] + (int)b0; b < cptr[q] + (int)b1; ++b) {
          if (!upper) {
            for (int i = bst[b]; i < bst[b + 1]; ++i) {
              int l = 0;
              for (int k = rp[i]; k < dg[i]; ++k) l = std::max(l, lev[ci[k]] + 1);
              lev[i] = l;
              ml = std::max(ml, l);
            }
          } else {
            for (int i = bst[b + 1] - 1; i >= bst[b]; --i) {
              int l = 0;
              for (int k = dg[i] + 1; k < rp[i + 1]; ++k) l = std::max(l, lev[ci[k]] + 1);
              lev[i] = l;
              ml = std::max(ml, l);
            }
          }
        }
        std::lock_guard<std::mutex> g(mu);
        maxl = std::max(maxl, ml);
      }, 256);
    }
    if (!upper) bucket_levels(lev, maxl, n, Lptr, Lrows);
    else bucket_levels(lev, maxl, n, Uptr, Urows);
  }
}

// Whether the segment sweeps (one workgroup per segment, the segments side by
// side) beat one launch per level of the whole matrix, estimated on the
// forward triangle, and the workgroup size of the segment sweeps.  A
// workgroup spends a fixed time per level of its segment (kSegLevelUs for
// 1024 threads, less for the smaller groups, whose barriers are cheaper) plus
// kSegEntryUs per entry; the segments run side by side up to the number of
// workgroups the chip holds at once.  One launch per level costs about
// kLevelLaunchUs.  Narrow levels (1D devices, per-energy slabs) take small
// workgroups; wide levels, where one CU runs out of bandwidth, the launches.
bool segment_sweeps_pay(shesolve_ctx *c, const std::vector<int> &Lptr, const std::vector<int> &Lrows) {
  const auto &sp = c->fseg_ptr;
  const int S = (int)sp.size() - 1, nl = (int)Lptr.size() - 1;
  if (S < 1 || nl < 1) return false;
  // the sweeps address a segment's FP64 outputs through 32-bit buffer offsets
  for (int sg = 0; sg < S; ++sg)
    if ((int64_t)(sp[sg + 1] - sp[sg]) * (int64_t)sizeof(double) >= ((int64_t)1 << 31)) return false;
  // per segment: its rows, the number of distinct levels holding them, and
  // its forward entries (each row counted with its diagonal); on host threads
  const int n = c->n;
  std::vector<int> lev(S, 0), rows(S, 0), level_of(n);
  std::vector<double> entries_cost(S, 0.0);
  host_parallel(nl, [&](int64_t l0, int64_t l1) {
    for (int64_t l = l0; l < l1; ++l)
      for (int q = Lptr[l]; q < Lptr[l + 1]; ++q) level_of[Lrows[q]] = (int)l;
  }, 64);
  host_parallel(S, [&](int64_t s0, int64_t s1) {
    std::vector<int> stamp(nl, -1);
    for (int sg = (int)s0; sg < (int)s1; ++sg) {
      int cnt = 0;
      for (int i = sp[sg]; i < sp[sg + 1]; ++i)
        if (stamp[level_of[i]] != sg) {
          stamp[level_of[i]] = sg;
          ++cnt;
        }
      lev[sg] = cnt;
      rows[sg] = sp[sg + 1] - sp[sg];
      int64_t ents = 0;
      for (int i = sp[sg]; i < sp[sg + 1]; ++i) ents += 1 + (c->h_fdg[i] - c->h_frp[i]);
      entries_cost[sg] = kSegEntryUs * (double)ents;
    }
  }, 1);
  // workgroup size from the mean level width (rows per level of a segment)
  int64_t tot_rows = 0, tot_lev = 0;
  for (int sg = 0; sg < S; ++sg) { tot_rows += rows[sg]; tot_lev += lev[sg]; }
  const double width = tot_lev ? (double)tot_rows / (double)tot_lev : 1.0;
  c->seg_nt = width <= 16.0 ? 64 : (width <= 128.0 ? 256 : 1024);
  const double level_us = c->seg_nt >= 1024 ? kSegLevelUs : (c->seg_nt >= 256 ? kSegLevelUs256 : kSegLevelUs64);
  const double slots = 256.0 * (c->seg_nt >= 1024 ? 2 : (c->seg_nt >= 256 ? 8 : 16));
  double mx = 0.0, sum = 0.0;
  for (int sg = 0; sg < S; ++sg) {
    const double t = level_us * lev[sg] + entries_cost[sg];
    mx = std::max(mx, t);
    sum += t;
  }
  const double seg = std::max(mx, sum / slots);
  return seg < kLevelLaunchUs * nl;
}

// Segment-local level schedules of both triangles of the factor: for each
// segment (a contiguous row range that no factor entry leaves), its rows
// ordered by level, ascending row within a level.
int segment_levels(shesolve_ctx *c) {
  const int n = c->n;
  const std::vector<int> &sp = c->fseg_ptr;
  const int S = (int)sp.size() - 1;
  const auto &rp = c->h_frp, &dg = c->h_fdg;
  const auto &ci = c->h_fci;
  std::vector<int> lev(n);
  // (segments on host threads: a check, then per triangle a pass that levels
  // each segment's rows and one that places them after the prefix sums)
  std::atomic<int64_t> bad{-1};
  host_parallel(S, [&](int64_t s0, int64_t s1) {
    for (int s = (int)s0; s < (int)s1; ++s)
      for (int i = sp[s]; i < sp[s + 1]; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k)
          if (ci[k] < sp[s] || ci[k] >= sp[s + 1]) bad = ((int64_t)i << 32) | (uint32_t)ci[k];
  }, 1);
  if (bad >= 0) {
    set_err("factor entry (%d, %d) couples two segments", (int)(bad >> 32), (int)(bad & 0xffffffff));
    return 2;
  }
  c->seg_levels_max = 0;
  for (int upper = 0; upper < 2; ++upper) {
    std::vector<int> nl(S, 0);
    host_parallel(S, [&](int64_t s0, int64_t s1) {
      for (int s = (int)s0; s < (int)s1; ++s) {
        const int r0 = sp[s], r1 = sp[s + 1];
        int maxl = -1;
        if (!upper) {
          for (int i = r0; i < r1; ++i) {
            int l = 0;
            for (int k = rp[i]; k < dg[i]; ++k) l = std::max(l, lev[ci[k]] + 1);
            lev[i] = l;
            maxl = std::max(maxl, l);
          }
        } else {
          for (int i = r1 - 1; i >= r0; --i) {
            int l = 0;
            for (int k = dg[i] + 1; k < rp[i + 1]; ++k) l = std::max(l, lev[ci[k]] + 1);
            lev[i] = l;
            maxl = std::max(maxl, l);
          }
        }
        nl[s] = maxl + 1;
      }
    }, 1);
    std::vector<int> seg(S + 1, 0), lvl, rows(n);
    for (int s = 0; s < S; ++s) {
      seg[s + 1] = seg[s] + nl[s];
      c->seg_levels_max = std::max(c->seg_levels_max, nl[s]);
    }
    lvl.assign(seg[S] + 1, 0);
    host_parallel(S, [&](int64_t s0, int64_t s1) {
      std::vector<int> cnt;
      for (int s = (int)s0; s < (int)s1; ++s) {
        const int r0 = sp[s], r1 = sp[s + 1], m = nl[s];
        cnt.assign(m + 1, 0);
        for (int i = r0; i < r1; ++i) cnt[lev[i] + 1]++;
        for (int l = 0; l < m; ++l) cnt[l + 1] += cnt[l];
        // the segment's rows occupy positions r0 .. r1 of the schedule
        for (int l = 0; l < m; ++l) lvl[seg[s] + l + 1] = r0 + cnt[l + 1];
        for (int i = r0; i < r1; ++i) rows[r0 + cnt[lev[i]]++] = i;
      }
    }, 1);
    int *&dseg = upper ? c->d_segU : c->d_segL;
    int *&dlvl = upper ? c->d_lvlU : c->d_lvlL;
    int *&drows = upper ? c->d_rowsU : c->d_rowsL;
    HIPCHK(hipMalloc(&dseg, sizeof(int) * seg.size()));
    HIPCHK(hipMalloc(&dlvl, sizeof(int) * lvl.size()));
    HIPCHK(hipMalloc(&drows, sizeof(int) * std::max<size_t>(rows.size(), 1)));
    HIPCHK(hipMemcpy(dseg, seg.data(), sizeof(int) * seg.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dlvl, lvl.data(), sizeof(int) * lvl.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(drows, rows.data(), sizeof(int) * rows.size(), hipMemcpyHostToDevice));
    (upper ? c->h_segrowsU : c->h_segrowsL).swap(rows);
    c->h_lvlq[upper].swap(lvl);
    c->h_segq[upper].swap(seg);
  }
  if (!c->seg_ptr.empty()) {
    // batch: the segment's system holds its convergence mask
    HIPCHK(hipMalloc(&c->d_fseg_sys, sizeof(int) * S));
    HIPCHK(hipMemcpy(c->d_fseg_sys, c->fseg_sys.data(), sizeof(int) * S, hipMemcpyHostToDevice));
  }
  c->nseg = S;
  return 0;
}

// Level-of-fill pattern of ILU(k) over rows [r0, r1): entries of A have level
// 0; eliminating row j from row i gives column c > j the level
// lev(i,j) + lev(j,c) + 1, kept (minimum over j) when <= k.  Rows in order, the
// lower part of a row in ascending column order (a min-heap of pending
// columns).  The rows may only reach columns in [r0, r1) (a diagonal block).
// Output row lengths in len[i - r0], columns and the A-entry of each factor
// entry (-1 for fill) appended to fci / pos.
int iluk_rows(const std::vector<int> &rp, const hvec<int> &ci, int r0, int r1, int k, std::vector<int> &len,
              std::vector<int> &fci, std::vector<int> &pos, bool drop_outside = false) {
  const int m = r1 - r0;
  std::vector<int> frp(m + 1, 0), udg(m);
  std::vector<unsigned char> flev;
  std::vector<int> lev(m, -1), src(m, -1), touched;
  std::priority_queue<int, std::vector<int>, std::greater<int>> heap;
  len.assign(m, 0);
  if (k <= 1) {
    // Levels 0 and 1 without the heap: a level-1 entry (i, c) needs lev(i, j) =
    // lev(j, c) = 0, so row i is the block's part of A's row i plus, for each j < i
    // in it, the block's part of A's row j right of j (level 1 where A has no entry)
    for (int i = 0; i < m; ++i) {
      touched.clear();
      const int gi = r0 + i;
      for (int q = rp[gi]; q < rp[gi + 1]; ++q) {
        const int cl = ci[q] - r0;
        if (cl < 0 || cl >= m) {
          if (drop_outside) continue;
          return -1;
        }
        lev[cl] = 0;
        src[cl] = q;
        touched.push_back(cl);
      }
      if (k == 1) {
        const size_t na = touched.size();
        for (size_t t = 0; t < na && touched[t] < i; ++t) {
          const int j = touched[t], gj = r0 + j;
          for (int q = rp[gj]; q < rp[gj + 1]; ++q) {
            const int col = ci[q] - r0;
            if (col <= j || col >= m || lev[col] >= 0) continue;
            lev[col] = 1;
            touched.push_back(col);
          }
        }
        std::sort(touched.begin(), touched.end());
      }
      int dg = -1;
      for (int col : touched) {
        if (col == i) dg = (int)fci.size();
        fci.push_back(col + r0);
        pos.push_back(src[col]);
        lev[col] = -1;
        src[col] = -1;
      }
      if (dg < 0) return -2 - gi;
      len[i] = (int)touched.size();
    }
    return 0;
  }
  for (int i = 0; i < m; ++i) {
    touched.clear();
    const int gi = r0 + i;
    for (int q = rp[gi]; q < rp[gi + 1]; ++q) {
      const int cl = ci[q] - r0;
      if (cl < 0 || cl >= m) {
        if (drop_outside) continue;  // block ILU(k): only the diagonal block is factorised
        return -1;
      }
      lev[cl] = 0;
      src[cl] = q;
      touched.push_back(cl);
      if (cl < i) heap.push(cl);
    }
    while (!heap.empty()) {
      const int j = heap.top();
      heap.pop();
      const int lj = lev[j];
      for (int q = udg[j] + 1; q < frp[j + 1]; ++q) {
        const int col = fci[q] - r0, nl = lj + flev[q] + 1;
        if (nl > k) continue;
        if (lev[col] < 0) {
          lev[col] = nl;
          touched.push_back(col);
          if (col < i) heap.push(col);
        } else if (nl < lev[col]) {
          lev[col] = nl;
        }
      }
    }
    std::sort(touched.begin(), touched.end());
    udg[i] = -1;
    for (int col : touched) {
      if (col == i) udg[i] = (int)fci.size();
      fci.push_back(col + r0);
      flev.push_back((unsigned char)lev[col]);
      pos.push_back(src[col]);
      lev[col] = -1;
      src[col] = -1;
    }
    if (udg[i] < 0) return -2 - gi;
    frp[i + 1] = (int)fci.size();
    len[i] = frp[i + 1] - frp[i];
  }
  return 0;
}

// ILU(k) pattern of the whole matrix: one diagonal block per factor segment
// (systems of a batch, or per-energy slabs for BLOCK_ILUK, built in parallel,
// each with its own level), entries outside a segment dropped when
// drop_outside.  pos maps factor entries to A's entries, -1 for fill.
int iluk_pattern(shesolve_ctx *c, hvec<int> &pos, bool drop_outside) {
  const int n = c->n;
  const std::vector<int> &seg = c->fseg_ptr;
  const int S = (int)seg.size() - 1;
  std::vector<int> kseg(S, c->prm.fill_level);
  const int nsys = c->seg_ptr.empty() ? 1 : (int)c->seg_ptr.size() - 1;
  if ((int)c->seg_fill.size() == nsys)
    for (int q = 0; q < S; ++q) kseg[q] = c->seg_fill[c->fseg_sys[q]];
  if (*std::max_element(kseg.begin(), kseg.end()) <= 1) {
    // ILU(0) and ILU(1): two passes over the segments on host threads (row
    // sizes, then the rows written in place), no per-segment buffers.  A
    // level-1 entry (i, c) needs lev(i, j) = lev(j, c) = 0, so row i is the
    // segment's part of A's row i plus, for each j < i in it, the segment's part
    // of A's row j right of j (level 1 where A has no entry).
    PhaseClock phase(c->stream);
    int mmax = 0;
    for (int q = 0; q < S; ++q) mmax = std::max(mmax, seg[q + 1] - seg[q]);
    std::atomic<int64_t> out_seg{-1}, no_diag{-1};
    c->h_frp.assign(n + 1, 0);
    auto pass = [&](bool fill) {
      host_parallel(S, [&](int64_t s0, int64_t s1) {
        std::vector<int> lev(mmax, -1), src(mmax, -1), touched;
        for (int q = (int)s0; q < (int)s1; ++q) {
          const int r0 = seg[q], m = seg[q + 1] - r0, k = kseg[q];
          for (int gi = r0; gi < r0 + m; ++gi) {
            const int i = gi - r0;
            touched.clear();
            bool outside = false;
            for (int e = c->h_rp[gi]; e < c->h_rp[gi + 1]; ++e) {
              const int cl = c->h_ci[e] - r0;
              if (cl < 0 || cl >= m) {
                outside = true;
                continue;
              }
              lev[cl] = 0;
              src[cl] = e;
              touched.push_back(cl);
            }
            if (outside && !drop_outside) out_seg = q;
            if (k == 1) {
              const size_t na = touched.size();
              for (size_t t = 0; t < na && touched[t] < i; ++t) {
                const int j = touched[t], gj = r0 + j;
                for (int e = c->h_rp[gj]; e < c->h_rp[gj + 1]; ++e) {
                  const int col = c->h_ci[e] - r0;
                  if (col <= j || col >= m || lev[col] >= 0) continue;
                  lev[col] = 1;
                  touched.push_back(col);
                }
              }
              std::sort(touched.begin(), touched.end());
            }
            bool dg = false;
            int w = fill ? c->h_frp[gi] : 0;
            for (int col : touched) {
              dg = dg || col == i;
              if (fill) {
                c->h_fci[w] = col + r0;
                pos[w] = src[col];
                ++w;
              }
              lev[col] = -1;
              src[col] = -1;
            }
            if (!dg) no_diag = gi;
            if (!fill) c->h_frp[gi + 1] = (int)touched.size();
          }
        }
      }, 1);
    };
    pass(false);
    if (out_seg >= 0) { set_err("ILU(k): an entry of segment %d leaves its diagonal block", (int)out_seg.load()); return 2; }
    if (no_diag >= 0) { set_err("ILU(k): row %d has no diagonal entry", (int)no_diag.load()); return 2; }
    int64_t total = 0;
    for (int i = 0; i < n; ++i) {
      total += c->h_frp[i + 1];
      if (total > 0x7fffffff) { set_err("ILU(k) factor too large for 32-bit indices"); return 2; }
      c->h_frp[i + 1] = (int)total;
    }
    c->h_fci.resize(total);
    pos.resize(total);
    pass(true);
    phase("    level 0/1 patterns");
    return 0;
  }
  std::vector<std::vector<int>> lens(S), fcis(S), poss(S);
  std::vector<int> err(S, 0);
  std::atomic<int> next{0};
  auto work = [&]() {
    for (;;) {
      const int s = next.fetch_add(1);
      if (s >= S) break;
      err[s] = iluk_rows(c->h_rp, c->h_ci, seg[s], seg[s + 1], kseg[s], lens[s], fcis[s], poss[s], drop_outside);
    }
  };
  const int nt = std::max(1, std::min<int>(S, (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()))));
  PhaseClock phase(c->stream);
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) th.emplace_back(work);
  work();
  for (auto &t : th) t.join();
  phase("    segment patterns");
  int64_t total = 0;
  for (int s = 0; s < S; ++s) {
    if (err[s] == -1) { set_err("ILU(k): an entry of segment %d leaves its diagonal block", s); return 2; }
    if (err[s] < -1) { set_err("ILU(k): row %d has no diagonal entry", -(err[s] + 2)); return 2; }
    total += (int64_t)fcis[s].size();
  }
  if (total > 0x7fffffff) { set_err("ILU(k) factor too large for 32-bit indices"); return 2; }
  c->h_frp.assign(n + 1, 0);
  c->h_fci.resize(total);
  pos.resize(total);
  // each segment's rows go to their place in parallel
  std::vector<int64_t> off(S + 1, 0);
  for (int s = 0; s < S; ++s) off[s + 1] = off[s] + (int64_t)fcis[s].size();
  host_parallel(S, [&](int64_t s0, int64_t s1) {
    for (int s = (int)s0; s < (int)s1; ++s) {
      int64_t w = off[s];
      for (int i = 0; i < seg[s + 1] - seg[s]; ++i) c->h_frp[seg[s] + i + 1] = (int)(w += lens[s][i]);
      std::copy(fcis[s].begin(), fcis[s].end(), c->h_fci.begin() + off[s]);
      std::copy(poss[s].begin(), poss[s].end(), pos.begin() + off[s]);
      std::vector<int>().swap(fcis[s]);
      std::vector<int>().swap(poss[s]);
    }
  }, 1);
  phase("    merge");
  return 0;
}

// Grouped sweep layout of one triangle of a block multicolour factor (split
// pattern rp/ci in the solve space, block starts bst): per row its entries
// outside the row's block (CSR, and per block the CSR position of its first
// row), per block the tile slots of its own entries (code = block index
// within an aligned run of kBmcGroupMax blocks of the colour, times kBmcTile,
// plus the slot), each with its position in the triangle's value array and
// its column, for the value gathers after a refactor.
int build_bmc_part(shesolve_ctx *c, const std::vector<int> &bst, bool upper, const hvec<int> &rp,
                   const hvec<int> &ci, shesolve_ctx::BmcPart &P) {
  const int n = c->n, nb = (int)bst.size() - 1;
  if (nb < 1 || bst[nb] != n) {
    set_err("block multicolour layout: blocks do not cover the rows");
    return -1;
  }
  PhaseClock phase(c->stream);
  std::vector<int> orp(n + 1, 0), iptr(nb + 1, 0), bK(nb + 1, 0);
  hvec<int> oci, opos, ipos, icol;
  hvec<unsigned short> icode;
  // first block of each block's colour (the colours are consecutive block ranges)
  std::vector<int> bq0(nb, 0);
  for (int q = 0; q + 1 < (int)c->bmc_cptr.size(); ++q)
    for (int blk = c->bmc_cptr[q]; blk < c->bmc_cptr[q + 1]; ++blk) bq0[blk] = c->bmc_cptr[q];
  // two passes over the blocks in parallel: count each row's coupling entries
  // and each block's own entries, then fill at the prefix sums
  std::atomic<int64_t> bad{-1};
  host_parallel(nb, [&](int64_t b0, int64_t b1) {
    for (int blk = (int)b0; blk < (int)b1; ++blk) {
      const int r0 = bst[blk], r1 = bst[blk + 1];
      int in = 0;
      for (int r = r0; r < r1; ++r) {
        int out = 0;
        for (int k = rp[r]; k < rp[r + 1]; ++k) {
          const int j = ci[k];
          if (j >= r0 && j < r1) {
            if (upper ? j <= r : j >= r) bad = ((int64_t)r << 32) | (uint32_t)j;
            ++in;
          } else {
            ++out;
          }
        }
        orp[r + 1] = out;
      }
      iptr[blk + 1] = in;
    }
  }, 1024);
  if (bad >= 0) {
    set_err("block multicolour layout: entry (%d, %d) is not in its triangle", (int)(bad >> 32), (int)(bad & 0xffffffff));
    return -1;
  }
  phase("    layout counts");
  for (int r = 0; r < n; ++r) orp[r + 1] += orp[r];
  for (int blk = 0; blk < nb; ++blk) iptr[blk + 1] += iptr[blk];
  for (int blk = 0; blk <= nb; ++blk) bK[blk] = orp[bst[blk]];
  oci.resize(orp[n]);
  opos.resize(orp[n]);
  icode.resize(iptr[nb]);
  ipos.resize(iptr[nb]);
  icol.resize(iptr[nb]);
  host_parallel(nb, [&](int64_t b0, int64_t b1) {
    for (int blk = (int)b0; blk < (int)b1; ++blk) {
      const int sub = (blk - bq0[blk]) % kBmcGroupMax, r0 = bst[blk], r1 = bst[blk + 1];
      int wi = iptr[blk];
      for (int r = r0; r < r1; ++r) {
        int wo = orp[r];
        for (int k = rp[r]; k < rp[r + 1]; ++k) {
          const int j = ci[k];
          if (j >= r0 && j < r1) {
            icode[wi] = (unsigned short)(sub * kBmcTile + bmc_slot(upper, j - r0, r - r0));
            ipos[wi] = k;
            icol[wi] = j;
            ++wi;
          } else {
            oci[wo] = j;
            opos[wo] = k;
            ++wo;
          }
        }
      }
    }
  }, 1024);
  phase("    layout fill");
  P.no = (int64_t)oci.size();
  P.ni = (int64_t)ipos.size();
  // the sweep addresses these arrays through 32-bit buffer offsets
  if (std::max<int64_t>(P.no, P.ni) * (int64_t)sizeof(double) >= ((int64_t)1 << 31) ||
      (int64_t)(n + 1) * (int64_t)sizeof(double) >= ((int64_t)1 << 31)) {
    set_err("block multicolour layout: %lld coupling and %lld in-block entries exceed the sweep's 2 GiB arrays",
            (long long)P.no, (long long)P.ni);
    return -1;
  }
  const size_t no = std::max<size_t>(oci.size(), 1), ni = std::max<size_t>(ipos.size(), 1);
  HIPCHK(hipMalloc(&P.d_orp, sizeof(int) * (n + 1)));
  HIPCHK(hipMalloc(&P.d_bK, sizeof(int) * (nb + 1)));
  HIPCHK(hipMalloc(&P.d_oci, sizeof(int) * (no + kVecPad)));
  HIPCHK(hipMemset(P.d_oci, 0, sizeof(int) * (no + kVecPad)));
  HIPCHK(hipMalloc(&P.d_opos, sizeof(int) * no));
  HIPCHK(hipMalloc(&P.d_ov64, sizeof(double) * no));
  HIPCHK(hipMalloc(&P.d_ov32, sizeof(float) * (no + kVecPad)));
  HIPCHK(hipMemset(P.d_ov32, 0, sizeof(float) * (no + kVecPad)));
  HIPCHK(hipMalloc(&P.d_iptr, sizeof(int) * (nb + 1)));
  HIPCHK(hipMalloc(&P.d_ipos, sizeof(int) * ni));
  HIPCHK(hipMalloc(&P.d_icol, sizeof(int) * ni));
  HIPCHK(hipMalloc(&P.d_icode, sizeof(unsigned short) * ni));
  HIPCHK(hipMalloc(&P.d_iv64, sizeof(double) * ni));
  HIPCHK(hipMalloc(&P.d_iv32, sizeof(float) * ni));
  HIPCHK(hipMemcpy(P.d_orp, orp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(P.d_bK, bK.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice));
  if (P.no) {
    HIPCHK(hipMemcpy(P.d_oci, oci.data(), sizeof(int) * P.no, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(P.d_opos, opos.data(), sizeof(int) * P.no, hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemcpy(P.d_iptr, iptr.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice));
  if (P.ni) {
    HIPCHK(hipMemcpy(P.d_ipos, ipos.data(), sizeof(int) * P.ni, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(P.d_icol, icol.data(), sizeof(int) * P.ni, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(P.d_icode, icode.data(), sizeof(unsigned short) * P.ni, hipMemcpyHostToDevice));
  }
  return 0;
}

// Values of the grouped sweep layout from the split factor's triangle values:
// coupling entries as they are, in-block entries negated and, for the backward
// triangle (dinv given), divided by their column's diagonal.
void bmc_part_values(shesolve_ctx *c, const shesolve_ctx::BmcPart &P, const double *src, const double *dinv) {
  if (P.no) {
    hipLaunchKernelGGL(k_gather, dim3(grid_for(P.no, kBlock * 4)), dim3(kBlock), 0, c->stream, P.no, P.d_opos, src,
                       P.d_ov64);
    hipLaunchKernelGGL((k_cast<double, float>), dim3(grid_for(P.no, kBlock * 4)), dim3(kBlock), 0, c->stream, P.no,
                       (const double *)P.d_ov64, P.d_ov32);
  }
  if (P.ni)
    hipLaunchKernelGGL(k_bmc_inblock, dim3(grid_for(P.ni, kBlock)), dim3(kBlock), 0, c->stream, P.ni, P.d_ipos,
                       P.d_icol, src, dinv, P.d_iv64, P.d_iv32);
}

int analyse_and_factor(shesolve_ctx *c) {
  const int n = c->n;
  if (n <= 0) return 0;
  PhaseClock phase(c->stream);
  // factor pattern (in the factor's own index space) and the gather map from A
  if (c->prm.precond == SHESOLVE_PRECOND_NONE) return 0;
  hvec<int> pos;
  std::vector<int> perm;
  // Segments of the factor: the systems of a batch (or the whole matrix), cut
  // further into the blocks of the block preconditioners when every block is
  // one contiguous row range inside one system (the per-energy slabs): the
  // block factor never couples two blocks, so each is swept on its own.
  {
    std::vector<int> sys = c->seg_ptr.empty() ? std::vector<int>{0, n} : c->seg_ptr;
    c->fseg_ptr = sys;
    const bool block = (c->prm.precond == SHESOLVE_PRECOND_BLOCK_ILU0 ||
                        c->prm.precond == SHESOLVE_PRECOND_BLOCK_ILUK) && !c->blk_of_row.empty();
    if (block) {
      std::vector<int> runs(1, 0);
      std::vector<char> seen;
      bool ok = true;
      for (int i = 1; i <= n && ok; ++i) {
        if (i < n && c->blk_of_row[i] == c->blk_of_row[i - 1]) continue;
        const int b = c->blk_of_row[i - 1];
        if (b >= (int)seen.size()) seen.resize(b + 1, 0);
        if (seen[b]) ok = false;  // a block in two pieces
        seen[b] = 1;
        runs.push_back(i);
      }
      // every system boundary must be a block boundary
      for (int b : sys) ok = ok && std::binary_search(runs.begin(), runs.end(), b);
      if (ok) c->fseg_ptr = runs;
      else if (c->prm.precond == SHESOLVE_PRECOND_BLOCK_ILUK) {
        set_err("BLOCK_ILUK needs block ranges that are contiguous row ranges inside one system");
        return 2;
      }
    }
    const int S = (int)c->fseg_ptr.size() - 1;
    c->fseg_sys.assign(S, 0);
    for (int q = 0; q < S; ++q)
      c->fseg_sys[q] = (int)(std::upper_bound(sys.begin(), sys.end(), c->fseg_ptr[q]) - sys.begin()) - 1;
  }
  if (c->prm.precond == SHESOLVE_PRECOND_BLOCK_ILUK) {
    if (int e = iluk_pattern(c, pos, true)) return e;
  } else if (c->prm.precond == SHESOLVE_PRECOND_BLOCK_ILU0 && !c->blk_of_row.empty()) {
    c->h_frp.assign(n + 1, 0);
    c->h_fci.clear();
    c->h_fci.reserve(c->h_ci.size());
    for (int i = 0; i < n; ++i) {
      for (int k = c->h_rp[i]; k < c->h_rp[i + 1]; ++k)
        if (c->blk_of_row[c->h_ci[k]] == c->blk_of_row[i]) {
          c->h_fci.push_back(c->h_ci[k]);
          pos.push_back(k);
        }
      c->h_frp[i + 1] = (int)c->h_fci.size();
    }
  } else if (c->prm.precond == SHESOLVE_PRECOND_MC_ILU0 || c->prm.precond == SHESOLVE_PRECOND_BMC_ILU0) {
    // greedy colouring of the rows in natural order on the symmetrised
    // pattern: rows of one colour are mutually independent in both
    // triangles, so each colour is one level of both sweeps; rows sorted by
    // colour (stable) give the multicolour ordering new -> old.  The block
    // form colours blocks of kBmcRows consecutive rows instead (two blocks
    // conflict when an entry links them); a block keeps its rows together and
    // in natural order, and the blocks of one colour are independent.
    std::vector<int> colour(n, -1);
    std::vector<int> mark(1024, -1);
    int ncol = 0;
    if (c->prm.precond == SHESOLVE_PRECOND_MC_ILU0) {
      std::vector<int> trp(n + 1, 0), tci(c->h_ci.size());
      for (int i = 0; i < n; ++i)
        for (int k = c->h_rp[i]; k < c->h_rp[i + 1]; ++k) trp[c->h_ci[k] + 1]++;
      for (int i = 0; i < n; ++i) trp[i + 1] += trp[i];
      {
        std::vector<int> fillp(trp.begin(), trp.end() - 1);
        for (int i = 0; i < n; ++i)
          for (int k = c->h_rp[i]; k < c->h_rp[i + 1]; ++k) tci[fillp[c->h_ci[k]]++] = i;
      }
      for (int i = 0; i < n; ++i) {
        for (int k = c->h_rp[i]; k < c->h_rp[i + 1]; ++k) {
          const int cj = colour[c->h_ci[k]];
          if (cj >= 0) mark[cj] = i;
        }
        for (int k = trp[i]; k < trp[i + 1]; ++k) {
          const int cj = colour[tci[k]];
          if (cj >= 0) mark[cj] = i;
        }
        int q = 0;
        while (q < 1023 && mark[q] == i) ++q;
        colour[i] = q;
        ncol = std::max(ncol, q + 1);
      }
    } else {
      // on the graph of the blocks: the blocks that the entries of a block's
      // rows reach (built in parallel), and for each block the earlier blocks
      // whose entries reach it; a block takes the smallest colour that no
      // earlier neighbour in either direction holds
      const int nb = (n + kBmcRows - 1) / kBmcRows;
      std::vector<std::vector<int>> reach(nb);
      host_parallel(nb, [&](int64_t b0, int64_t b1) {
        std::vector<int> tmp;
        for (int bi = (int)b0; bi < (int)b1; ++bi) {
          tmp.clear();
          const int i0 = bi * kBmcRows, i1 = std::min(n, i0 + kBmcRows);
          for (int k = c->h_rp[i0]; k < c->h_rp[i1]; ++k) {
            const int bj = c->h_ci[k] / kBmcRows;
            if (bj != bi) tmp.push_back(bj);
          }
          std::sort(tmp.begin(), tmp.end());
          tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
          reach[bi] = tmp;
        }
      }, 1024);
      std::vector<int> tptr(nb + 1, 0), tsrc;
      for (int bj = 0; bj < nb; ++bj)
        for (int b : reach[bj])
          if (b > bj) tptr[b + 1]++;
      for (int b = 0; b < nb; ++b) tptr[b + 1] += tptr[b];
      tsrc.resize(tptr[nb]);
      {
        std::vector<int> fillp(tptr.begin(), tptr.end() - 1);
        for (int bj = 0; bj < nb; ++bj)
          for (int b : reach[bj])
            if (b > bj) tsrc[fillp[b]++] = bj;
      }
      std::vector<int> bcol(nb, -1);
      for (int bi = 0; bi < nb; ++bi) {
        for (int b : reach[bi])
          if (b < bi) mark[bcol[b]] = bi;
        for (int k = tptr[bi]; k < tptr[bi + 1]; ++k) mark[bcol[tsrc[k]]] = bi;
        int q = 0;
        while (q < 1023 && mark[q] == bi) ++q;
        bcol[bi] = q;
        ncol = std::max(ncol, q + 1);
      }
      for (int i = 0; i < n; ++i) colour[i] = bcol[i / kBmcRows];
    }
    phase("  colouring");
    c->ncolors = ncol;
    c->h_colour = colour;
    std::vector<int> start(ncol + 1, 0);
    for (int i = 0; i < n; ++i) start[colour[i] + 1]++;
    for (int q = 0; q < ncol; ++q) start[q + 1] += start[q];
    perm.assign(n, 0);
    std::vector<int> iperm(n);
    for (int i = 0; i < n; ++i) {
      const int p = start[colour[i]]++;
      perm[p] = i;
      iperm[i] = p;
    }
    phase("  order");
    c->h_frp.assign(n + 1, 0);
    for (int p = 0; p < n; ++p) c->h_frp[p + 1] = c->h_frp[p] + (c->h_rp[perm[p] + 1] - c->h_rp[perm[p]]);
    c->h_fci.resize(c->h_ci.size());
    pos.resize(c->h_ci.size());
    host_parallel(n, [&](int64_t p0, int64_t p1) {
      std::vector<std::pair<int, int>> tmp;
      for (int p = (int)p0; p < (int)p1; ++p) {
        const int i = perm[p];
        tmp.clear();
        for (int k = c->h_rp[i]; k < c->h_rp[i + 1]; ++k) tmp.emplace_back(iperm[c->h_ci[k]], k);
        std::sort(tmp.begin(), tmp.end());
        int w = c->h_frp[p];
        for (auto &e : tmp) {
          c->h_fci[w] = e.first;
          pos[w] = e.second;
          ++w;
        }
      }
    });
  } else if (c->prm.precond == SHESOLVE_PRECOND_ILUK && (c->prm.fill_level > 0 || !c->seg_fill.empty())) {
    if (int e = iluk_pattern(c, pos, false)) return e;
  } else {
    c->h_frp = c->h_rp;
    c->h_fci = c->h_ci;
  }
  phase("  factor pattern");
  c->factor_is_A = pos.empty();
  c->h_perm = perm;
  c->fnnz = c->h_frp[n];
  c->h_fdg.assign(n, -1);
  int maxlen = 0, nodiag = n;
  {
    std::mutex mu;
    host_parallel(n, [&](int64_t i0, int64_t i1) {
      int ml = 0, nd = n;
      for (int i = (int)i0; i < (int)i1; ++i) {
        ml = std::max(ml, c->h_frp[i + 1] - c->h_frp[i]);
        for (int k = c->h_frp[i]; k < c->h_frp[i + 1]; ++k)
          if (c->h_fci[k] == i) { c->h_fdg[i] = k; break; }
        if (c->h_fdg[i] < 0) nd = std::min(nd, i);
      }
      std::lock_guard<std::mutex> g(mu);
      maxlen = std::max(maxlen, ml);
      nodiag = std::min(nodiag, nd);
    });
  }
  if (nodiag < n) {
    set_err("ILU0: row %d has no diagonal entry", nodiag);
    return 2;
  }
  c->fmaxlen = maxlen;
  if (maxlen > kMaxRowLDS) {
    set_err("ILU0: row length %d exceeds %d", maxlen, kMaxRowLDS);
    return 2;
  }
  phase("  diagonal");
  std::vector<int> Lptr, Lrows, Uptr, Urows;
  bool colour_levels = false;
  std::vector<int> bst;
  if (!perm.empty() && c->prm.precond == SHESOLVE_PRECOND_BMC_ILU0) {
    // blocks in the solve space: a block's rows stay consecutive (one colour,
    // natural order), so a block starts where the natural block changes
    c->bmc_cptr.assign(1, 0);
    for (int p = 0; p < n; ++p) {
      if (p > 0 && perm[p] / kBmcRows == perm[p - 1] / kBmcRows) continue;
      if (p > 0 && c->h_colour[perm[p]] != c->h_colour[perm[p - 1]]) c->bmc_cptr.push_back((int)bst.size());
      bst.push_back(p);
    }
    c->bmc_cptr.push_back((int)bst.size());
    bst.push_back(n);
    for (size_t k = 0; k + 1 < bst.size(); ++k)
      if (bst[k + 1] - bst[k] > kBmcRows) {
        set_err("block multicolour ordering: block %zu holds %d rows", k, bst[k + 1] - bst[k]);
        return -1;
      }
    HIPCHK(hipMalloc(&c->d_bmc_start, sizeof(int) * bst.size()));
    HIPCHK(hipMemcpyAsync(c->d_bmc_start, bst.data(), sizeof(int) * bst.size(), hipMemcpyHostToDevice, c->stream));
  }
  const bool bmc = !bst.empty();
  if (!perm.empty() && !bmc) {
    // multicolour ordering: colour q occupies the contiguous range
    // [cstart[q], cstart[q+1]) of the factor; use whole colours as levels
    // (forward: ascending colours, backward: descending) when every factor
    // entry links different colours
    std::vector<int> cstart(1, 0), colour_of(n);
    for (int p = 0; p < n; ++p) {
      if (p > 0 && c->h_colour[perm[p]] != c->h_colour[perm[p - 1]]) cstart.push_back(p);
      colour_of[p] = (int)cstart.size() - 1;
    }
    cstart.push_back(n);
    std::atomic<bool> valid{true};
    host_parallel(n, [&](int64_t p0, int64_t p1) {
      for (int p = (int)p0; p < (int)p1 && valid.load(std::memory_order_relaxed); ++p)
        for (int k = c->h_frp[p]; k < c->h_frp[p + 1]; ++k)
          if (c->h_fci[k] != p && colour_of[c->h_fci[k]] == colour_of[p]) { valid = false; break; }
    });
    colour_levels = valid;
    if (valid) {
      const int nq = (int)cstart.size() - 1;
      Lptr.assign(cstart.begin(), cstart.end());
      Lrows.resize(n);
      for (int p = 0; p < n; ++p) Lrows[p] = p;
      Uptr.assign(1, 0);
      Urows.clear();
      for (int q = nq - 1; q >= 0; --q) {
        for (int p = cstart[q]; p < cstart[q + 1]; ++p) Urows.push_back(p);
        Uptr.push_back((int)Urows.size());
      }
    }
  }
  phase("  colour levels");
  if (bmc) compute_levels_blocked(c->h_frp, c->h_fci, c->h_fdg, n, bst, c->bmc_cptr, Lptr, Lrows, Uptr, Urows);
  else if (!colour_levels &&
           !(perm.empty() && c->fseg_ptr.size() > 2 &&
             compute_levels_segments(c->h_frp, c->h_fci, c->h_fdg, n, c->fseg_ptr, Lptr, Lrows, Uptr, Urows)))
    compute_levels(c->h_frp, c->h_fci, c->h_fdg, n, Lptr, Lrows, Uptr, Urows);
  phase("  levels");
  if (perm.empty()) {
    // factor segments (systems of a batch, per-energy slabs or the whole
    // matrix): one workgroup walks each segment's levels, unless one launch
    // per level of the whole matrix is estimated faster
    if (segment_sweeps_pay(c, Lptr, Lrows))
      if (int e = segment_levels(c)) return e;
    phase("  segment levels");
  }
  c->lvL.ptr = Lptr;
  c->lvU.ptr = Uptr;
  HIPCHK(hipMalloc(&c->lvL.d_rows, sizeof(int) * std::max(n, 1)));
  HIPCHK(hipMalloc(&c->lvU.d_rows, sizeof(int) * std::max(n, 1)));
  HIPCHK(hipMemcpyAsync(c->lvL.d_rows, Lrows.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->lvU.d_rows, Urows.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
  if (c->factor_is_A) {
    c->d_frp = c->d_rp;
    c->d_fci = c->d_ci;
  } else {
    HIPCHK(hipMalloc(&c->d_frp, sizeof(int) * (n + 1)));
    HIPCHK(hipMalloc(&c->d_fci, sizeof(int) * (std::max<int64_t>(c->fnnz, 1) + kVecPad)));
    HIPCHK(hipMemcpyAsync(c->d_frp, c->h_frp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_fci, c->h_fci.data(), sizeof(int) * c->fnnz, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMalloc(&c->d_fpos, sizeof(int) * std::max<int64_t>(c->fnnz, 1)));
    HIPCHK(hipMemcpyAsync(c->d_fpos, pos.data(), sizeof(int) * c->fnnz, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  if (!perm.empty()) {
    HIPCHK(hipMalloc(&c->d_perm, sizeof(int) * n));
    HIPCHK(hipMemcpyAsync(c->d_perm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMalloc(&c->z64, sizeof(double) * n));
    HIPCHK(hipMalloc(&c->z32, sizeof(float) * n));
  }
  HIPCHK(hipMalloc(&c->d_fdg, sizeof(int) * std::max(n, 1)));
  HIPCHK(hipMemcpyAsync(c->d_fdg, c->h_fdg.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMalloc(&c->d_f64, sizeof(double) * std::max<int64_t>(c->fnnz, 1)));
  c->flpr = pick_lpr((double)c->fnnz / std::max(n, 1) / 2.0);
  phase("  uploads");
  // split pattern of the factor
  hvec<int> Lrp(n + 1, 0), Urp(n + 1, 0), Lci, Uci;
  for (int i = 0; i < n; ++i) {
    Lrp[i + 1] = Lrp[i] + (c->h_fdg[i] - c->h_frp[i]);
    Urp[i + 1] = Urp[i] + (c->h_frp[i + 1] - c->h_fdg[i] - 1);
  }
  Lci.resize(Lrp[n]);
  Uci.resize(Urp[n]);
  host_parallel(n, [&](int64_t i0, int64_t i1) {
    const auto f = c->h_fci.begin();
    for (int i = (int)i0; i < (int)i1; ++i) {
      std::copy(f + c->h_frp[i], f + c->h_fdg[i], Lci.begin() + Lrp[i]);
      std::copy(f + c->h_fdg[i] + 1, f + c->h_frp[i + 1], Uci.begin() + Urp[i]);
    }
  });
  c->nL = Lci.size();
  c->nU = Uci.size();
  std::vector<int> loLrp, loUrp;
  if (c->nseg > 0) {
    // segment sweeps: store each triangle in level order; the natural-indexed
    // "row pointer" becomes the start of each row in that storage (k_split
    // writes through it), lrp[q] the start of the q-th level-ordered row
    // columns become positions in the same schedule (the sweep's window index)
    auto reorder = [&](hvec<int> &rp, hvec<int> &ci, const std::vector<int> &rows,
                       std::vector<int> &lrp) {
      hvec<int> start(n + 1), lci(ci.size());
      std::vector<int> qpos(n);
      lrp.assign(n + 1, 0);
      // row lengths in schedule order, their prefix sums in 16 slices (slice
      // sums first), then the columns mapped and placed, all on host threads
      constexpr int kSlices = 16;
      int64_t part[kSlices + 1] = {0};
      std::vector<std::thread> th;
      for (int t = 0; t < kSlices; ++t)
        th.emplace_back([&, t] {
          int64_t acc = 0;
          for (int q = (int)((int64_t)n * t / kSlices); q < (int)((int64_t)n * (t + 1) / kSlices); ++q) {
            qpos[rows[q]] = q;
            acc += rp[rows[q] + 1] - rp[rows[q]];
          }
          part[t + 1] = acc;
        });
      for (auto &x : th) x.join();
      th.clear();
      for (int t = 0; t < kSlices; ++t) part[t + 1] += part[t];
      for (int t = 0; t < kSlices; ++t)
        th.emplace_back([&, t] {
          int w = (int)part[t];
          for (int q = (int)((int64_t)n * t / kSlices); q < (int)((int64_t)n * (t + 1) / kSlices); ++q) {
            const int row = rows[q];
            start[row] = w;
            lrp[q] = w;
            for (int k = rp[row]; k < rp[row + 1]; ++k) lci[w++] = qpos[ci[k]];
          }
        });
      for (auto &x : th) x.join();
      const int w = (int)part[kSlices];
      lrp[n] = w;
      start[n] = w;
      rp.swap(start);
      ci.swap(lci);
    };
    phase("  split");
    reorder(Lrp, Lci, c->h_segrowsL, loLrp);
    reorder(Urp, Uci, c->h_segrowsU, loUrp);
    phase("  level order");
    for (int upper = 0; upper < 2; ++upper) {
      const std::vector<int> &lv = c->h_lvlq[upper], &lrp = upper ? loUrp : loLrp;
      const hvec<int> &lci = upper ? Uci : Lci;
      int reach = 0;
      std::mutex mu;
      host_parallel((int64_t)lv.size() - 1, [&](int64_t l0, int64_t l1) {
        int r = 0;
        for (int64_t li = l0; li < l1; ++li)
          for (int q = lv[li]; q < lv[li + 1]; ++q)
            for (int k = lrp[q]; k < lrp[q + 1]; ++k) r = std::max(r, lv[li + 1] - lci[k]);
        std::lock_guard<std::mutex> g(mu);
        reach = std::max(reach, r);
      }, 1024);
      c->seg_reach[upper] = reach;
    }
    {
      // chunks of the chain sweeps: consecutive levels of a segment while they
      // hold at most kChainRows rows and kChainEnts entries
      // (each segment's chunks on host threads, then concatenated)
      std::atomic<bool> ok{true};
      std::vector<int> ch[2], cp[2];
      for (int upper = 0; upper < 2 && ok; ++upper) {
        const std::vector<int> &lv = c->h_lvlq[upper], &sg = c->h_segq[upper], &lrp = upper ? loUrp : loLrp;
        const int S = (int)sg.size() - 1;
        std::vector<std::vector<int>> sch(std::max(S, 0));
        host_parallel(S, [&](int64_t s0, int64_t s1) {
          for (int s = (int)s0; s < (int)s1 && ok; ++s) {
            int la = sg[s];
            while (la < sg[s + 1]) {
              int lb = la;
              while (lb < sg[s + 1] && lv[lb + 1] - lv[la] <= kChainRows && lrp[lv[lb + 1]] - lrp[lv[la]] <= kChainEnts)
                ++lb;
              if (lb == la) { ok = false; break; }  // a level wider than a chunk
              const int hd[8] = {la, lb, lv[la], lv[lb], lrp[lv[la]], lrp[lv[lb]], 0, 0};
              sch[s].insert(sch[s].end(), hd, hd + 8);
              la = lb;
            }
          }
        }, 1);
        if (!ok) break;
        cp[upper].assign(S + 1, 0);
        for (int s = 0; s < S; ++s) cp[upper][s + 1] = cp[upper][s] + (int)(sch[s].size() / 8);
        ch[upper].resize((size_t)cp[upper][S] * 8);
        host_parallel(S, [&](int64_t s0, int64_t s1) {
          for (int s = (int)s0; s < (int)s1; ++s)
            std::copy(sch[s].begin(), sch[s].end(), ch[upper].begin() + (size_t)cp[upper][s] * 8);
        }, 1);
      }
      if (ok) {
        int **dch[2] = {&c->d_chunkL, &c->d_chunkU}, **dcp[2] = {&c->d_cptrL, &c->d_cptrU};
        for (int upper = 0; upper < 2; ++upper) {
          HIPCHK(hipMalloc(dch[upper], sizeof(int) * std::max<size_t>(ch[upper].size(), 8)));
          HIPCHK(hipMalloc(dcp[upper], sizeof(int) * cp[upper].size()));
          HIPCHK(hipMemcpy(*dch[upper], ch[upper].data(), sizeof(int) * ch[upper].size(), hipMemcpyHostToDevice));
          HIPCHK(hipMemcpy(*dcp[upper], cp[upper].data(), sizeof(int) * cp[upper].size(), hipMemcpyHostToDevice));
        }
      }
      if (!c->fseg_sys.empty()) {
        // positions of segment s: [lv[sg[s]], lv[sg[s + 1]]) in both schedules
        const std::vector<int> &lv = c->h_lvlq[0], &sg = c->h_segq[0];
        std::vector<int> qs((n + kBlock - 1) / kBlock, -2);
        for (size_t s = 0; s + 1 < sg.size(); ++s) {
          const int sy = c->fseg_sys[s], q0 = lv[sg[s]], q1 = lv[sg[s + 1]];
          for (int b = q0 / kBlock; q0 < q1 && b <= (q1 - 1) / kBlock; ++b) qs[b] = (qs[b] == -2 || qs[b] == sy) ? sy : -1;
        }
        for (int &e : qs) e = e == -2 ? -1 : e;
        HIPCHK(hipMalloc(&c->d_qsys, sizeof(int) * qs.size()));
        HIPCHK(hipMemcpy(c->d_qsys, qs.data(), sizeof(int) * qs.size(), hipMemcpyHostToDevice));
      }
      for (int upper = 0; upper < 2; ++upper) {
        std::vector<int>().swap(c->h_lvlq[upper]);
        std::vector<int>().swap(c->h_segq[upper]);
      }
    }
    phase("  reach, chunks");
    if (getenv("SHESOLVE_VERBOSE"))
      fprintf(stderr, "[shesolve] segment sweeps: %d segments, %d threads, reach %d / %d positions\n", c->nseg,
              c->seg_nt, c->seg_reach[0], c->seg_reach[1]);
    {
      // lanes per row: enough that kSegJ entries per lane cover the longest
      // rows, at most kSegLprMax (longer rows loop over the rest)
      int mx = 0;
      for (int q = 0; q < n; ++q) mx = std::max(mx, std::max(loLrp[q + 1] - loLrp[q], loUrp[q + 1] - loUrp[q]));
      c->seg_lpr = mx <= 2 * kSegJ ? 2 : (mx <= 4 * kSegJ ? 4 : 8);
      c->seg_lpr = std::min(c->seg_lpr, kSegLprMax);
    }
    {
      std::vector<int> posL(n), mapLU(n);
      for (int q = 0; q < n; ++q) posL[c->h_segrowsL[q]] = q;
      for (int q = 0; q < n; ++q) mapLU[q] = posL[c->h_segrowsU[q]];
      HIPCHK(hipMalloc(&c->d_mapLU, sizeof(int) * n));
      HIPCHK(hipMemcpy(c->d_mapLU, mapLU.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMalloc(&c->d_dq64, sizeof(double) * n));
    HIPCHK(hipMalloc(&c->d_tq64, sizeof(double) * n));
    HIPCHK(hipMalloc(&c->d_zq64, sizeof(double) * n));
    HIPCHK(hipMalloc(&c->d_dq32, sizeof(float) * n));
    HIPCHK(hipMalloc(&c->d_tq32, sizeof(float) * n));
    HIPCHK(hipMalloc(&c->d_zq32, sizeof(float) * n));
    std::vector<int>().swap(c->h_segrowsL);
    std::vector<int>().swap(c->h_segrowsU);
    HIPCHK(hipMalloc(&c->d_loLrp, sizeof(int) * (n + 1)));
    HIPCHK(hipMalloc(&c->d_loUrp, sizeof(int) * (n + 1)));
    HIPCHK(hipMemcpy(c->d_loLrp, loLrp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_loUrp, loUrp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
  }
  phase("  segment uploads");
  // contiguous levels get CSR-stream row blocks
  auto runs = [&](const Levels &lv, const std::vector<int> &rows, const hvec<int> &rp,
                  std::vector<shesolve_ctx::LevelRun> &out, std::vector<int> &blk) {
    out.clear();
    blk.clear();
    for (int l = 0; l < lv.count(); ++l) {
      const int b = lv.ptr[l], e = lv.ptr[l + 1];
      bool contiguous = e > b;
      for (int q = b + 1; q < e && contiguous; ++q) contiguous = rows[q] == rows[q - 1] + 1;
      shesolve_ctx::LevelRun r{0, 0, 0, 0};
      if (contiguous) {
        r.r0 = rows[b];
        r.r1 = rows[e - 1] + 1;
        r.blk_off = (int)blk.size();
        std::vector<int> bb;
        make_blocks(rp.data(), r.r0, r.r1, kSweepItems * kStreamThreads, bb, kSweepBlockRows);
        // stored as boundaries: blk_cnt blocks need blk_cnt + 1 entries
        for (int v : bb) blk.push_back(v);
        r.blk_cnt = (int)bb.size() - 1;
      }
      out.push_back(r);
    }
  };
  std::vector<int> lblk, ublk;
  c->c0 = 0;
  if (c->nseg == 0) {
    runs(c->lvL, Lrows, Lrp, c->runL, lblk);
    runs(c->lvU, Urows, Urp, c->runU, ublk);
    bool all_runs = !c->runL.empty() && c->runL[0].r0 == 0 && Lrp[c->runL[0].r1] == 0;
    for (const auto &r : c->runL) all_runs = all_runs && r.blk_cnt > 0;
    for (const auto &r : c->runU) all_runs = all_runs && r.blk_cnt > 0;
    if (all_runs && c->runL.size() > 1) c->c0 = c->runL[0].r1;
  }
  phase("  level runs");
  HIPCHK(hipMalloc(&c->d_Lrp, sizeof(int) * (n + 1)));
  HIPCHK(hipMalloc(&c->d_Urp, sizeof(int) * (n + 1)));
  HIPCHK(hipMalloc(&c->d_Lci, sizeof(int) * std::max<int64_t>(c->nL, 1)));
  HIPCHK(hipMalloc(&c->d_Uci, sizeof(int) * std::max<int64_t>(c->nU, 1)));
  HIPCHK(hipMalloc(&c->d_L64, sizeof(double) * std::max<int64_t>(c->nL, 1)));
  HIPCHK(hipMalloc(&c->d_U64, sizeof(double) * std::max<int64_t>(c->nU, 1)));
  HIPCHK(hipMalloc(&c->d_L32, sizeof(float) * std::max<int64_t>(c->nL, 1)));
  HIPCHK(hipMalloc(&c->d_U32, sizeof(float) * std::max<int64_t>(c->nU, 1)));
  HIPCHK(hipMalloc(&c->d_dinv64, sizeof(double) * n));
  HIPCHK(hipMalloc(&c->d_dinv32, sizeof(float) * n));
  HIPCHK(hipMalloc(&c->d_lblk, sizeof(int) * std::max<size_t>(lblk.size(), 1)));
  HIPCHK(hipMalloc(&c->d_ublk, sizeof(int) * std::max<size_t>(ublk.size(), 1)));
  HIPCHK(hipMemcpyAsync(c->d_Lrp, Lrp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d_Urp, Urp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, c->stream));
  if (c->nL) HIPCHK(hipMemcpyAsync(c->d_Lci, Lci.data(), sizeof(int) * c->nL, hipMemcpyHostToDevice, c->stream));
  if (c->nU) HIPCHK(hipMemcpyAsync(c->d_Uci, Uci.data(), sizeof(int) * c->nU, hipMemcpyHostToDevice, c->stream));
  if (!lblk.empty()) HIPCHK(hipMemcpyAsync(c->d_lblk, lblk.data(), sizeof(int) * lblk.size(), hipMemcpyHostToDevice, c->stream));
  if (!ublk.empty()) HIPCHK(hipMemcpyAsync(c->d_ublk, ublk.data(), sizeof(int) * ublk.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  phase("  triangle uploads");
  if (!c->bmc_cptr.empty()) {
    // the two triangles' layouts are independent: the backward one on a thread of its own
    // (its error message, kept per thread, is carried back to this one)
    int eu = 0;
    std::string eu_msg;
    std::thread tu([&] {
      eu = build_bmc_part(c, bst, true, Urp, Uci, c->bmcU);
      if (eu) eu_msg = g_err;
    });
    const int el = build_bmc_part(c, bst, false, Lrp, Lci, c->bmcL);
    tu.join();
    if (el) return el;
    if (eu) {
      g_err = eu_msg;
      return eu;
    }
    phase("  grouped layout");
  }
  return 0;
}

// Copies A's values into the factor pattern (gather for block ILU) and factorises.
int refactor(shesolve_ctx *c) {
  if (c->prm.precond == SHESOLVE_PRECOND_NONE || !c->d_frp) return 0;
  if (c->solve_perm) {
    HIPCHK(hipMemcpyAsync(c->d_f64, c->d_sv64, sizeof(double) * c->fnnz, hipMemcpyDeviceToDevice, c->stream));
  } else if (c->factor_is_A) {
    HIPCHK(hipMemcpyAsync(c->d_f64, c->d_v64, sizeof(double) * c->nnz, hipMemcpyDeviceToDevice, c->stream));
  } else {
    hipLaunchKernelGGL(k_gather, dim3(grid_for(c->fnnz, kBlock * 4)), dim3(kBlock), 0, c->stream, c->fnnz, c->d_fpos,
                       c->d_v64, c->d_f64);
  }
  const double drop = c->prm.precond == SHESOLVE_PRECOND_ILUK ? c->prm.drop_tol : 0.0;
  const bool short_rows = c->fmaxlen <= kShortRow;
  for (int l = 0; l < c->lvL.count(); ++l) {
    int cnt = c->lvL.ptr[l + 1] - c->lvL.ptr[l];
    if (short_rows)
      hipLaunchKernelGGL(k_ilu0_short, dim3(grid_for((int64_t)cnt * kShortLanes, 256)), dim3(256), 0, c->stream, cnt,
                         c->lvL.d_rows + c->lvL.ptr[l], c->d_frp, c->d_fci, c->d_fdg, c->d_f64, drop);
    else
      hipLaunchKernelGGL(k_ilu0_level, dim3(cnt), dim3(64), 0, c->stream, cnt, c->lvL.d_rows + c->lvL.ptr[l],
                         c->d_frp, c->d_fci, c->d_fdg, c->d_f64, drop);
  }
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_split, dim3(grid_for(c->n, kBlock)), dim3(kBlock), 0, c->stream, c->n, c->d_frp, c->d_fdg,
                     c->d_f64, c->d_Lrp, c->d_Urp, c->d_L64, c->d_U64, c->d_dinv64);
  if (c->nL)
    hipLaunchKernelGGL((k_cast<double, float>), dim3(grid_for(c->nL, kBlock * 4)), dim3(kBlock), 0, c->stream, c->nL,
                       c->d_L64, c->d_L32);
  if (c->nU)
    hipLaunchKernelGGL((k_cast<double, float>), dim3(grid_for(c->nU, kBlock * 4)), dim3(kBlock), 0, c->stream, c->nU,
                       c->d_U64, c->d_U32);
  hipLaunchKernelGGL((k_cast<double, float>), dim3(grid_for(c->n, kBlock * 4)), dim3(kBlock), 0, c->stream,
                     (int64_t)c->n, c->d_dinv64, c->d_dinv32);
  if (!c->bmc_cptr.empty()) {
    bmc_part_values(c, c->bmcL, c->d_L64, nullptr);
    bmc_part_values(c, c->bmcU, c->d_U64, c->d_dinv64);
  }
  if (c->nseg > 0) {
    // 1/diagonal in the backward schedule's order
    const int g = grid_for(c->n, kBlock);
    hipLaunchKernelGGL((k_qgather<double, double>), dim3(g), dim3(kBlock), 0, c->stream, c->n, c->d_rowsU,
                       c->d_dinv64, c->d_dq64);
    hipLaunchKernelGGL((k_qgather<double, float>), dim3(g), dim3(kBlock), 0, c->stream, c->n, c->d_rowsU,
                       c->d_dinv64, c->d_dq32);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int upload_values(shesolve_ctx *c) {
  if (c->solve_perm) {
    hipLaunchKernelGGL(k_gather, dim3(grid_for(c->fnnz, kBlock * 4)), dim3(kBlock), 0, c->stream, c->fnnz, c->d_fpos,
                       c->d_v64, c->d_sv64);
    hipLaunchKernelGGL((k_cast<double, float>), dim3(grid_for(c->fnnz, kBlock * 4)), dim3(kBlock), 0, c->stream,
                       c->fnnz, c->d_sv64, c->d_sv32);
  } else {
    hipLaunchKernelGGL((k_cast<double, float>), dim3(grid_for(c->nnz, kBlock * 4)), dim3(kBlock), 0, c->stream,
                       c->nnz, c->d_v64, c->d_v32);
  }
  HIPCHK(hipGetLastError());
  return refactor(c);
}

// Descriptors of the pipelined SpMV for the row blocks `blocks` (boundaries)
// or `pairs` ((first, end) per block) of the CSR with row pointer rp.
int build_desc(shesolve_ctx *c, const int *blocks, const int *pairs, int nblk, const int *rp,
               const std::vector<int> &hrp) {
  dfree(c->d_desc);
  c->desc_src = nullptr;
  c->desc_nblk = 0;
  c->desc_maxrow = 0;
  for (size_t i = 0; i + 1 < hrp.size(); ++i) c->desc_maxrow = std::max(c->desc_maxrow, hrp[i + 1] - hrp[i]);
  if (nblk <= 0) return 0;
  HIPCHK(hipMalloc(&c->d_desc, sizeof(int4) * nblk));
  hipLaunchKernelGGL(k_block_desc, dim3(grid_for(nblk, kBlock)), dim3(kBlock), 0, c->stream, nblk, blocks, pairs, rp,
                     c->d_desc);
  HIPCHK(hipGetLastError());
  c->desc_src = pairs ? pairs : blocks;
  c->desc_nblk = nblk;
  return 0;
}

// 16-bit column codes of the solve-space matrix for the coded pipelined SpMV,
// for the row blocks `flat` ((first, end) row pairs).  Block multicolour
// ordering only: its 32-row blocks keep consecutive natural rows together and in
// order, so column col lies in the block starting at col - perm[col] % 32.  Per
// row block, the distinct block starts of its columns, in order of first
// appearance, form its table (when a row block needs more than kCodeSlots, no
// codes are built and the SpMV reads the column indices); an entry's code is
// slot << 5 | perm[col] % 32.
int build_codes(shesolve_ctx *c, const std::vector<int> &flat) {
  free_codes(c);
  if (c->prm.precond != SHESOLVE_PRECOND_BMC_ILU0 || !c->solve_perm || c->h_perm.size() != (size_t)c->n) return 0;
  const int nb = (int)(flat.size() / 2);
  const std::vector<int> &rp = c->h_frp, &perm = c->h_perm;
  const hvec<int> &ci = c->h_fci;
  std::vector<std::vector<int>> tabs(nb);
  std::vector<unsigned short> code((size_t)c->fnnz + kVecPad, 0);
  std::atomic<bool> ok{true};
  const int nbn = (c->n + 31) / 32;  // natural 32-row blocks
  auto work = [&](int k0, int k1) {
    // a column's slot in the row block's table, by its natural block: seen[]
    // holds the row block that last met that natural block (tables are in
    // order of first appearance)
    std::vector<int> seen(nbn, -1);
    std::vector<unsigned char> slot_of(nbn, 0);
    for (int k = k0; k < k1 && ok.load(std::memory_order_relaxed); ++k) {
      const int e0 = rp[flat[2 * k]], e1 = rp[flat[2 * k + 1]];
      std::vector<int> &T = tabs[k];
      for (int e = e0; e < e1; ++e) {
        const int nat = perm[ci[e]], nbk = nat >> 5, off = nat & 31;
        if (seen[nbk] != k) {
          if ((int)T.size() == kCodeSlots) { ok = false; break; }
          seen[nbk] = k;
          slot_of[nbk] = (unsigned char)T.size();
          T.push_back(ci[e] - off);
        }
        code[e] = (unsigned short)(slot_of[nbk] << 5 | off);
      }
    }
  };
  const int nth = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (int w = 0; w < nth; ++w) th.emplace_back(work, (int)((int64_t)nb * w / nth), (int)((int64_t)nb * (w + 1) / nth));
  for (auto &x : th) x.join();
  if (!ok) return 0;
  std::vector<int> tab, desc((size_t)nb * 8, 0);
  for (int k = 0; k < nb; ++k) {
    const int r0 = flat[2 * k], r1 = flat[2 * k + 1];
    const int t0 = (int)tab.size();
    tab.insert(tab.end(), tabs[k].begin(), tabs[k].end());
    const int v[8] = {r0, r1, rp[r0], rp[r1], t0, (int)tab.size(), 0, 0};
    std::copy(v, v + 8, desc.begin() + (size_t)k * 8);
  }
  HIPCHK(hipMalloc(&c->d_code, sizeof(unsigned short) * code.size()));
  HIPCHK(hipMalloc(&c->d_tab, sizeof(int) * std::max<size_t>(tab.size(), 1)));
  HIPCHK(hipMalloc(&c->d_cdesc, sizeof(int) * std::max<size_t>(desc.size(), 8)));
  HIPCHK(hipMemcpy(c->d_code, code.data(), sizeof(unsigned short) * code.size(), hipMemcpyHostToDevice));
  if (!tab.empty()) HIPCHK(hipMemcpy(c->d_tab, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_cdesc, desc.data(), sizeof(int) * desc.size(), hipMemcpyHostToDevice));
  c->ntab = (int64_t)tab.size();
  c->code_src = c->s_pairs;
  c->code_nblk = nb;
  return 0;
}

// Solve space of a multicolour factor: the permuted matrix (the factor's
// pattern), its CSR-stream row blocks, and the permuted b / x of a solve.
int setup_solve_space(shesolve_ctx *c) {
  c->solve_perm = c->d_perm != nullptr && c->seg_ptr.empty();
  if (!c->solve_perm) {
    c->s_rp = c->d_rp;
    c->s_ci = c->d_ci;
    c->s_blk = c->d_ablk;
    c->s_nblk = c->nablk;
    return build_desc(c, c->d_ablk, nullptr, c->nablk, c->d_rp, c->h_rp);
  }
  dfree(c->d_v32);  // the natural-order FP32 copy served the launch autotuner only
  HIPCHK(hipMalloc(&c->d_sv64, sizeof(double) * (std::max<int64_t>(c->fnnz, 1) + kVecPad)));
  HIPCHK(hipMalloc(&c->d_sv32, sizeof(float) * (std::max<int64_t>(c->fnnz, 1) + kVecPad)));
  HIPCHK(hipMalloc(&c->pb64, sizeof(double) * std::max(c->n, 1)));
  HIPCHK(hipMalloc(&c->px64, sizeof(double) * std::max(c->n, 1)));
  // Row blocks of the permuted matrix, inside each colour, dealt in the order
  // of their first row's natural position: the SpMV then walks all colours
  // through the natural index space together, so the gathered x entries (the
  // rows' neighbours, near them in natural order) stay in L2 instead of being
  // fetched again for every colour.
  auto build_pairs = [&](int items, std::vector<int> &flat) {
    std::vector<std::pair<int, int>> pr;
    int start = 0;
    for (int q = 1; q <= c->n; ++q) {
      if (q < c->n && c->h_colour[c->h_perm[q]] == c->h_colour[c->h_perm[start]]) continue;
      std::vector<int> blk;
      make_blocks(c->h_frp.data(), start, q, spmv_cap(items), blk);
      for (size_t k = 0; k + 1 < blk.size(); ++k) pr.emplace_back(blk[k], blk[k + 1]);
      start = q;
    }
    std::stable_sort(pr.begin(), pr.end(), [&](const std::pair<int, int> &x, const std::pair<int, int> &y) {
      return c->h_perm[x.first] < c->h_perm[y.first];
    });
    flat.resize(2 * pr.size());
    for (size_t k = 0; k < pr.size(); ++k) {
      flat[2 * k] = pr[k].first;
      flat[2 * k + 1] = pr[k].second;
    }
  };
  auto install = [&](const std::vector<int> &flat) -> int {
    free_codes(c);
    dfree(c->d_sblk);
    HIPCHK(hipMalloc(&c->d_sblk, sizeof(int) * std::max<size_t>(flat.size(), 2)));
    HIPCHK(hipMemcpy(c->d_sblk, flat.data(), sizeof(int) * flat.size(), hipMemcpyHostToDevice));
    c->s_blk = c->d_sblk;
    c->s_pairs = c->d_sblk;
    c->s_nblk = (int)(flat.size() / 2);
    return build_desc(c, nullptr, c->d_sblk, c->s_nblk, c->d_frp, c->h_frp);
  };
  c->s_rp = c->d_frp;
  c->s_ci = c->d_fci;
  c->spmv_lpr = 0;  // the interleaved block order needs the CSR-stream kernel
  std::vector<int> flat;
  if (!c->prm.autotune || c->tune_reuse) {
    if (c->tune_reuse) {
      c->items = c->tuned_sitems;
      c->xcd = c->tuned_sxcd;
    }
    build_pairs(c->items, flat);
    if (int e = install(flat)) return e;
    return build_codes(c, flat);
  }
  // Entries per thread and block-to-XCD mapping of the permuted matrix, timed
  // on it as the solve launches it (fused dot product, persistent grid), with
  // 4-byte column indices: the column codes are built once, for the winner
  // (with codes on every candidate the 161x81 L=3 device picked the same shape,
  // 8 entries and the strided mapping, at 320 against 329-370 us).  The
  // natural matrix's choice does not carry over (contiguous eighths of the
  // interleaved order give each XCD its own eighth of the natural index space
  // in every colour), and plain launches do not tell: on the 161x81 L=3
  // device two mappings ran 362 and 375 us plain but 414 and 560 us fused.
  PhaseClock phase(c->stream);
  hipLaunchKernelGGL((k_fill<float>), dim3(kGrid), dim3(kBlock), 0, c->stream, (int64_t)c->fnnz, c->d_sv32, 1.0f);
  hipLaunchKernelGGL((k_fill<float>), dim3(kGrid), dim3(kBlock), 0, c->stream, (int64_t)c->n, c->v32[2], 1.0f);
  float best = 1e30f;
  int best_items = c->items, best_xcd = c->xcd;
  std::vector<int> best_flat;
  const int cand[3] = {4, 8, 16};
  for (int ii = 0; ii < 3; ++ii) {
    build_pairs(cand[ii], flat);
    if (int e = install(flat)) return e;
    c->items = cand[ii];
    for (int xm = 0; xm < 2; ++xm) {
      c->xcd = xm;
      launch_spmv<float, float, float>(c, 0, c->d_sv32, c->v32[0], c->v32[1], nullptr, c->v32[2], nullptr, c->d_part);
      HIPCHK(hipEventRecord(c->ev0, c->stream));
      for (int r = 0; r < 5; ++r)
        launch_spmv<float, float, float>(c, 0, c->d_sv32, c->v32[0], c->v32[1], nullptr, c->v32[2], nullptr,
                                         c->d_part);
      HIPCHK(hipEventRecord(c->ev1, c->stream));
      HIPCHK(hipEventSynchronize(c->ev1));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
      c->tune_us[ii * 2 + xm] = ms * 1e3 / 5;  // the solve space's candidates replace the natural matrix's
      if (ms < best) {
        best = ms;
        best_items = cand[ii];
        best_xcd = xm;
        best_flat = flat;
      }
    }
  }
  c->items = best_items;
  c->xcd = best_xcd;
  c->tuned_sitems = best_items;
  c->tuned_sxcd = best_xcd;
  if (int e = install(best_flat)) return e;
  phase("  solve-space tuning");
  const int e = build_codes(c, best_flat);
  phase("  column codes");
  return e;
}

// One BiCGStab solve (right preconditioned) on the device, precision T.
// b, x device vectors of precision T.  Returns iterations; status in *st.
// confirm_true: on convergence of the recurrence residual, confirm with the
// explicit residual (FP64 solves).  FP32 inner solves of the refinement loop
// skip it: the outer FP64 residual is the acceptance test there.
template <typename TV, typename T>
int bicgstab_dev(shesolve_ctx *c, const TV *Av, const TV *Mv, const T *b, T *x, double tol, int max_iters,
                 int *st, double *res_out, bool confirm_true = true) {
  const int n = c->n;
  T **V;
  if constexpr (sizeof(T) == 8) V = (T **)c->v64; else V = (T **)c->v32;
  T *r = V[0], *r0 = V[1], *p = V[2], *v = V[3], *ph = V[4], *s = V[5], *sh = V[6], *t = V[7];
  double *part = c->d_part;
  Scalars hs{};
  const int vg = kGrid;
  // r = b - A x ; r0 = r ; (r,r)
  launch_spmv<TV, T, T>(c, c->lpr, Av, x, r, b, nullptr, nullptr, nullptr);
  {
    KTimer kt(c, 3);
    hipLaunchKernelGGL((k_copy_dot<T>), dim3(vg), dim3(kBlock), 0, c->stream, n, r, r0, part);
  }
  // bnorm comes from b, not from the initial residual
  hipLaunchKernelGGL((k_fill<T>), dim3(vg), dim3(kBlock), 0, c->stream, (int64_t)n, p, (T)0);
  hipLaunchKernelGGL((k_fill<T>), dim3(vg), dim3(kBlock), 0, c->stream, (int64_t)n, v, (T)0);
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, c->stream, 0, kGrid, part, c->d_sc, c->d_norm);
  // |b|
  hipLaunchKernelGGL((k_copy_dot<T>), dim3(vg), dim3(kBlock), 0, c->stream, n, b, s, part);
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, c->stream, 4, kGrid, part, c->d_sc, c->d_norm);
  double bn = 0;
  HIPCHK(hipMemcpyAsync(&hs, c->d_sc, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(&bn, c->d_norm, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  double r_init = hs.bnorm;  // ||r_0||
  // store ||b|| as the normaliser
  {
    Scalars tmp = hs;
    tmp.bnorm = bn;
    tmp.res = bn > 0 ? r_init / bn : 0.0;
    tmp.tol = tol;
    tmp.halt = 0;
    HIPCHK(hipMemcpyAsync(c->d_sc, &tmp, sizeof(Scalars), hipMemcpyHostToDevice, c->stream));
    hs = tmp;
  }
  *st = 1;
  if (bn == 0.0) {
    hipLaunchKernelGGL((k_fill<T>), dim3(vg), dim3(kBlock), 0, c->stream, (int64_t)n, x, (T)0);
    HIPCHK(hipStreamSynchronize(c->stream));
    *st = 0;
    *res_out = 0.0;
    return 0;
  }
  if (hs.res <= tol) {
    *st = 0;
    *res_out = hs.res;
    return 0;
  }
  // Iterations are enqueued one ahead of the host's check of their scalars (a
  // pinned ring of two, each copy followed by an event), so the device does
  // not wait for the host between iterations; an iteration enqueued behind a
  // halt changes nothing (Scalars::halt).  The launches are not captured into
  // a graph: replaying one measured no faster (334 ms per solve either way on
  // the 161x81 L=3 device), and a capture in progress makes the synchronous
  // copies of other handles on other host threads fail.
  auto iter_launches = [&]() -> int {
    {
      KTimer kt(c, 3);
      hipLaunchKernelGGL((k_update_p<T>), dim3(vg), dim3(kBlock), 0, c->stream, n, p, r, v, c->d_sc);
    }
    launch_precond<TV, T>(c, Mv, p, ph);
    launch_spmv<TV, T, T>(c, c->lpr, Av, ph, v, nullptr, r0, nullptr, part);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, c->stream, 1, c->spmv_grid, part, c->d_sc, c->d_norm);
    {
      KTimer kt(c, 3);
      hipLaunchKernelGGL((k_update_s<T>), dim3(vg), dim3(kBlock), 0, c->stream, n, s, r, v, c->d_sc);
    }
    launch_precond<TV, T>(c, Mv, s, sh);
    launch_spmv<TV, T, T>(c, c->lpr, Av, sh, t, nullptr, s, t, part);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, c->stream, 2, c->spmv_grid, part, c->d_sc, c->d_norm);
    {
      KTimer kt(c, 3);
      hipLaunchKernelGGL((k_update_xr<T>), dim3(vg), dim3(kBlock), 0, c->stream, n, x, r, ph, sh, s, t, r0, c->d_sc,
                         part);
    }
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, c->stream, 3, kGrid, part, c->d_sc, c->d_norm);
    HIPCHK(hipGetLastError());
    return 0;
  };
  auto enqueue = [&](int k) -> int {
    if (int e = iter_launches()) return e;
    HIPCHK(hipMemcpyAsync(&c->h_ring[k & 1], c->d_sc, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipEventRecord(c->iter_ev[k & 1], c->stream));
    return 0;
  };
  int it = 0, enq = 0, best_it = 0;
  double best_res = hs.res;
  while (true) {
    while (enq < max_iters && enq < it + 2) {
      if (int e = enqueue(enq + 1)) return e;
      ++enq;
    }
    if (it >= enq) break;  // max_iters reached and checked
    HIPCHK(hipEventSynchronize(c->iter_ev[(it + 1) & 1]));
    ++it;
    hs = c->h_ring[it & 1];
    if (!std::isfinite(hs.res)) {
      *st = 2;
      break;
    }
    if (hs.res <= tol && !confirm_true) {
      *st = 0;
      break;
    }
    if (!confirm_true && hs.res > 1e4) {
      // FP32 inner solve diverging: give the refinement loop the chance to reject it
      *st = 2;
      break;
    }
    if (!confirm_true) {
      // FP32 inner solve stalled: no 1 % gain over kStallWindow iterations
      if (hs.res < 0.99 * best_res) {
        best_res = hs.res;
        best_it = it;
      } else if (it - best_it >= kStallWindow) {
        *st = 2;
        break;
      }
    }
    if (hs.res <= tol || hs.breakdown) {
      // the device halted: the iteration enqueued behind this one changed nothing
      HIPCHK(hipStreamSynchronize(c->stream));
      enq = it;
    }
    if (hs.res <= tol) {
      // confirm with the explicit residual r = b - A x
      launch_spmv<TV, T, T>(c, c->lpr, Av, x, r, b, nullptr, nullptr, nullptr);
      hipLaunchKernelGGL((k_copy_dot<T>), dim3(vg), dim3(kBlock), 0, c->stream, n, r, r0, part);
      hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, c->stream, 4, kGrid, part, c->d_sc, c->d_norm);
      double rn = 0;
      HIPCHK(hipMemcpyAsync(&rn, c->d_norm, sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      hs.res = rn / bn;
      if (hs.res <= tol) {
        *st = 0;
        break;
      }
      hs.breakdown = 1;  // restart from the true residual (r0 = r already)
    }
    if (hs.breakdown) {
      // restart: r0 = r, p = v = 0, rho = (r0,r), alpha = omega = 1
      hipLaunchKernelGGL((k_copy_dot<T>), dim3(vg), dim3(kBlock), 0, c->stream, n, r, r0, part);
      hipLaunchKernelGGL((k_fill<T>), dim3(vg), dim3(kBlock), 0, c->stream, (int64_t)n, p, (T)0);
      hipLaunchKernelGGL((k_fill<T>), dim3(vg), dim3(kBlock), 0, c->stream, (int64_t)n, v, (T)0);
      hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, c->stream, 0, kGrid, part, c->d_sc, c->d_norm);
      Scalars tmp{};
      HIPCHK(hipMemcpyAsync(&tmp, c->d_sc, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      tmp.bnorm = bn;
      tmp.res = std::sqrt(tmp.rr) / bn;
      tmp.tol = tol;
      HIPCHK(hipMemcpyAsync(c->d_sc, &tmp, sizeof(Scalars), hipMemcpyHostToDevice, c->stream));
      if (tmp.rr == 0.0) { *st = 0; hs.res = 0; break; }
    }
  }
  // an iteration still enqueued behind a host-side stop (FP32 stall or
  // divergence; the device had not halted) ran its updates: report its state,
  // which x now holds
  HIPCHK(hipStreamSynchronize(c->stream));
  if (enq > it && !c->h_ring[it & 1].halt) {
    it = enq;
    hs = c->h_ring[it & 1];
  }
  *res_out = hs.res;
  return it;
}

// FP64 residual norm ||b - A x|| into d_norm[slot] and r64 = b - A x
int residual64(shesolve_ctx *c, const double *b, const double *x, double *r, double *out) {
  launch_spmv<double, double, double>(c, c->lpr, sv64(c), x, r, b, nullptr, nullptr, nullptr);
  hipLaunchKernelGGL((k_copy_dot<double>), dim3(kGrid), dim3(kBlock), 0, c->stream, c->n, r, r, c->d_part);
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, c->stream, 4, kGrid, c->d_part, c->d_sc, c->d_norm);
  HIPCHK(hipMemcpyAsync(out, c->d_norm, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int solve_core(shesolve_ctx *c, const double *b, double *x, int warm, shesolve_report *rep);

// One record of the refinement history; the relative error against the
// reference solution when one is set (partial sums added on the host in a
// fixed order).
int ir_record(shesolve_ctx *c, const double *x, int step, int inner, int kind, double rel_res) {
  shesolve_ir_step rec{};
  rec.step = step;
  rec.inner_iterations = inner;
  rec.kind = kind;
  rec.rel_residual = rel_res;
  rec.rel_error = std::numeric_limits<double>::quiet_NaN();
  if (c->xref_cur) {
    hipLaunchKernelGGL(k_err_partials, dim3(kGrid), dim3(kBlock), 0, c->stream, c->n, x, c->xref_cur, c->d_part);
    std::vector<double> h(2 * kGrid);
    HIPCHK(hipMemcpyAsync(h.data(), c->d_part, sizeof(double) * kGrid, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(h.data() + kGrid, c->d_part + kPartStride, sizeof(double) * kGrid, hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double dd = 0, rr = 0;
    for (int i = 0; i < kGrid; ++i) {
      dd += h[i];
      rr += h[kGrid + i];
    }
    rec.rel_error = rr > 0 ? std::sqrt(dd / rr) : (dd > 0 ? std::numeric_limits<double>::infinity() : 0.0);
  }
  c->ir_hist.push_back(rec);
  return 0;
}

// Solve in the solve space: b and x go through the multicolour permutation
// once per solve when the matrix is stored permuted.
int solve_device(shesolve_ctx *c, const double *b, double *x, int warm, shesolve_report *rep) {
  c->ir_hist.clear();
  c->xref_cur = c->d_xref;
  if (!c->solve_perm) return solve_core(c, b, x, warm, rep);
  const int g = grid_for(c->n, kBlock);
  if (c->d_xref) {
    if (!c->d_xref_s) HIPCHK(hipMalloc(&c->d_xref_s, sizeof(double) * c->n));
    hipLaunchKernelGGL((k_permute<double>), dim3(g), dim3(kBlock), 0, c->stream, c->n, c->d_perm, c->d_xref,
                       c->d_xref_s, 0);
    c->xref_cur = c->d_xref_s;
  }
  hipLaunchKernelGGL((k_permute<double>), dim3(g), dim3(kBlock), 0, c->stream, c->n, c->d_perm, b, c->pb64, 0);
  if (warm)
    hipLaunchKernelGGL((k_permute<double>), dim3(g), dim3(kBlock), 0, c->stream, c->n, c->d_perm, x, c->px64, 0);
  if (int e = solve_core(c, c->pb64, c->px64, warm, rep)) return e;
  hipLaunchKernelGGL((k_permute<double>), dim3(g), dim3(kBlock), 0, c->stream, c->n, c->d_perm, c->px64, x, 1);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int solve_core(shesolve_ctx *c, const double *b, double *x, int warm, shesolve_report *rep) {
  auto t0 = std::chrono::steady_clock::now();
  shesolve_report r{};
  if (!c->d_rp) {
    set_err("no matrix set");
    return -1;
  }
  if (int e = ensure_vectors(c)) return e;
  const int n = c->n;
  if (warm) {
    // an initial guess worse than zero (||b - A x0|| > ||b||) is dropped
    double rn = 0, bn = 0;
    if (int e = residual64(c, b, x, c->v64[8], &rn)) return e;
    hipLaunchKernelGGL((k_copy_dot<double>), dim3(kGrid), dim3(kBlock), 0, c->stream, n, b, c->v64[0], c->d_part);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, c->stream, 4, kGrid, c->d_part, c->d_sc, c->d_norm);
    HIPCHK(hipMemcpyAsync(&bn, c->d_norm, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!(rn <= bn)) warm = 0;
  }
  if (!warm) hipLaunchKernelGGL((k_fill<double>), dim3(kGrid), dim3(kBlock), 0, c->stream, (int64_t)n, x, 0.0);
  if (c->prm.precision == SHESOLVE_FP64) {
    int st = 1;
    double res = 0;
    int it = bicgstab_dev<double, double>(c, sv64(c), c->d_f64, b, x, c->prm.tol, c->prm.max_iters, &st, &res);
    if (it < 0) return it;
    r.iterations = it;
    r.status = st;
    r.rel_residual = res;
    if (st != 0) {
      // the recurrence residual of an unconverged solve is not ||b - A x||: recompute it
      double rn = 0, bn = 0;
      if (int e = residual64(c, b, x, c->v64[8], &rn)) return e;
      hipLaunchKernelGGL((k_copy_dot<double>), dim3(kGrid), dim3(kBlock), 0, c->stream, n, b, c->v64[0], c->d_part);
      hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, c->stream, 4, kGrid, c->d_part, c->d_sc, c->d_norm);
      HIPCHK(hipMemcpyAsync(&bn, c->d_norm, sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      r.rel_residual = bn > 0 ? rn / bn : 0.0;
    }
    if (int e = ir_record(c, x, 0, it, 2, r.rel_residual)) return e;
  } else {
    // FP32 BiCGStab inside FP64 iterative refinement
    double *r64 = c->v64[8];
    float *r32 = c->v32[8];
    double bn = 0;
    // ||b||
    hipLaunchKernelGGL((k_copy_dot<double>), dim3(kGrid), dim3(kBlock), 0, c->stream, n, b, c->v64[0], c->d_part);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, c->stream, 4, kGrid, c->d_part, c->d_sc, c->d_norm);
    HIPCHK(hipMemcpyAsync(&bn, c->d_norm, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (bn == 0.0) {
      hipLaunchKernelGGL((k_fill<double>), dim3(kGrid), dim3(kBlock), 0, c->stream, (int64_t)n, x, 0.0);
      HIPCHK(hipStreamSynchronize(c->stream));
      r.status = 0;
    } else {
      r.status = 1;
      double prev_rn = -1.0;
      bool fallback = false, inner_failed = false;
      const int inner_cap = std::min(c->prm.max_iters, 1000);
      int first_it = -1;
      double *xprev = c->v64[6];  // idle during the FP32 inner solves
      int last_inner = 0;
      for (int k = 0; k <= c->prm.max_ir_steps; ++k) {
        double rn = 0;
        if (int e = residual64(c, b, x, r64, &rn)) return e;
        const bool rejected = prev_rn >= 0 && !(rn < prev_rn);
        if (int e = ir_record(c, x, r.ir_steps, last_inner, rejected ? 1 : 0, rn / bn)) return e;
        if (rejected) {
          // the last correction did not reduce the FP64 residual: undo it and
          // finish with FP64 BiCGStab from the best iterate
          HIPCHK(hipMemcpyAsync(x, xprev, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
          fallback = true;
          break;
        }
        r.rel_residual = rn / bn;
        if (r.rel_residual <= c->prm.tol) {
          r.status = 0;
          break;
        }
        if (inner_failed) {
          // the last FP32 solve stalled or broke down: keep its (useful)
          // correction and finish with FP64 BiCGStab
          fallback = true;
          break;
        }
        if (k == c->prm.max_ir_steps || r.iterations >= c->prm.max_iters) break;
        r.ir_steps++;
        prev_rn = rn;
        HIPCHK(hipMemcpyAsync(xprev, x, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
        // r32 = r64 / ||r64|| ; d = 0 ; inner solve ; x += ||r|| * d
        hipLaunchKernelGGL(k_cast_scale, dim3(kGrid), dim3(kBlock), 0, c->stream, n, r64, r32, c->d_norm);
        // the correction d (n floats) lives in the FP64 slot 7, idle during the FP32 solve
        float *dd = reinterpret_cast<float *>(c->v64[7]);
        hipLaunchKernelGGL((k_fill<float>), dim3(kGrid), dim3(kBlock), 0, c->stream, (int64_t)n, dd, 0.0f);
        int st = 1;
        double ires = 0;
        const int cap = first_it < 0 ? inner_cap : std::max(100, 4 * first_it);
        int it = bicgstab_dev<float, float>(c, sv32(c), (const float *)nullptr, r32, dd, c->prm.inner_tol,
                                            std::min(cap, c->prm.max_iters - r.iterations), &st, &ires, false);
        if (it < 0) return it;
        if (first_it < 0) first_it = it;
        last_inner = it;
        r.iterations += it;
        if (st != 0) inner_failed = true;
        HIPCHK(hipMemcpyAsync(c->d_norm, &rn, sizeof(double), hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_axpy_f32, dim3(kGrid), dim3(kBlock), 0, c->stream, n, x, dd, c->d_norm);
      }
      if (fallback) {
        // the FP64 solve is the safety net: it gets the full iteration budget.
        // x is the best iterate so far (residual r.rel_residual); it is kept in
        // r64 (idle from here on) and returned instead should the FP64 solve
        // stop unconverged with a larger true residual
        const double best_rel = r.rel_residual;
        HIPCHK(hipMemcpyAsync(r64, x, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
        int st = 1;
        double res = 0;
        int it = bicgstab_dev<double, double>(c, sv64(c), c->d_f64, b, x, c->prm.tol, c->prm.max_iters, &st, &res);
        if (it < 0) return it;
        r.iterations += it;
        r.status = st;
        r.rel_residual = res;
        if (st != 0) {
          double rn_true = 0;
          if (int e = residual64(c, b, x, c->v64[0], &rn_true)) return e;
          r.rel_residual = rn_true / bn;
          if (!(r.rel_residual <= best_rel)) {
            HIPCHK(hipMemcpyAsync(x, r64, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
            r.rel_residual = best_rel;
          }
        }
        if (int e = ir_record(c, x, r.ir_steps, it, 2, r.rel_residual)) return e;
      }
    }
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  r.setup_ms = c->last_setup_ms;
  r.solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (rep) *rep = r;
  return 0;
}

int set_pattern_host(shesolve_ctx *c, int64_t n, int64_t nnz, const int64_t *rp, const int32_t *ci, int nblocks,
                     const int64_t *ranges) {
  if (n <= 0 || n > 0x7fffffff) { set_err("n out of range"); return 2; }
  if (nnz < 0 || nnz > 0x7fffffff) { set_err("nnz must fit a 32-bit index"); return 3; }
  if (rp[0] != 0 || rp[n] != nnz) { set_err("row_ptr inconsistent with nnz"); return 4; }
  for (int64_t i = 0; i < n; ++i)
    if (rp[i + 1] < rp[i]) { set_err("row_ptr decreases at row %lld", (long long)i); return 4; }
  c->h_rp.resize(n + 1);
  if (ci != c->h_ci.data()) c->h_ci.resize(nnz);
  // copies and checks in parallel; the first row with a bad column is reported
  std::atomic<int64_t> bad{n};
  host_parallel(n, [&](int64_t i0, int64_t i1) {
    if (ci != c->h_ci.data()) std::copy(ci + rp[i0], ci + rp[i1], c->h_ci.begin() + rp[i0]);
    for (int64_t i = i0; i < i1; ++i) {
      c->h_rp[i] = (int)rp[i];
      for (int64_t k = rp[i]; k < rp[i + 1]; ++k)
        if (ci[k] < 0 || ci[k] >= n || (k > rp[i] && ci[k] <= ci[k - 1])) {
          int64_t cur = bad.load();
          while (i < cur && !bad.compare_exchange_weak(cur, i)) {
          }
          break;
        }
    }
  });
  c->h_rp[n] = (int)rp[n];
  if (bad < n) {
    set_err("column indices must be in range and strictly ascending per row (row %lld)", (long long)bad.load());
    return 5;
  }
  c->blk_of_row.clear();
  if (c->prm.precond == SHESOLVE_PRECOND_BLOCK_ILU0 || c->prm.precond == SHESOLVE_PRECOND_BLOCK_ILUK) {
    if (nblocks > 0 && !ranges) { set_err("block ranges missing"); return 7; }
    c->blk_of_row.assign(n, -1);
    for (int bk = 0; bk < nblocks; ++bk)
      for (int64_t i = std::max<int64_t>(0, ranges[2 * bk]); i < std::min<int64_t>(n, ranges[2 * bk + 1]); ++i)
        if (c->blk_of_row[i] < 0) c->blk_of_row[i] = bk;
    for (int64_t i = 0; i < n; ++i)
      if (c->blk_of_row[i] < 0) c->blk_of_row[i] = nblocks;
  }
  c->n = (int)n;
  c->nnz = nnz;
  return 0;
}

int ensure_vectors(shesolve_ctx *c);

// Row blocks of the whole matrix for `cap` entries per block: per system in a
// batch (no block straddles two systems), else over all rows.
void matrix_blocks(const shesolve_ctx *c, int cap, std::vector<int> &out) {
  out.clear();
  if (c->seg_ptr.empty()) {
    make_blocks(c->h_rp.data(), 0, c->n, cap, out);
    return;
  }
  std::vector<int> one;
  for (size_t s = 0; s + 1 < c->seg_ptr.size(); ++s) {
    one.clear();
    make_blocks(c->h_rp.data(), c->seg_ptr[s], c->seg_ptr[s + 1], cap, one);
    out.insert(out.end(), out.empty() ? one.begin() : one.begin() + 1, one.end());
  }
}

// Autotuner of the CSR-stream launch shape: entries per thread (4, 8, 16) x
// block-to-XCD mapping (strided, contiguous), each timed on three FP32 SpMV
// launches of this matrix (values cast below); the fastest is kept for the
// SpMV and for the sweeps of the factors.
int tune_stream(shesolve_ctx *c) {
  const int n = c->n;
  const int cand_items[3] = {4, 8, 16};
  std::vector<int> best_blocks;
  double best = 1e300;
  int best_items = 8, best_xcd = 1;
  HIPCHK(hipMalloc(&c->d_v32, sizeof(float) * (std::max<int64_t>(c->nnz, 1) + kVecPad)));
  hipLaunchKernelGGL((k_cast<double, float>), dim3(grid_for(c->nnz, kBlock * 4)), dim3(kBlock), 0, c->stream, c->nnz,
                     c->d_v64, c->d_v32);
  if (int e = ensure_vectors(c)) return e;
  hipLaunchKernelGGL((k_fill<float>), dim3(kGrid), dim3(kBlock), 0, c->stream, (int64_t)n, c->v32[0], 1.0f);
  hipLaunchKernelGGL((k_fill<float>), dim3(kGrid), dim3(kBlock), 0, c->stream, (int64_t)n, c->v32[2], 1.0f);
  // every candidate is timed as the solve launches it: with the fused dot
  // product on the persistent grid (plain launches rank the mappings differently)
  const bool reuse = c->tune_reuse;
  for (int ii = 0; ii < 3; ++ii) {
    if (!c->prm.autotune && cand_items[ii] != 8) continue;
    if (reuse && cand_items[ii] != c->tuned_items) continue;
    std::vector<int> ab;
    matrix_blocks(c, spmv_cap(cand_items[ii]), ab);
    int *d_ab = nullptr;
    HIPCHK(hipMalloc(&d_ab, sizeof(int) * ab.size()));
    HIPCHK(hipMemcpyAsync(d_ab, ab.data(), sizeof(int) * ab.size(), hipMemcpyHostToDevice, c->stream));
    for (int xm = 0; xm < 2; ++xm) {
      if (!c->prm.autotune && xm != 0) continue;
      if (reuse && xm != c->tuned_xcd) continue;
      c->items = cand_items[ii];
      c->xcd = xm;
      c->d_ablk = d_ab;
      c->nablk = (int)ab.size() - 1;
      launch_spmv<float, float, float>(c, 0, c->d_v32, c->v32[0], c->v32[1], nullptr, c->v32[2], nullptr, c->d_part);
      HIPCHK(hipEventRecord(c->ev0, c->stream));
      for (int r = 0; r < 3; ++r)
        launch_spmv<float, float, float>(c, 0, c->d_v32, c->v32[0], c->v32[1], nullptr, c->v32[2], nullptr,
                                         c->d_part);
      HIPCHK(hipEventRecord(c->ev1, c->stream));
      HIPCHK(hipEventSynchronize(c->ev1));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
      c->tune_us[ii * 2 + xm] = ms * 1e3 / 3;
      if (ms < best) {
        best = ms;
        best_items = cand_items[ii];
        best_xcd = xm;
        best_blocks = ab;
      }
    }
    c->d_ablk = nullptr;
    HIPCHK(hipFree(d_ab));
  }
  // vector-CSR candidates (lanes per row 4, 8, 16)
  int best_lpr = reuse ? c->tuned_lpr : 0;
  if (c->prm.autotune && !reuse) {
    c->items = best_items;
    c->xcd = best_xcd;
    const int cand_lpr[3] = {4, 8, 16};
    for (int li = 0; li < 3; ++li) {
      c->spmv_lpr = cand_lpr[li];
      launch_spmv<float, float, float>(c, 0, c->d_v32, c->v32[0], c->v32[1], nullptr, c->v32[2], nullptr, c->d_part);
      HIPCHK(hipEventRecord(c->ev0, c->stream));
      for (int r = 0; r < 3; ++r)
        launch_spmv<float, float, float>(c, 0, c->d_v32, c->v32[0], c->v32[1], nullptr, c->v32[2], nullptr,
                                         c->d_part);
      HIPCHK(hipEventRecord(c->ev1, c->stream));
      HIPCHK(hipEventSynchronize(c->ev1));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
      c->tune_us[6 + li] = ms * 1e3 / 3;
      if (ms < best) {
        best = ms;
        best_lpr = cand_lpr[li];
      }
    }
  }
  c->spmv_lpr = best_lpr;
  c->items = best_items;
  c->xcd = best_xcd;
  if (c->prm.autotune && !reuse) {
    c->tuned_n = n;
    c->tuned_precond = c->prm.precond;
    c->tuned_items = best_items;
    c->tuned_xcd = best_xcd;
    c->tuned_lpr = best_lpr;
  }
  c->nablk = (int)best_blocks.size() - 1;
  HIPCHK(hipMalloc(&c->d_ablk, sizeof(int) * best_blocks.size()));
  HIPCHK(hipMemcpyAsync(c->d_ablk, best_blocks.data(), sizeof(int) * best_blocks.size(), hipMemcpyHostToDevice,
                        c->stream));
  if (!c->seg_ptr.empty()) {
    std::vector<int> bsys(std::max(c->nablk, 1), 0);
    for (int k = 0; k < c->nablk; ++k)
      bsys[k] = (int)(std::upper_bound(c->seg_ptr.begin(), c->seg_ptr.end(), best_blocks[k]) - c->seg_ptr.begin()) - 1;
    HIPCHK(hipMalloc(&c->d_blk_sys, sizeof(int) * bsys.size()));
    HIPCHK(hipMemcpy(c->d_blk_sys, bsys.data(), sizeof(int) * bsys.size(), hipMemcpyHostToDevice));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int set_matrix_body(shesolve_ctx *c, bool device_src, int64_t n, int64_t nnz, const int64_t *rp,
                    const int32_t *ci, const double *val, int nblocks, const int64_t *ranges, int val_device);

// Stores a matrix and prepares the solver; on any failure the handle is left
// without a matrix (every solve then reports "no matrix set"), never with a
// half-built one.
int set_matrix_common(shesolve_ctx *c, bool device_src, int64_t n, int64_t nnz, const int64_t *rp,
                      const int32_t *ci, const double *val, int nblocks, const int64_t *ranges,
                      int val_device = -1) {
  if (val_device < 0) val_device = device_src ? 1 : 0;
  HIPCHK(hipSetDevice(c->device));
  free_matrix(c);
  if (n <= 0 || n > 0x7fffffff) { set_err("n out of range"); return 2; }
  const int e = set_matrix_body(c, device_src, n, nnz, rp, ci, val, nblocks, ranges, val_device);
  if (e) {
    hipStreamSynchronize(c->stream);
    free_matrix(c);
  }
  return e;
}

// Device -> host copy through the context's two pinned chunks: the copy of one
// chunk into the (pageable) destination overlaps the transfer of the next,
// where a copy straight into pageable memory moves a few GB/s.
int staged_d2h(shesolve_ctx *c, void *dst, const void *src, size_t bytes) {
  constexpr size_t kChunk = (size_t)8 << 20;
  for (int k = 0; k < 2; ++k) {
    if (!c->stage[k]) HIPCHK(hipHostMalloc(&c->stage[k], kChunk));
    if (!c->stage_ev[k]) HIPCHK(hipEventCreateWithFlags(&c->stage_ev[k], hipEventDisableTiming));
  }
  const size_t nch = (bytes + kChunk - 1) / kChunk;
  for (size_t k = 0; k <= nch; ++k) {
    if (k < nch) {
      const size_t off = k * kChunk, len = std::min(kChunk, bytes - off);
      HIPCHK(hipMemcpyAsync(c->stage[k & 1], (const char *)src + off, len, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipEventRecord(c->stage_ev[k & 1], c->stream));
    }
    if (k > 0) {
      const size_t j = k - 1, off = j * kChunk, len = std::min(kChunk, bytes - off);
      HIPCHK(hipEventSynchronize(c->stage_ev[j & 1]));
      std::memcpy((char *)dst + off, c->stage[j & 1], len);
    }
  }
  return 0;
}

int set_matrix_body(shesolve_ctx *c, bool device_src, int64_t n, int64_t nnz, const int64_t *rp,
                    const int32_t *ci, const double *val, int nblocks, const int64_t *ranges, int val_device) {
  auto t0 = std::chrono::steady_clock::now();
  std::vector<int64_t> hrp;
  if (device_src) {
    hrp.resize(n + 1);
    if (int e = staged_d2h(c, hrp.data(), rp, sizeof(int64_t) * (n + 1))) return e;
    nnz = hrp[n];
    if (hrp[0] != 0 || nnz < 0 || nnz > 0x7fffffff) { set_err("row_ptr inconsistent: nnz %lld", (long long)nnz); return 4; }
    // the column indices go straight into the analysis copy
    c->h_ci.resize(nnz);
    if (int e = staged_d2h(c, c->h_ci.data(), ci, sizeof(int32_t) * nnz)) return e;
    rp = hrp.data();
    ci = c->h_ci.data();
  }
  // SHESOLVE_VERBOSE: the duration of each set-up phase on stderr
  static const bool verbose = getenv("SHESOLVE_VERBOSE") != nullptr;
  auto tp = t0;
  auto phase = [&](const char *what) {
    if (!verbose) return;
    hipStreamSynchronize(c->stream);
    const auto t = std::chrono::steady_clock::now();
    fprintf(stderr, "[shesolve] set-up %s: %.1f ms\n", what, std::chrono::duration<double, std::milli>(t - tp).count());
    tp = t;
  };
  phase("pattern copy");
  if (int e = set_pattern_host(c, n, nnz, rp, ci, nblocks, ranges)) return e;
  phase("pattern");
  HIPCHK(hipMalloc(&c->d_rp, sizeof(int) * (n + 1)));
  HIPCHK(hipMalloc(&c->d_ci, sizeof(int) * (std::max<int64_t>(nnz, 1) + kVecPad)));
  HIPCHK(hipMalloc(&c->d_v64, sizeof(double) * (std::max<int64_t>(nnz, 1) + kVecPad)));
  HIPCHK(hipMemcpyAsync(c->d_rp, c->h_rp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d_ci, c->h_ci.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d_v64, val, sizeof(double) * nnz, val_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                        c->stream));
  c->lpr = c->prm.lanes_per_row > 0 ? c->prm.lanes_per_row : pick_lpr((double)nnz / n);
  phase("upload");
  c->tune_reuse = c->prm.autotune && c->tuned_n > 0 && c->tuned_precond == c->prm.precond && c->seg_ptr.empty() &&
                  std::llabs((long long)n - (long long)c->tuned_n) * 10 <= (long long)c->tuned_n;
  if (int e = tune_stream(c)) return e;
  phase("stream tuning");
  if (int e = analyse_and_factor(c)) return e;
  phase("analysis");
  if (int e = setup_solve_space(c)) return e;
  phase("solve space");
  if (int e = upload_values(c)) return e;
  phase("values");
  if (int e = ensure_vectors(c)) return e;
  c->last_setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

}  // namespace


// ================================================================ batched solves
struct shesolve_batch_ctx {
  shesolve_ctx *c = nullptr;       // the concatenated block-diagonal system
  int S = 0;
  std::vector<int> sys_ptr;        // row offsets of the systems
  std::vector<int64_t> val_ptr;    // value offsets of the systems
  int *d_sys_ptr = nullptr, *d_row_sys = nullptr, *d_mask = nullptr, *d_caps = nullptr;
  int *d_rsys = nullptr;  // system of each kBlock rows (-1: two systems)
  BSys *d_bs = nullptr;
  double *d_bpart = nullptr, *d_nrm = nullptr;
  double *x64 = nullptr, *b64 = nullptr, *xprev = nullptr;
  float *dd = nullptr;
  std::vector<BSys> h_bs;
  std::vector<shesolve_report> rep;
  // one iteration: [0] FP64, [1] FP32, [2] FP64 with the FP32 factor
  hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};
  bool use_graph = true;
  int poll = 8;                    // iterations between host checks of the masks
  bool verbose = getenv("SHESOLVE_VERBOSE") != nullptr;  // progress lines on stderr
  double last_setup_ms = 0;
};

namespace {

void batch_free(shesolve_batch_ctx *b) {
  if (b->c) b->c->seg_mask = nullptr;
  for (auto &g : b->graph)
    if (g) { hipGraphExecDestroy(g); g = nullptr; }
  dfree(b->d_sys_ptr); dfree(b->d_row_sys); dfree(b->d_mask); dfree(b->d_caps); dfree(b->d_bs); dfree(b->d_rsys);
  dfree(b->d_bpart); dfree(b->d_nrm); dfree(b->x64); dfree(b->b64); dfree(b->xprev); dfree(b->dd);
}

template <typename T>
void bdot(shesolve_batch_ctx *b, const int *mask, const T *a0, const T *b0, const T *a1, const T *b1) {
  shesolve_ctx *c = b->c;
  KTimer kt(c, 3);
  if (a1)
    hipLaunchKernelGGL((k_bdot<T, 2>), dim3(kBatchParts, b->S), dim3(kBlock), 0, c->stream, b->S, b->d_sys_ptr,
                       b->d_bs, mask, a0, b0, a1, b1, b->d_bpart);
  else
    hipLaunchKernelGGL((k_bdot<T, 1>), dim3(kBatchParts, b->S), dim3(kBlock), 0, c->stream, b->S, b->d_sys_ptr,
                       b->d_bs, mask, a0, b0, a1, b1, b->d_bpart);
}
void bfin(shesolve_batch_ctx *b, int op) {
  hipLaunchKernelGGL(k_bfinalize, dim3((b->S + 63) / 64), dim3(64), 0, b->c->stream, op, b->S, b->d_bpart, b->d_bs,
                     b->d_nrm);
}
template <typename T>
void bupd(shesolve_batch_ctx *b, int op, T *a, T *bb, T *cw, const T *c1, const T *c2, const T *c3, const T *c4) {
  shesolve_ctx *c = b->c;
  KTimer kt(c, 3);
  hipLaunchKernelGGL((k_bupdate<T>), dim3(kGrid), dim3(kBlock), 0, c->stream, op, c->n, b->d_row_sys, b->d_rsys,
                     b->d_bs, a, bb, cw, c1, c2, c3, c4);
}
void bref(shesolve_batch_ctx *b, int op, double *x, const double *src, float *f) {
  hipLaunchKernelGGL(k_bref, dim3(kGrid), dim3(kBlock), 0, b->c->stream, op, b->c->n, b->d_row_sys, b->d_rsys,
                     b->d_mask, x, src, f, b->d_nrm);
}
int upload_mask(shesolve_batch_ctx *b, const std::vector<int> &m) {
  HIPCHK(hipMemcpyAsync(b->d_mask, m.data(), sizeof(int) * b->S, hipMemcpyHostToDevice, b->c->stream));
  return 0;
}
// Per-system norms sqrt((u,u)) of the masked systems into d_nrm and `out`.
template <typename T>
int bnorms(shesolve_batch_ctx *b, const T *u, std::vector<double> &out) {
  bdot<T>(b, b->d_mask, u, u, (const T *)nullptr, (const T *)nullptr);
  bfin(b, 5);
  out.resize(b->S);
  HIPCHK(hipMemcpyAsync(out.data(), b->d_nrm, sizeof(double) * b->S, hipMemcpyDeviceToHost, b->c->stream));
  HIPCHK(hipStreamSynchronize(b->c->stream));
  return 0;
}

// One iteration of the batched right-preconditioned BiCGStab (kernels only,
// so that it can be captured into a graph).  m32: an FP64 iteration that
// applies the FP32 copy of the factor (FP64 vectors in and out).
template <typename TV, typename T>
void bbicg_iter(shesolve_batch_ctx *b, T *x, bool m32 = false) {
  shesolve_ctx *c = b->c;
  T **V;
  const TV *Av, *Mv;
  if constexpr (sizeof(T) == 8) { V = (T **)c->v64; Av = (const TV *)c->d_v64; Mv = (const TV *)c->d_f64; }
  else { V = (T **)c->v32; Av = (const TV *)c->d_v32; Mv = nullptr; }
  T *r = V[0], *r0 = V[1], *p = V[2], *v = V[3], *ph = V[4], *s = V[5], *sh = V[6], *t = V[7];
  // the SpMVs of the iteration skip the row blocks of converged systems
  c->spmv_skip = c->d_blk_sys ? b->d_bs : nullptr;
  bupd<T>(b, 0, p, nullptr, nullptr, r, v, nullptr, nullptr);
  if (sizeof(T) == 8 && m32) launch_precond<float, T>(c, nullptr, p, ph);
  else launch_precond<TV, T>(c, Mv, p, ph);
  launch_spmv<TV, T, T>(c, c->lpr, Av, ph, v, nullptr, nullptr, nullptr, nullptr);
  bdot<T>(b, nullptr, r0, v, (const T *)nullptr, (const T *)nullptr);
  bfin(b, 1);
  bupd<T>(b, 1, s, nullptr, nullptr, r, v, nullptr, nullptr);
  if (sizeof(T) == 8 && m32) launch_precond<float, T>(c, nullptr, s, sh);
  else launch_precond<TV, T>(c, Mv, s, sh);
  launch_spmv<TV, T, T>(c, c->lpr, Av, sh, t, nullptr, nullptr, nullptr, nullptr);
  bdot<T>(b, nullptr, t, s, t, t);
  bfin(b, 2);
  bupd<T>(b, 2, x, r, nullptr, ph, sh, s, t);
  bdot<T>(b, nullptr, r, r, r0, r);
  bfin(b, 3);
  c->spmv_skip = nullptr;
}

// Batched BiCGStab on the systems selected by mask (host vector), from the
// initial guess in x.  Per-system results in b->h_bs.
template <typename TV, typename T>
int bbicg_run(shesolve_batch_ctx *b, const T *rhs, T *x, double tol, int max_iters, double div_limit,
              const std::vector<int> &mask, const std::vector<int> *caps = nullptr) {
  shesolve_ctx *c = b->c;
  T **V;
  const TV *Av;
  if constexpr (sizeof(T) == 8) { V = (T **)c->v64; Av = (const TV *)c->d_v64; }
  else { V = (T **)c->v32; Av = (const TV *)c->d_v32; }
  T *r = V[0], *r0 = V[1], *p = V[2], *v = V[3];
  if (int e = upload_mask(b, mask)) return e;
  if (caps) HIPCHK(hipMemcpyAsync(b->d_caps, caps->data(), sizeof(int) * b->S, hipMemcpyHostToDevice, c->stream));
  // FP32 inner solves (finite divergence limit) also stop when they stall
  hipLaunchKernelGGL(k_binit, dim3((b->S + 63) / 64), dim3(64), 0, c->stream, b->S, b->d_bs, b->d_mask, tol, max_iters,
                     div_limit, div_limit < 1e300 ? kBatchStallWindow : 0, caps ? (const int *)b->d_caps : nullptr);
  launch_spmv<TV, T, T>(c, c->lpr, Av, x, r, rhs, nullptr, nullptr, nullptr);
  bupd<T>(b, 3, r0, p, v, r, nullptr, nullptr, nullptr);
  bdot<T>(b, nullptr, r, r, rhs, rhs);
  bfin(b, 0);
  HIPCHK(hipGetLastError());
  // the FP64 solve of the systems an FP32+IR batch hands over applies the FP32
  // factor: the Krylov recurrence is what has to be FP64 for them, the factor's
  // rounding is far below its ILU error, and the sweeps read half the bytes
  const bool m32 = sizeof(T) == 8 && c->prm.precision == SHESOLVE_FP32_IR;
  const int gi = sizeof(T) == 8 ? (m32 ? 2 : 0) : 1;
  const bool graph = b->use_graph && !c->prof;
  if (graph && !b->graph[gi]) {
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    bbicg_iter<TV, T>(b, x, m32);
    HIPCHK(hipStreamEndCapture(c->stream, &g));
    HIPCHK(hipGraphInstantiate(&b->graph[gi], g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
  }
  b->h_bs.resize(b->S);
  for (int done = 0; done < max_iters;) {
    HIPCHK(hipMemcpyAsync(b->h_bs.data(), b->d_bs, sizeof(BSys) * b->S, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    int nact = 0;
    for (auto &q : b->h_bs) nact += q.active != 0;
    if (!nact) break;
    if (b->verbose && done % (8 * b->poll) == 0 && done > 0)
      fprintf(stderr, "[shesolve batch]   %s iteration %d: %d systems active\n", sizeof(T) == 8 ? "FP64" : "FP32",
              done, nact);
    const int k = std::min(b->poll, max_iters - done);
    for (int i = 0; i < k; ++i) {
      if (graph) HIPCHK(hipGraphLaunch(b->graph[gi], c->stream));
      else bbicg_iter<TV, T>(b, x, m32);
    }
    done += k;
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(b->h_bs.data(), b->d_bs, sizeof(BSys) * b->S, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// True FP64 residuals ||b_s - A_s x_s|| of the masked systems (r64 = b - A x).
int bresidual(shesolve_batch_ctx *b, std::vector<double> &rn) {
  shesolve_ctx *c = b->c;
  double *r64 = c->v64[8];
  launch_spmv<double, double, double>(c, c->lpr, c->d_v64, b->x64, r64, b->b64, nullptr, nullptr, nullptr);
  return bnorms<double>(b, r64, rn);
}

int batch_solve_internal(shesolve_batch_ctx *b, int warm) {
  shesolve_ctx *c = b->c;
  const int S = b->S;
  const auto &prm = c->prm;
  b->rep.assign(S, shesolve_report{});
  std::vector<int> all(S, 1), act(S, 0), fb(S, 0);
  std::vector<double> bn, rn, prev(S, -1.0);
  // systems handed to the FP64 solve: relative residual of their best iterate,
  // which xprev keeps while that solve runs
  std::vector<double> best(S, -1.0);
  if (int e = upload_mask(b, all)) return e;
  if (int e = bnorms<double>(b, b->b64, bn)) return e;
  if (warm) {
    // systems whose initial guess is worse than zero (||b - A x0|| > ||b||) start from zero
    if (int e = bresidual(b, rn)) return e;
    std::vector<int> drop(S, 0);
    for (int s = 0; s < S; ++s) drop[s] = !(rn[s] <= bn[s]);
    if (int e = upload_mask(b, drop)) return e;
    bref(b, 3, b->x64, nullptr, nullptr);
    if (int e = upload_mask(b, all)) return e;
  } else {
    bref(b, 3, b->x64, nullptr, nullptr);
  }
  std::vector<int> zero_rhs(S, 0);
  bool anyz = false;
  for (int s = 0; s < S; ++s)
    if (bn[s] == 0.0) { zero_rhs[s] = 1; anyz = true; }
  if (anyz) {
    if (int e = upload_mask(b, zero_rhs)) return e;
    bref(b, 3, b->x64, nullptr, nullptr);
  }
  for (int s = 0; s < S; ++s) {
    b->rep[s].status = zero_rhs[s] ? 0 : 1;
    act[s] = !zero_rhs[s];
  }
  // FP64 BiCGStab on the systems in `sel` (cleared as they finish): a system
  // whose recurrence converges is confirmed with its explicit residual and
  // restarted from it when that is still above the tolerance, at most four
  // runs, within the iteration budget.  iters_base[s] counts iterations spent
  // before (the FP32 refinement of a handed-over system): the budget of the
  // FP64 runs is max_iters of their own.
  auto fp64_runs = [&](std::vector<int> &sel, const std::vector<int> &iters_base) -> int {
    for (int attempt = 0; attempt < 4; ++attempt) {
      bool any = false;
      for (int s = 0; s < S; ++s) any |= sel[s] != 0;
      if (!any) break;
      int cap = 0;
      std::vector<int> caps(S, 0);
      for (int s = 0; s < S; ++s)
        if (sel[s]) {
          caps[s] = prm.max_iters - (b->rep[s].iterations - iters_base[s]);
          cap = std::max(cap, caps[s]);
        }
      if (cap <= 0) break;
      if (int e = bbicg_run<double, double>(b, b->b64, b->x64, prm.tol, cap, 1e300, sel, &caps)) return e;
      for (int s = 0; s < S; ++s)
        if (sel[s]) {
          b->rep[s].iterations += b->h_bs[s].iters;
          b->rep[s].status = b->h_bs[s].status;
        }
      if (int e = upload_mask(b, sel)) return e;
      if (int e = bresidual(b, rn)) return e;
      for (int s = 0; s < S; ++s) {
        if (!sel[s]) continue;
        b->rep[s].rel_residual = rn[s] / bn[s];
        if (b->rep[s].rel_residual <= prm.tol) { b->rep[s].status = 0; sel[s] = 0; }
        else if (b->rep[s].status != 0 || b->rep[s].iterations - iters_base[s] >= prm.max_iters) sel[s] = 0;
        else b->rep[s].status = 1;  // recurrence converged but the true residual did not: restart
      }
    }
    return 0;
  };
  if (prm.precision == SHESOLVE_FP64) {
    if (int e = fp64_runs(act, std::vector<int>(S, 0))) return e;
  } else {
    float *r32 = c->v32[8];
    // an FP32 inner solve that stalls (iteration cap) or breaks down hands its
    // system to the FP64 solve after its correction has been checked
    std::vector<int> inner_failed(S, 0), first_it(S, -1), caps(S, 0);
    const int inner_cap = std::min(prm.max_iters, 1000);
    for (int k = 0; k <= prm.max_ir_steps; ++k) {
      if (int e = upload_mask(b, act)) return e;
      if (int e = bresidual(b, rn)) return e;
      std::vector<int> revert(S, 0);
      bool anyrev = false, any = false;
      for (int s = 0; s < S; ++s) {
        if (!act[s]) continue;
        if (prev[s] >= 0 && !(rn[s] < prev[s])) {
          // the last correction did not reduce the FP64 residual: undo it and
          // finish this system with FP64 BiCGStab
          revert[s] = 1;
          anyrev = true;
          fb[s] = 1;
          act[s] = 0;
          continue;
        }
        b->rep[s].rel_residual = rn[s] / bn[s];
        if (b->rep[s].rel_residual <= prm.tol) {
          b->rep[s].status = 0;
          act[s] = 0;
        } else if (inner_failed[s]) {
          // a stalled FP32 solve whose correction still cut the FP64 residual
          // tenfold keeps refining in FP32 (the next step restarts from the
          // new residual); otherwise the system moves to FP64
          if (rn[s] < kStallKeep * prev[s]) {
            inner_failed[s] = 0;
          } else {
            fb[s] = 1;
            act[s] = 0;
          }
        } else if (k == prm.max_ir_steps || b->rep[s].iterations >= prm.max_iters) {
          if (b->rep[s].iterations < prm.max_iters) fb[s] = 1;  // refinement steps used up
          act[s] = 0;
        }
        any |= act[s] != 0;
      }
      if (anyrev) {
        if (int e = upload_mask(b, revert)) return e;
        bref(b, 2, b->x64, b->xprev, nullptr);
      }
      if (!any) break;
      if (int e = upload_mask(b, act)) return e;
      // d_nrm holds ||r_s|| of the masked systems from bresidual
      bref(b, 2, b->xprev, b->x64, nullptr);
      bref(b, 0, nullptr, c->v64[8], r32);
      bref(b, 4, nullptr, nullptr, b->dd);
      for (int s = 0; s < S; ++s)
        if (act[s]) { prev[s] = rn[s]; b->rep[s].ir_steps++; }
      // later corrections get at most 2x the iterations the system's first
      // FP32 solve needed (at least 40): a correction that stops making
      // progress goes to FP64 early (on the 96-diode family a cap of
      // max(100, 4x) took 41.8 s against 28.7 s, the FP32 solves that stall
      // running longer before the same hand-over)
      for (int s = 0; s < S; ++s) caps[s] = first_it[s] < 0 ? inner_cap : std::max(40, 2 * first_it[s]);
      if (int e = bbicg_run<float, float>(b, r32, b->dd, prm.inner_tol, inner_cap, 1e4, act, &caps)) return e;
      for (int s = 0; s < S; ++s)
        if (act[s] && first_it[s] < 0) first_it[s] = b->h_bs[s].iters;
      int nact = 0, maxit = 0, nfail = 0;
      for (int s = 0; s < S; ++s)
        if (act[s]) {
          b->rep[s].iterations += b->h_bs[s].iters;
          if (b->h_bs[s].status != 0) inner_failed[s] = 1;
          ++nact;
          maxit = std::max(maxit, b->h_bs[s].iters);
          nfail += b->h_bs[s].status != 0;
        }
      if (b->verbose)
        fprintf(stderr, "[shesolve batch] refinement step %d: %d systems, inner iterations max %d, %d stalled\n", k,
                nact, maxit, nfail);
      // x += ||r_s|| d  (bbicg_run left d_mask = act and d_nrm untouched)
      bref(b, 1, b->x64, nullptr, b->dd);
    }
    bool anyfb = false;
    for (int s = 0; s < S; ++s) anyfb |= fb[s] != 0;
    if (anyfb && b->verbose) {
      int nfb = 0;
      for (int s = 0; s < S; ++s) nfb += fb[s];
      fprintf(stderr, "[shesolve batch] FP64 solve for %d systems\n", nfb);
    }
    if (anyfb) {
      if (int e = upload_mask(b, fb)) return e;
      bref(b, 2, b->xprev, b->x64, nullptr);
      for (int s = 0; s < S; ++s)
        if (fb[s]) best[s] = b->rep[s].rel_residual;
      std::vector<int> base(S, 0);
      for (int s = 0; s < S; ++s) base[s] = b->rep[s].iterations;
      std::vector<int> sel = fb;
      if (int e = fp64_runs(sel, base)) return e;
    }
  }
  // final explicit residuals of every system with a nonzero right-hand side
  for (int s = 0; s < S; ++s) act[s] = !zero_rhs[s];
  if (int e = upload_mask(b, act)) return e;
  if (int e = bresidual(b, rn)) return e;
  std::vector<int> restore(S, 0);
  bool anyrestore = false;
  for (int s = 0; s < S; ++s) {
    if (zero_rhs[s]) { b->rep[s].rel_residual = 0; b->rep[s].status = 0; continue; }
    b->rep[s].rel_residual = rn[s] / bn[s];
    if (b->rep[s].rel_residual <= prm.tol) b->rep[s].status = 0;
    else if (b->rep[s].status == 0) b->rep[s].status = 1;
    if (b->rep[s].status != 0 && best[s] >= 0 && !(b->rep[s].rel_residual <= best[s])) {
      // the FP64 solve ended worse than the refinement: its best iterate is returned
      restore[s] = 1;
      anyrestore = true;
      b->rep[s].rel_residual = best[s];
    }
  }
  if (anyrestore) {
    if (int e = upload_mask(b, restore)) return e;
    bref(b, 2, b->x64, b->xprev, nullptr);
  }
  return 0;
}

int batch_set_systems(shesolve_batch_ctx *b, bool device_src, int count, const int64_t *n, const int64_t *const *rp,
                      const int32_t *const *ci, const double *const *val, const int *nblocks,
                      const int64_t *const *ranges) {
  auto t0 = std::chrono::steady_clock::now();
  shesolve_ctx *c = b->c;
  HIPCHK(hipSetDevice(c->device));
  if (count <= 0) { set_err("count must be positive"); return 2; }
  if (!n) { set_err("n is NULL"); return 3; }
  if (c->prm.precond == SHESOLVE_PRECOND_MC_ILU0 || c->prm.precond == SHESOLVE_PRECOND_BMC_ILU0) {
    set_err("the multicolour orderings are not available for batches (use ILU0, BLOCK_ILU0 or ILUK)");
    return -3;
  }
  if (!rp) { set_err("row_ptr is NULL"); return 4; }
  if (!ci) { set_err("col is NULL"); return 5; }
  if (!val) { set_err("val is NULL"); return 6; }
  batch_free(b);
  int64_t N = 0, NNZ = 0;
  std::vector<int64_t> nnz(count);
  std::vector<std::vector<int64_t>> hrp(count);
  for (int s = 0; s < count; ++s) {
    if (n[s] <= 0 || !rp[s] || !ci[s] || !val[s]) {
      set_err("system %d: empty or missing arrays", s);
      return 3;
    }
    hrp[s].resize(n[s] + 1);
    if (device_src) {
      if (int e = staged_d2h(c, hrp[s].data(), rp[s], sizeof(int64_t) * (n[s] + 1))) return e;
    } else {
      std::copy(rp[s], rp[s] + n[s] + 1, hrp[s].begin());
    }
    nnz[s] = hrp[s][n[s]] - hrp[s][0];
    if (hrp[s][0] != 0 || nnz[s] < 0) { set_err("system %d: row_ptr must start at 0", s); return 4; }
    N += n[s];
    NNZ += nnz[s];
  }
  if (N > 0x7fffffff || NNZ > 0x7fffffff) {
    set_err("batch too large for 32-bit indices: %lld rows, %lld entries", (long long)N, (long long)NNZ);
    return 2;
  }
  // the concatenated pattern: each system's columns straight into place (the
  // columns then shifted and checked per system on host threads), the column
  // array becoming the analysis copy without another copy
  hvec<int64_t> crp(N + 1);
  hvec<int32_t> cci(NNZ);
  std::vector<int64_t> cranges;
  b->S = count;
  b->sys_ptr.assign(count + 1, 0);
  b->val_ptr.assign(count + 1, 0);
  for (int s = 0; s < count; ++s) {
    b->sys_ptr[s + 1] = (int)(b->sys_ptr[s] + n[s]);
    b->val_ptr[s + 1] = b->val_ptr[s] + nnz[s];
    if (device_src) {
      if (int e = staged_d2h(c, cci.data() + b->val_ptr[s], ci[s], sizeof(int32_t) * nnz[s])) return e;
    }
    const int64_t r0 = b->sys_ptr[s];
    if (nblocks && ranges && nblocks[s] > 0 && ranges[s]) {
      for (int q = 0; q < nblocks[s]; ++q) {
        cranges.push_back(r0 + std::max<int64_t>(0, std::min<int64_t>(n[s], ranges[s][2 * q])));
        cranges.push_back(r0 + std::max<int64_t>(0, std::min<int64_t>(n[s], ranges[s][2 * q + 1])));
      }
    } else {
      cranges.push_back(r0);
      cranges.push_back(r0 + n[s]);
    }
  }
  std::atomic<int> bad_sys{-1};
  host_parallel(count, [&](int64_t s0, int64_t s1) {
    for (int s = (int)s0; s < (int)s1; ++s) {
      const int64_t r0 = b->sys_ptr[s], v0 = b->val_ptr[s];
      for (int64_t i = 0; i <= n[s]; ++i) crp[r0 + i] = v0 + hrp[s][i];
      const int32_t *src = device_src ? cci.data() + v0 : ci[s];
      for (int64_t k = 0; k < nnz[s]; ++k) {
        const int32_t j = src[k];
        if (j < 0 || j >= n[s]) bad_sys = s;
        cci[v0 + k] = (int32_t)(j + r0);
      }
    }
  }, 1);
  if (bad_sys >= 0) { set_err("system %d: column index out of range", bad_sys.load()); return 5; }
  // values: one concatenated buffer (device staging for device sources)
  double *dval = nullptr;
  std::vector<double> hval;
  if (device_src) {
    HIPCHK(hipMalloc(&dval, sizeof(double) * std::max<int64_t>(NNZ, 1)));
    for (int s = 0; s < count; ++s)
      HIPCHK(hipMemcpyAsync(dval + b->val_ptr[s], val[s], sizeof(double) * nnz[s], hipMemcpyDeviceToDevice,
                            c->stream));
  } else {
    hval.resize(NNZ);
    for (int s = 0; s < count; ++s) std::copy(val[s], val[s] + nnz[s], hval.begin() + b->val_ptr[s]);
  }
  c->seg_ptr = b->sys_ptr;
  if (!c->seg_fill.empty() && (int)c->seg_fill.size() != count) {
    set_err("%d fill levels set for %d systems", (int)c->seg_fill.size(), count);
    return 2;
  }
  c->h_ci.swap(cci);
  int e = set_matrix_common(c, false, N, NNZ, crp.data(), c->h_ci.data(), device_src ? dval : hval.data(),
                            (int)(cranges.size() / 2), cranges.data(), device_src ? 1 : 0);
  if (dval) hipFree(dval);
  if (e) return e;
  std::vector<int> row_sys(N);
  host_parallel(count, [&](int64_t s0, int64_t s1) {
    for (int s = (int)s0; s < (int)s1; ++s) std::fill(row_sys.begin() + b->sys_ptr[s], row_sys.begin() + b->sys_ptr[s + 1], s);
  }, 1);
  HIPCHK(hipMalloc(&b->d_sys_ptr, sizeof(int) * (count + 1)));
  HIPCHK(hipMalloc(&b->d_row_sys, sizeof(int) * N));
  HIPCHK(hipMalloc(&b->d_mask, sizeof(int) * count));
  HIPCHK(hipMalloc(&b->d_caps, sizeof(int) * count));
  HIPCHK(hipMalloc(&b->d_bs, sizeof(BSys) * count));
  HIPCHK(hipMalloc(&b->d_bpart, sizeof(double) * 2 * count * kBatchParts));
  HIPCHK(hipMalloc(&b->d_nrm, sizeof(double) * count));
  HIPCHK(hipMalloc(&b->x64, sizeof(double) * N));
  HIPCHK(hipMalloc(&b->b64, sizeof(double) * N));
  HIPCHK(hipMalloc(&b->xprev, sizeof(double) * N));
  HIPCHK(hipMalloc(&b->dd, sizeof(float) * N));
  HIPCHK(hipMemcpy(b->d_sys_ptr, b->sys_ptr.data(), sizeof(int) * (count + 1), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->d_row_sys, row_sys.data(), sizeof(int) * N, hipMemcpyHostToDevice));
  {
    std::vector<int> rs((N + kBlock - 1) / kBlock);
    for (size_t blk = 0; blk < rs.size(); ++blk) {
      const int i0 = (int)blk * kBlock, i1 = (int)std::min<int64_t>(N, i0 + kBlock);
      rs[blk] = row_sys[i0] == row_sys[i1 - 1] ? row_sys[i0] : -1;
    }
    HIPCHK(hipMalloc(&b->d_rsys, sizeof(int) * rs.size()));
    HIPCHK(hipMemcpy(b->d_rsys, rs.data(), sizeof(int) * rs.size(), hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemset(b->d_bs, 0, sizeof(BSys) * count));
  c->seg_mask = b->d_bs;
  b->last_setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

int batch_update_values(shesolve_batch_ctx *b, bool device_src, const double *const *val) {
  shesolve_ctx *c = b->c;
  if (!c->d_rp) { set_err("no systems set"); return -1; }
  if (!val) { set_err("val is NULL"); return 2; }
  HIPCHK(hipSetDevice(c->device));
  for (int s = 0; s < b->S; ++s) {
    if (!val[s]) { set_err("system %d: values missing", s); return 2; }
    HIPCHK(hipMemcpyAsync(c->d_v64 + b->val_ptr[s], val[s], sizeof(double) * (b->val_ptr[s + 1] - b->val_ptr[s]),
                          device_src ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
  }
  return upload_values(c);
}

int batch_solve(shesolve_batch_ctx *b, bool device_src, const double *const *rhs, double *const *x, int warm,
                shesolve_report *reports) {
  auto t0 = std::chrono::steady_clock::now();
  shesolve_ctx *c = b->c;
  if (!c->d_rp) { set_err("no systems set"); return -1; }
  if (!rhs) { set_err("b is NULL"); return 2; }
  if (!x) { set_err("x is NULL"); return 3; }
  HIPCHK(hipSetDevice(c->device));
  const hipMemcpyKind in = device_src ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  const hipMemcpyKind out = device_src ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  for (int s = 0; s < b->S; ++s) {
    if (!rhs[s] || !x[s]) { set_err("system %d: b or x missing", s); return 2; }
    const size_t bytes = sizeof(double) * (b->sys_ptr[s + 1] - b->sys_ptr[s]);
    HIPCHK(hipMemcpyAsync(b->b64 + b->sys_ptr[s], rhs[s], bytes, in, c->stream));
    if (warm) HIPCHK(hipMemcpyAsync(b->x64 + b->sys_ptr[s], x[s], bytes, in, c->stream));
  }
  if (!device_src) {
    // reject non-finite right-hand sides (src/solvers/solve.hpp:55-59)
    for (int s = 0; s < b->S; ++s)
      for (int i = 0; i < b->sys_ptr[s + 1] - b->sys_ptr[s]; ++i)
        if (!std::isfinite(rhs[s][i])) { set_err("system %d: right-hand side is not finite", s); return 2; }
  }
  if (int e = batch_solve_internal(b, warm)) return e;
  for (int s = 0; s < b->S; ++s) {
    const size_t bytes = sizeof(double) * (b->sys_ptr[s + 1] - b->sys_ptr[s]);
    HIPCHK(hipMemcpyAsync(x[s], b->x64 + b->sys_ptr[s], bytes, out, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (reports)
    for (int s = 0; s < b->S; ++s) {
      reports[s] = b->rep[s];
      reports[s].setup_ms = b->last_setup_ms;
      reports[s].solve_ms = ms;
    }
  return 0;
}

}  // namespace

// ================================================================ C ABI
extern "C" {

void shesolve_default_params(shesolve_params *p) {
  if (!p) return;
  p->tol = 1e-10;
  p->max_iters = 1000;
  p->precond = SHESOLVE_PRECOND_ILU0;
  p->precision = SHESOLVE_FP64;
  p->inner_tol = 1e-2;
  p->max_ir_steps = 30;
  p->lanes_per_row = 0;
  p->fill_level = 1;
  p->drop_tol = 0.0;
  p->autotune = 1;
}

const char *shesolve_last_error(void) { return g_err.c_str(); }

static int check_params(const shesolve_params *p) {
  if (!(p->tol > 0)) { set_err("tol must be positive"); return 1; }
  if (p->max_iters <= 0) { set_err("max_iters must be positive"); return 1; }
  if (p->precond != SHESOLVE_PRECOND_NONE && p->precond != SHESOLVE_PRECOND_ILU0 &&
      p->precond != SHESOLVE_PRECOND_BLOCK_ILU0 && p->precond != SHESOLVE_PRECOND_MC_ILU0 &&
      p->precond != SHESOLVE_PRECOND_BMC_ILU0 && p->precond != SHESOLVE_PRECOND_ILUK &&
      p->precond != SHESOLVE_PRECOND_BLOCK_ILUK) {
    set_err("unknown preconditioner %d", p->precond);
    return 1;
  }
  if (p->precision != SHESOLVE_FP64 && p->precision != SHESOLVE_FP32_IR) { set_err("unknown precision"); return 1; }
  if (p->fill_level < 0 || p->fill_level > 8) { set_err("fill_level must be in [0, 8]"); return 1; }
  if (!(p->drop_tol >= 0) || p->drop_tol >= 1) { set_err("drop_tol must be in [0, 1)"); return 1; }
  if (p->lanes_per_row != 0 && (p->lanes_per_row < 2 || p->lanes_per_row > 64 ||
                                (p->lanes_per_row & (p->lanes_per_row - 1)))) {
    set_err("lanes_per_row must be 0 or a power of two in [2, 64]");
    return 1;
  }
  return 0;
}

int shesolve_create(shesolve_handle *h, int device, const shesolve_params *p) {
  if (!h) { set_err("handle pointer is NULL"); return 1; }
  int cnt = 0;
  if (hipGetDeviceCount(&cnt) != hipSuccess || cnt == 0) { set_err("no HIP device available"); return -2; }
  if (device < 0 || device >= cnt) { set_err("device %d out of range", device); return 2; }
  shesolve_params prm;
  shesolve_default_params(&prm);
  if (p) {
    if (check_params(p)) return 3;
    prm = *p;
  }
  HIPCHK(hipSetDevice(device));
  shesolve_ctx *c = new shesolve_ctx();
  c->device = device;
  c->prm = prm;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    set_err("stream/event creation failed");
    delete c;
    return -10;
  }
  *h = c;
  return 0;
}

int shesolve_destroy(shesolve_handle h) {
  if (!h) return 1;
  hipSetDevice(h->device);
  hipStreamSynchronize(h->stream);
  free_matrix(h);
  for (auto &p : h->v64) dfree(p);
  for (auto &p : h->v32) dfree(p);
  dfree(h->d_part);
  dfree(h->d_sc);
  dfree(h->d_norm);
  if (h->h_ring) hipHostFree(h->h_ring);
  for (auto &e : h->iter_ev)
    if (e) hipEventDestroy(e);
  hipEventDestroy(h->ev0);
  hipEventDestroy(h->ev1);
  for (int k = 0; k < 2; ++k) {
    if (h->stage[k]) hipHostFree(h->stage[k]);
    if (h->stage_ev[k]) hipEventDestroy(h->stage_ev[k]);
  }
  for (auto &e : h->evpool) hipEventDestroy(e);
  hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

int shesolve_set_params(shesolve_handle h, const shesolve_params *p) {
  if (!h) return 1;
  if (!p) return 2;
  if (check_params(p)) return 2;
  bool refactor_needed = h->d_rp && p->precond != h->prm.precond;
  h->prm = *p;
  if (h->d_rp && p->lanes_per_row > 0) h->lpr = p->lanes_per_row;
  if (refactor_needed) {
    set_err("changing the preconditioner needs a new shesolve_set_matrix");
    return -3;
  }
  return 0;
}

int shesolve_set_matrix(shesolve_handle h, int64_t n, int64_t nnz, const int64_t *row_ptr, const int32_t *col,
                        const double *val, int nblocks, const int64_t *block_ranges) {
  if (!h) return 1;
  if (!row_ptr) return 4;
  if (!col && nnz > 0) return 5;
  if (!val && nnz > 0) return 6;
  return set_matrix_common(h, false, n, nnz, row_ptr, col, val, nblocks, block_ranges);
}

int shesolve_set_matrix_device(shesolve_handle h, int64_t n, int64_t nnz, const int64_t *d_row_ptr,
                               const int32_t *d_col, const double *d_val, int nblocks, const int64_t *block_ranges) {
  if (!h) return 1;
  if (!d_row_ptr) return 4;
  if (!d_col) return 5;
  if (!d_val) return 6;
  return set_matrix_common(h, true, n, nnz, d_row_ptr, d_col, d_val, nblocks, block_ranges);
}

int shesolve_update_values(shesolve_handle h, const double *val) {
  if (!h) return 1;
  if (!val) return 2;
  if (!h->d_v64) { set_err("no matrix set"); return -1; }
  auto t0 = std::chrono::steady_clock::now();
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipMemcpyAsync(h->d_v64, val, sizeof(double) * h->nnz, hipMemcpyHostToDevice, h->stream));
  int e = upload_values(h);
  h->last_setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return e;
}

int shesolve_update_values_device(shesolve_handle h, const double *d_val) {
  if (!h) return 1;
  if (!d_val) return 2;
  if (!h->d_v64) { set_err("no matrix set"); return -1; }
  auto t0 = std::chrono::steady_clock::now();
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipMemcpyAsync(h->d_v64, d_val, sizeof(double) * h->nnz, hipMemcpyDeviceToDevice, h->stream));
  int e = upload_values(h);
  h->last_setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return e;
}

int shesolve_solve(shesolve_handle h, const double *b, double *x, int warm_start, shesolve_report *report) {
  if (!h) return 1;
  if (!b) return 2;
  if (!x) return 3;
  if (!h->d_rp) { set_err("no matrix set"); return -1; }
  HIPCHK(hipSetDevice(h->device));
  for (int i = 0; i < h->n; ++i)
    if (!std::isfinite(b[i])) {
      // check_vector_for_valid_entries (util/checks.hpp:110-130) throws on NaN/Inf
      set_err("right-hand side holds a non-finite value at %d", i);
      return 2;
    }
  double *db = nullptr, *dx = nullptr;
  HIPCHK(hipMalloc(&db, sizeof(double) * h->n));
  HIPCHK(hipMalloc(&dx, sizeof(double) * h->n));
  HIPCHK(hipMemcpyAsync(db, b, sizeof(double) * h->n, hipMemcpyHostToDevice, h->stream));
  if (warm_start) HIPCHK(hipMemcpyAsync(dx, x, sizeof(double) * h->n, hipMemcpyHostToDevice, h->stream));
  int e = solve_device(h, db, dx, warm_start, report);
  if (e == 0) {
    hipMemcpyAsync(x, dx, sizeof(double) * h->n, hipMemcpyDeviceToHost, h->stream);
    if (hipStreamSynchronize(h->stream) != hipSuccess) { set_err("copy back failed"); e = -10; }
  }
  hipFree(db);
  hipFree(dx);
  return e;
}

int shesolve_solve_device(shesolve_handle h, const double *d_b, double *d_x, int warm_start, shesolve_report *report) {
  if (!h) return 1;
  if (!d_b) return 2;
  if (!d_x) return 3;
  HIPCHK(hipSetDevice(h->device));
  return solve_device(h, d_b, d_x, warm_start, report);
}

int shesolve_spmv(shesolve_handle h, const double *x, double *y) {
  if (!h) return 1;
  if (!x) return 2;
  if (!y) return 3;
  if (!h->d_rp) { set_err("no matrix set"); return -1; }
  HIPCHK(hipSetDevice(h->device));
  if (int e = ensure_vectors(h)) return e;
  HIPCHK(hipMemcpyAsync(h->v64[0], x, sizeof(double) * h->n, hipMemcpyHostToDevice, h->stream));
  if (h->solve_perm) {
    const int g = grid_for(h->n, kBlock);
    hipLaunchKernelGGL((k_permute<double>), dim3(g), dim3(kBlock), 0, h->stream, h->n, h->d_perm, h->v64[0], h->v64[2], 0);
    launch_spmv<double, double, double>(h, h->lpr, sv64(h), h->v64[2], h->v64[3], nullptr, nullptr, nullptr, nullptr);
    hipLaunchKernelGGL((k_permute<double>), dim3(g), dim3(kBlock), 0, h->stream, h->n, h->d_perm, h->v64[3], h->v64[1], 1);
  } else {
    launch_spmv<double, double, double>(h, h->lpr, sv64(h), h->v64[0], h->v64[1], nullptr, nullptr, nullptr, nullptr);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(y, h->v64[1], sizeof(double) * h->n, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int shesolve_spmv_f32(shesolve_handle h, const float *x, float *y) {
  if (!h) return 1;
  if (!x) return 2;
  if (!y) return 3;
  if (!h->d_rp) { set_err("no matrix set"); return -1; }
  HIPCHK(hipSetDevice(h->device));
  if (int e = ensure_vectors(h)) return e;
  HIPCHK(hipMemcpyAsync(h->v32[0], x, sizeof(float) * h->n, hipMemcpyHostToDevice, h->stream));
  if (h->solve_perm) {
    const int g = grid_for(h->n, kBlock);
    hipLaunchKernelGGL((k_permute<float>), dim3(g), dim3(kBlock), 0, h->stream, h->n, h->d_perm, h->v32[0], h->v32[2], 0);
    launch_spmv<float, float, float>(h, h->lpr, sv32(h), h->v32[2], h->v32[3], nullptr, nullptr, nullptr, nullptr);
    hipLaunchKernelGGL((k_permute<float>), dim3(g), dim3(kBlock), 0, h->stream, h->n, h->d_perm, h->v32[3], h->v32[1], 1);
  } else {
    launch_spmv<float, float, float>(h, h->lpr, sv32(h), h->v32[0], h->v32[1], nullptr, nullptr, nullptr, nullptr);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(y, h->v32[1], sizeof(float) * h->n, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int shesolve_precond_apply(shesolve_handle h, const double *rr, double *z) {
  if (!h) return 1;
  if (!rr) return 2;
  if (!z) return 3;
  if (!h->d_rp) { set_err("no matrix set"); return -1; }
  HIPCHK(hipSetDevice(h->device));
  if (int e = ensure_vectors(h)) return e;
  HIPCHK(hipMemcpyAsync(h->v64[0], rr, sizeof(double) * h->n, hipMemcpyHostToDevice, h->stream));
  // every segment of a batch takes part, whatever its convergence mask says
  const BSys *mask = h->seg_mask;
  h->seg_mask = nullptr;
  if (h->solve_perm) {
    const int g = grid_for(h->n, kBlock);
    hipLaunchKernelGGL((k_permute<double>), dim3(g), dim3(kBlock), 0, h->stream, h->n, h->d_perm, h->v64[0], h->v64[2], 0);
    launch_precond<double, double>(h, h->d_f64, h->v64[2], h->v64[3]);
    hipLaunchKernelGGL((k_permute<double>), dim3(g), dim3(kBlock), 0, h->stream, h->n, h->d_perm, h->v64[3], h->v64[1], 1);
  } else {
    launch_precond<double, double>(h, h->d_f64, h->v64[0], h->v64[1]);
  }
  h->seg_mask = mask;
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(z, h->v64[1], sizeof(double) * h->n, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int64_t shesolve_factor_nnz(shesolve_handle h) { return h ? h->fnnz : -1; }

int shesolve_get_factors(shesolve_handle h, int64_t *row_ptr, int32_t *col, double *val) {
  if (!h) return 1;
  if (!row_ptr) return 2;
  if (!col) return 3;
  if (!val) return 4;
  if (!h->d_f64) { set_err("no factorisation"); return -1; }
  HIPCHK(hipSetDevice(h->device));
  for (int i = 0; i <= h->n; ++i) row_ptr[i] = h->h_frp[i];
  std::memcpy(col, h->h_fci.data(), sizeof(int32_t) * h->fnnz);
  HIPCHK(hipMemcpyAsync(val, h->d_f64, sizeof(double) * h->fnnz, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int shesolve_stream_config(shesolve_handle h, int32_t *items, int32_t *xcd, double *tune_us9) {
  if (!h) return 1;
  if (items) *items = h->spmv_lpr ? -h->spmv_lpr : h->items;
  if (xcd) *xcd = h->xcd;
  if (tune_us9)
    for (int k = 0; k < 9; ++k) tune_us9[k] = h->tune_us[k];
  return 0;
}

int shesolve_get_ordering(shesolve_handle h, int32_t *perm) {
  if (!h) return 1;
  if (!perm) return 2;
  for (int i = 0; i < h->n; ++i) perm[i] = h->h_perm.empty() ? i : h->h_perm[i];
  return 0;
}

int shesolve_levels(shesolve_handle h, int64_t *out2) {
  if (!h) return 1;
  if (!out2) return 2;
  out2[0] = h->lvL.count();
  out2[1] = h->lvU.count();
  return 0;
}

int shesolve_set_reference_solution(shesolve_handle h, int64_t n, const double *x_ref, int on_device) {
  if (!h) { set_err("handle is NULL"); return 1; }
  HIPCHK(hipSetDevice(h->device));
  if (!x_ref) {
    dfree(h->d_xref);
    dfree(h->d_xref_s);
    h->xref_cur = nullptr;
    return 0;
  }
  if (!h->d_rp) { set_err("no matrix set"); return -1; }
  if (n != h->n) { set_err("reference solution has %lld values, the matrix %d rows", (long long)n, h->n); return 2; }
  if (!h->d_xref) HIPCHK(hipMalloc(&h->d_xref, sizeof(double) * h->n));
  HIPCHK(hipMemcpyAsync(h->d_xref, x_ref, sizeof(double) * h->n,
                        on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int shesolve_ir_history(shesolve_handle h, int32_t cap, shesolve_ir_step *out, int32_t *count) {
  if (!h) { set_err("handle is NULL"); return 1; }
  if (cap < 0) { set_err("cap must not be negative"); return 2; }
  if (cap > 0 && !out) { set_err("out is NULL"); return 3; }
  const int m = std::min<int>(cap, (int)h->ir_hist.size());
  for (int i = 0; i < m; ++i) out[i] = h->ir_hist[i];
  if (count) *count = (int32_t)h->ir_hist.size();
  return 0;
}

int shesolve_set_profiling(shesolve_handle h, int on) {
  if (!h) return 1;
  if (on < 0 || on > 2) { set_err("profiling mode must be 0, 1 or 2"); return 2; }
  resolve_events(h);
  h->prof = on;
  for (int k = 0; k < 5; ++k) { h->ktime[k] = 0; h->kcount[k] = 0; }
  if (on == 2 && h->evpool.empty()) {
    h->evpool.resize(2 * 4096);
    h->evkind.resize(4096);
    for (auto &e : h->evpool)
      if (hipEventCreate(&e) != hipSuccess) { set_err("event creation failed"); return -10; }
  }
  return 0;
}

int shesolve_kernel_stats(shesolve_handle h, int kind, double *avg_us, int64_t *launches) {
  if (!h) return 1;
  if (kind < 0 || kind > 4) return 2;
  resolve_events(h);
  if (avg_us) *avg_us = h->kcount[kind] ? h->ktime[kind] / h->kcount[kind] : 0.0;
  if (launches) *launches = h->kcount[kind];
  return 0;
}

int shesolve_spmv_bytes(shesolve_handle h, int fp32, int64_t *bytes) {
  if (!h) return 1;
  if (!bytes) return 3;
  if (!h->d_rp) { set_err("no matrix set"); return -1; }
  const int64_t nz = h->solve_perm ? h->fnnz : h->nnz, sv = fp32 ? 4 : 8;
  const bool coded = h->d_cdesc && h->s_pairs && h->code_src == h->s_pairs && h->code_nblk == h->s_nblk;
  *bytes = nz * (sv + (coded ? 2 : 4)) + (coded ? h->ntab * 4 : 0) + (int64_t)(h->n + 1) * 4 + 2 * (int64_t)h->n * sv;
  return 0;
}

int shesolve_time_spmv(shesolve_handle h, int fp32, int reps, double *avg_us) {
  if (!h) return 1;
  if (reps <= 0) return 3;
  if (!avg_us) return 4;
  if (!h->d_rp) { set_err("no matrix set"); return -1; }
  HIPCHK(hipSetDevice(h->device));
  if (int e = ensure_vectors(h)) return e;
  resolve_events(h);
  const int was = h->prof;
  h->prof = 0;
  if (fp32) {
    hipLaunchKernelGGL((k_fill<float>), dim3(kGrid), dim3(kBlock), 0, h->stream, (int64_t)h->n, h->v32[0], 1.0f);
    launch_spmv<float, float, float>(h, h->lpr, sv32(h), h->v32[0], h->v32[1], nullptr, nullptr, nullptr, nullptr);
  } else {
    hipLaunchKernelGGL((k_fill<double>), dim3(kGrid), dim3(kBlock), 0, h->stream, (int64_t)h->n, h->v64[0], 1.0);
    launch_spmv<double, double, double>(h, h->lpr, sv64(h), h->v64[0], h->v64[1], nullptr, nullptr, nullptr,
                                        nullptr);
  }
  HIPCHK(hipEventRecord(h->ev0, h->stream));
  for (int r = 0; r < reps; ++r) {
    if (fp32)
      launch_spmv<float, float, float>(h, h->lpr, sv32(h), h->v32[0], h->v32[1], nullptr, nullptr, nullptr, nullptr);
    else
      launch_spmv<double, double, double>(h, h->lpr, sv64(h), h->v64[0], h->v64[1], nullptr, nullptr, nullptr,
                                          nullptr);
  }
  HIPCHK(hipEventRecord(h->ev1, h->stream));
  HIPCHK(hipEventSynchronize(h->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  *avg_us = ms * 1e3 / reps;
  h->prof = was;
  return 0;
}

// ---------------------------------------------------------------- batched solves
int shesolve_batch_create(shesolve_batch *h, int device, const shesolve_params *p) {
  if (!h) { set_err("handle pointer is NULL"); return 1; }
  shesolve_handle c = nullptr;
  if (int e = shesolve_create(&c, device, p)) return e;
  shesolve_batch_ctx *b = new shesolve_batch_ctx();
  b->c = c;
  *h = b;
  return 0;
}
int shesolve_batch_destroy(shesolve_batch h) {
  if (!h) return 1;
  hipSetDevice(h->c->device);
  hipStreamSynchronize(h->c->stream);
  batch_free(h);
  shesolve_destroy(h->c);
  delete h;
  return 0;
}
int shesolve_batch_set_params(shesolve_batch h, const shesolve_params *p) {
  if (!h) { set_err("handle is NULL"); return 1; }
  if (p && p->precond != h->c->prm.precond && h->c->d_rp) {
    set_err("the preconditioner kind is fixed once systems are set");
    return 2;
  }
  for (auto &g : h->graph)
    if (g) { hipGraphExecDestroy(g); g = nullptr; }
  return shesolve_set_params(h->c, p);
}
int shesolve_batch_set_systems(shesolve_batch h, int count, const int64_t *n, const int64_t *const *row_ptr,
                               const int32_t *const *col, const double *const *val, const int *nblocks,
                               const int64_t *const *block_ranges) {
  if (!h) { set_err("handle is NULL"); return 1; }
  return batch_set_systems(h, false, count, n, row_ptr, col, val, nblocks, block_ranges);
}
int shesolve_batch_set_systems_device(shesolve_batch h, int count, const int64_t *n, const int64_t *const *d_row_ptr,
                                      const int32_t *const *d_col, const double *const *d_val, const int *nblocks,
                                      const int64_t *const *block_ranges) {
  if (!h) { set_err("handle is NULL"); return 1; }
  return batch_set_systems(h, true, count, n, d_row_ptr, d_col, d_val, nblocks, block_ranges);
}
int shesolve_batch_update_values(shesolve_batch h, const double *const *val) {
  if (!h) { set_err("handle is NULL"); return 1; }
  return batch_update_values(h, false, val);
}
int shesolve_batch_update_values_device(shesolve_batch h, const double *const *d_val) {
  if (!h) { set_err("handle is NULL"); return 1; }
  return batch_update_values(h, true, d_val);
}
int shesolve_batch_solve(shesolve_batch h, const double *const *b, double *const *x, int warm_start,
                         shesolve_report *reports) {
  if (!h) { set_err("handle is NULL"); return 1; }
  return batch_solve(h, false, b, x, warm_start, reports);
}
int shesolve_batch_solve_device(shesolve_batch h, const double *const *d_b, double *const *d_x, int warm_start,
                                shesolve_report *reports) {
  if (!h) { set_err("handle is NULL"); return 1; }
  return batch_solve(h, true, d_b, d_x, warm_start, reports);
}
int shesolve_batch_set_fill_levels(shesolve_batch h, int count, const int *levels) {
  if (!h) { set_err("handle is NULL"); return 1; }
  if (count < 0) { set_err("count must be >= 0"); return 2; }
  if (count > 0 && !levels) { set_err("levels is NULL"); return 3; }
  for (int s = 0; s < count; ++s)
    if (levels[s] < 0 || levels[s] > 8) { set_err("fill level of system %d must be in [0, 8]", s); return 3; }
  h->c->seg_fill.assign(levels, levels + count);
  return 0;
}
int shesolve_batch_options(shesolve_batch h, int use_graph, int poll) {
  if (!h) { set_err("handle is NULL"); return 1; }
  if (poll < 1) { set_err("poll must be >= 1"); return 3; }
  h->use_graph = use_graph != 0;
  h->poll = poll;
  return 0;
}
int shesolve_batch_info(shesolve_batch h, int64_t *out6) {
  if (!h) { set_err("handle is NULL"); return 1; }
  if (!out6) { set_err("out is NULL"); return 2; }
  out6[0] = h->S;
  out6[1] = h->c->n;
  out6[2] = h->c->nnz;
  out6[3] = h->c->seg_levels_max;
  out6[4] = h->c->lvL.count();
  out6[5] = h->c->fnnz;
  return 0;
}
shesolve_handle shesolve_batch_solver(shesolve_batch h) { return h ? h->c : nullptr; }

}  // extern "C"

