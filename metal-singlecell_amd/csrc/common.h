// Shared device helpers and the C-ABI error state of libmsc_hip.
//
// Everything here is written for gfx950 (CDNA4): wave64, 64-bit lane masks, cross-lane moves
// by DPP where a pattern has one and through the LDS crossbar otherwise.  The library is built with -ffp-contract=off so every float operation
// in a kernel is separately rounded unless a builtin asks otherwise; the exact-distance and
// fixed-point parts of the neighbor-graph path rely on that.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace msc {

constexpr int kWave = 64;

// ---- error state (thread-local message, int status codes) --------------------------------
enum Status : int {
  kOk = 0,
  kInvalidArgument = 1,
  kHipError = 2,
  kWorkspaceTooSmall = 3,
  kUnsupported = 4,
  kInternal = 5,
};

void set_error(const char* fmt, ...);
const char* last_error();

#define MSC_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      ::msc::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
      return ::msc::kHipError;                                                          \
    }                                                                                   \
  } while (0)

#define MSC_LAUNCH_CHECK()                                                              \
  do {                                                                                  \
    hipError_t e_ = hipGetLastError();                                                  \
    if (e_ != hipSuccess) {                                                             \
      ::msc::set_error("%s:%d kernel launch: %s", __FILE__, __LINE__, hipGetErrorString(e_)); \
      return ::msc::kHipError;                                                          \
    }                                                                                   \
  } while (0)

#define MSC_REQUIRE(cond, ...)                 \
  do {                                         \
    if (!(cond)) {                             \
      ::msc::set_error(__VA_ARGS__);           \
      return ::msc::kInvalidArgument;          \
    }                                          \
  } while (0)

#define MSC_TRY(expr)              \
  do {                             \
    int rc_ = (expr);              \
    if (rc_ != ::msc::kOk) return rc_; \
  } while (0)

// ---- workspace carving -------------------------------------------------------------------
// A bump allocator over one caller-provided device buffer; every sub-buffer 256-B aligned.
struct Carve {
  char* base;
  size_t cap;
  size_t used = 0;
  bool sizing;  // base == nullptr: only count bytes
  Carve(void* p, size_t c) : base(static_cast<char*>(p)), cap(c), sizing(p == nullptr) {}
  template <typename T>
  T* take(size_t count) {
    size_t off = (used + 255) & ~size_t(255);
    used = off + count * sizeof(T);
    if (sizing) return nullptr;
    return used <= cap ? reinterpret_cast<T*>(base + off) : nullptr;
  }
  bool ok() const { return sizing || used <= cap; }
};

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---- device helpers ----------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t x, int m) {
  uint32_t lo = static_cast<uint32_t>(x), hi = static_cast<uint32_t>(x >> 32);
  lo = __shfl_xor(lo, m, kWave);
  hi = __shfl_xor(hi, m, kWave);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ uint64_t shfl_u64(uint64_t x, int src) {
  uint32_t lo = static_cast<uint32_t>(x), hi = static_cast<uint32_t>(x >> 32);
  lo = __shfl(lo, src, kWave);
  hi = __shfl(hi, src, kWave);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ uint64_t shfl_up_u64(uint64_t x, int d) {
  uint32_t lo = static_cast<uint32_t>(x), hi = static_cast<uint32_t>(x >> 32);
  lo = __shfl_up(lo, d, kWave);
  hi = __shfl_up(hi, d, kWave);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Value of lane l ^ J for a compile-time J, with all 64 lanes active, in VALU operations only.
// J = 1, 2: DPP quad_perm; J = 4: DPP row_half_mirror (lane i of an 8-lane half-row reads
// i ^ 7) then quad_perm [3,2,1,0] (reads i ^ 3); J = 8: DPP row_ror:8 (lane i of a 16-lane
// row reads (i + 8) mod 16);
// J = 16, 32: v_permlane16_swap / v_permlane32_swap of the value with itself, which yield
// (rows 0, 0, 2, 2 | rows 1, 1, 3, 3) and (lanes 0-31 twice | lanes 32-63 twice), and the
// lane's half picks the other one.  The DPP moves take old = 0 with bound_ctrl, so that the
// compiler can fold them into the instruction that uses them.
template <int J>
__device__ __forceinline__ uint32_t lane_xor(uint32_t v) {
  const int x = static_cast<int>(v);
  if constexpr (J == 1) {
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, true));
  } else if constexpr (J == 2) {
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, true));
  } else if constexpr (J == 4) {
    const int m7 = __builtin_amdgcn_update_dpp(0, x, 0x141, 0xF, 0xF, true);
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, m7, 0x1B, 0xF, 0xF, true));
  } else if constexpr (J == 8) {
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, x, 0x128, 0xF, 0xF, true));
  } else if constexpr (J == 16) {
    const auto p = __builtin_amdgcn_permlane16_swap(x, x, false, false);
    return static_cast<uint32_t>((__lane_id() & 16) ? p[0] : p[1]);
  } else {
    static_assert(J == 32, "lane_xor: J must be a power of two below 64");
    const auto p = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    return static_cast<uint32_t>((__lane_id() & 32) ? p[0] : p[1]);
  }
}
template <int J>
__device__ __forceinline__ uint64_t lane_xor_u64(uint64_t v) {
  const uint32_t lo = lane_xor<J>(static_cast<uint32_t>(v));
  const uint32_t hi = lane_xor<J>(static_cast<uint32_t>(v >> 32));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// One DPP move of a 64-bit value (both halves with the same control); lanes whose source is
// outside its row (bound_ctrl) or whose row is masked off read 0.
template <int CTRL, int ROWMASK, bool BOUND>
__device__ __forceinline__ int64_t dpp_i64(int64_t x) {
  const uint64_t u = static_cast<uint64_t>(x);
  const uint32_t lo = static_cast<uint32_t>(
      __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint32_t>(u)), CTRL, ROWMASK, 0xF, BOUND));
  const uint32_t hi = static_cast<uint32_t>(
      __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint32_t>(u >> 32)), CTRL, ROWMASK, 0xF, BOUND));
  return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
}

// Inclusive prefix sum of an int64 across the 64 lanes (all active): Hillis-Steele inside each
// 16-lane row by DPP row_shr 1, 2, 4, 8 (zero-filled), then row_bcast:15 adds each even row's
// total to the next row and row_bcast:31 adds lane 31's total to rows 2 and 3.
__device__ __forceinline__ int64_t wave_incl_sum_i64(int64_t x) {
  x += dpp_i64<0x111, 0xF, true>(x);
  x += dpp_i64<0x112, 0xF, true>(x);
  x += dpp_i64<0x114, 0xF, true>(x);
  x += dpp_i64<0x118, 0xF, true>(x);
  x += dpp_i64<0x142, 0xA, false>(x);
  x += dpp_i64<0x143, 0xC, false>(x);
  return x;
}

__device__ __forceinline__ uint64_t umin64(uint64_t a, uint64_t b) { return a < b ? a : b; }
__device__ __forceinline__ uint64_t umax64(uint64_t a, uint64_t b) { return a < b ? b : a; }

// Bitonic sort of one 64-bit key per lane, ascending across the 64 lanes of a wave (all lanes
// active).  Stages are unrolled at compile time so that each exchange is a lane_xor.
template <int K, int J>
__device__ __forceinline__ uint64_t wave_sort_stage(uint64_t x, int l) {
  const uint64_t y = lane_xor_u64<J>(x);
  const bool up = (l & K) == 0;
  const bool lower = (l & J) == 0;
  x = (lower == up) ? umin64(x, y) : umax64(x, y);
  if constexpr (J > 1) return wave_sort_stage<K, J / 2>(x, l);
  else return x;
}
template <int K>
__device__ __forceinline__ uint64_t wave_sort_from(uint64_t x, int l) {
  x = wave_sort_stage<K, K / 2>(x, l);
  if constexpr (K < kWave) return wave_sort_from<2 * K>(x, l);
  else return x;
}
__device__ __forceinline__ uint64_t wave_sort_u64(uint64_t x) { return wave_sort_from<2>(x, lane_id()); }

template <int J>
__device__ __forceinline__ uint64_t wave_merge_stage(uint64_t x, int l) {
  const uint64_t y = lane_xor_u64<J>(x);
  x = ((l & J) == 0) ? umin64(x, y) : umax64(x, y);
  if constexpr (J > 1) return wave_merge_stage<J / 2>(x, l);
  else return x;
}
// Merge: `cur` and `add` ascending across lanes; returns the 64 smallest of both, ascending.
__device__ __forceinline__ uint64_t wave_merge_min_u64(uint64_t cur, uint64_t add_sorted) {
  const int l = lane_id();
  const uint64_t rev = shfl_u64(add_sorted, kWave - 1 - l);
  return wave_merge_stage<kWave / 2>(umin64(cur, rev), l);  // the minima form a bitonic sequence
}

// Order-preserving map float -> uint32 (total order; -0 < +0 is irrelevant here).
__device__ __forceinline__ uint32_t float_ord(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord_float(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// Wave-wide sums (deterministic butterfly order).
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) v += __shfl_xor(v, m, kWave);
  return v;
}
__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) v += static_cast<int64_t>(shfl_xor_u64(static_cast<uint64_t>(v), m));
  return v;
}
__device__ __forceinline__ float wave_max_f32(float v) {
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) v = fmaxf(v, __shfl_xor(v, m, kWave));
  return v;
}

// ---- numerics shared bit-for-bit with the CPU oracle ---------------------------------------
// exp(x) in float32: Cody-Waite reduction x = n ln2 + r with the fdlibm two-part ln2, the
// degree-7 Taylor polynomial of e^r by Horner's rule (separately rounded mul/add), scaled
// by 2^n.  x < -87 gives 0, x > 88 gives +inf.
__device__ __forceinline__ float expf32(float x) {
  if (x != x) return x;
  if (x < -87.0f) return 0.0f;
  if (x > 88.0f) return __uint_as_float(0x7F800000u);
  const float nf = rintf(__fmul_rn(x, 1.44269504088896341f));
  float r = __fsub_rn(x, __fmul_rn(nf, 6.93145751953125e-01f));
  r = __fsub_rn(r, __fmul_rn(nf, 1.42860682030941723e-06f));
  float p = static_cast<float>(1.0 / 5040.0);
  p = __fadd_rn(__fmul_rn(p, r), static_cast<float>(1.0 / 720.0));
  p = __fadd_rn(__fmul_rn(p, r), static_cast<float>(1.0 / 120.0));
  p = __fadd_rn(__fmul_rn(p, r), static_cast<float>(1.0 / 24.0));
  p = __fadd_rn(__fmul_rn(p, r), static_cast<float>(1.0 / 6.0));
  p = __fadd_rn(__fmul_rn(p, r), 0.5f);
  p = __fadd_rn(__fmul_rn(p, r), 1.0f);
  p = __fadd_rn(__fmul_rn(p, r), 1.0f);
  return ldexpf(p, static_cast<int>(nf));
}

// splitmix64 finaliser and the per-(seed, phase, pass, vertex) coin stream.
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 30;
  x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 27;
  x *= 0x94D049BB133111EBull;
  x ^= x >> 31;
  return x;
}
__host__ __device__ __forceinline__ uint64_t stream_key(int64_t seed, uint64_t phase) {
  uint64_t s = static_cast<uint64_t>(seed & 0x7FFFFFFFFFFFFFFFll);
  return mix64(s * 0x9E3779B97F4A7C15ull + phase);
}
__host__ __device__ __forceinline__ uint64_t pass_key(uint64_t key, int pass) {
  return mix64(key + static_cast<uint64_t>(pass + 1) * 0xD1B54A32D192ED03ull);
}
__device__ __forceinline__ uint32_t coin_u32(uint64_t pkey, uint64_t v) {
  return static_cast<uint32_t>(mix64(pkey ^ (v * 0x9E3779B97F4A7C15ull)) >> 32);
}

// ---- wave-cooperative sorts of keys held in LDS -------------------------------------------
// For key sets wider than a wave (64 lanes): one wave sorts KP (a power of two >= 2) u64 keys
// in its own LDS region.  A wave's LDS operations complete in issue order, so between stages
// only a compiler-level ordering point is needed, not a block barrier.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// index of the lower element of compare pair p in a stage of the given stride
__device__ __forceinline__ int bitonic_lower(int p, int stride) {
  return ((p & ~(stride - 1)) << 1) | (p & (stride - 1));
}

// ascending bitonic sort of a[0..KP)
__device__ inline void wave_lds_sort_u64(uint64_t* a, int KP) {
  const int l = lane_id();
  for (int size = 2; size <= KP; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int p = l; p < (KP >> 1); p += kWave) {
        const int i = bitonic_lower(p, stride), j = i + stride;
        const uint64_t x = a[i], y = a[j];
        const bool up = (i & size) == 0;
        if ((x > y) == up) {
          a[i] = y;
          a[j] = x;
        }
      }
      wave_lds_sync();
    }
  }
}

// list[0..KP) <- the KP smallest keys of list and buf (both ascending), ascending: the
// elementwise minimum of list and reversed buf is bitonic, and half-cleaners sort it.
__device__ inline void wave_lds_merge_min_u64(uint64_t* list, const uint64_t* buf, int KP) {
  const int l = lane_id();
  for (int i = l; i < KP; i += kWave) {
    const uint64_t y = buf[KP - 1 - i];
    if (y < list[i]) list[i] = y;
  }
  wave_lds_sync();
  for (int stride = KP >> 1; stride > 0; stride >>= 1) {
    for (int p = l; p < (KP >> 1); p += kWave) {
      const int i = bitonic_lower(p, stride), j = i + stride;
      const uint64_t x = list[i], y = list[j];
      if (x > y) {
        list[i] = y;
        list[j] = x;
      }
    }
    wave_lds_sync();
  }
}

inline int pow2_at_least(int64_t x, int floor_value) {
  int p = floor_value;
  while (p < x) p <<= 1;
  return p;
}

}  // namespace msc
