// gpz_kernels.hip -- MI355X (gfx950) kernels of the GPZ particle compressor
// and the C ABI declared in include/gpz.h.
//
// One wave64 owns one 1024-point group.  Lane l holds the 16 points whose
// logical position in CUB's blocked sort input is L = 16*l + k, i.e. slot
// j = 32*(16*(l&1) + k) + (l>>1) of the group (the reference loads slot
// 32*i + t into thread t item i, src/cudaWarpCompress.cu:104-118, and
// BlockRadixSort sees thread t item i at position 32*t + i, :213).  Loading
// in that order lets a stable LSD radix sort reproduce the reference's tie
// order without a transpose; each load instruction still covers two full
// 128-byte lines.
//
// Pipeline per call (all on one stream, no host round trip):
//   seg_sample / seg_decide : device-side computeSegSize
//                             (src/cudaWarpCompress.cu:1424-1557)
//   encode                  : quantise, block id, LDS radix sort, runs,
//                             bit planes staged in LDS, copied to a fixed
//                             9344-byte slot (src/cudaWarpCompress.cu:58-435)
//   scan                    : exclusive scan of payload sizes, header
//                             (src/cudaWarpCompress.cu:1658-1683)
//   compact                 : meta + payload into the final stream
//                             (src/cudaWarpCompress.cu:32-56,1686-1688)
//   decode                  : inverse, one wave per group
//                             (src/cudaWarpCompress.cu:798-1214)
//
// Arithmetic follows the reference bit for bit: explicit fma, IEEE division
// and sqrt, saturating float->unsigned conversion (CUDA cvt.rzi), shifts by
// >= the width give 0 (PTX shr), u32 (FP32) / u64 (FP64) wrap-around.

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "gpz.h"

namespace {

// ------------------------------------------------ optional kernel timing
// When enabled (gpz_timing_enable), every launch is bracketed by a pair of
// hipEvents on its own stream; gpz_timing_read sums the elapsed times.
struct TimingState {
  std::mutex mu;
  bool on = false;
  struct Rec { std::string name; hipEvent_t a, b; };
  std::vector<Rec> pending;
  std::map<std::string, std::pair<double, uint64_t>> totals;
};
TimingState& timing() {
  static TimingState t;
  return t;
}
struct LaunchTimer {
  const char* name;
  hipStream_t s;
  hipEvent_t a = nullptr, b = nullptr;
  LaunchTimer(const char* n, hipStream_t st) : name(n), s(st) {
    if (!timing().on || !n) return;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess ||
        hipEventRecord(a, s) != hipSuccess)
      a = b = nullptr;
  }
  ~LaunchTimer() {
    if (!a) return;
    (void)hipEventRecord(b, s);
    std::lock_guard<std::mutex> g(timing().mu);
    timing().pending.push_back({name, a, b});
  }
};

constexpr uint32_t kGroup = 1024;
// s_waitcnt vmcnt(0) (expcnt and lgkmcnt left at their maxima), gfx9 encoding
constexpr int kWaitVmcnt0 = 0x0F70;
constexpr uint32_t kSlot = 9344;       // max payload bytes of one group
constexpr uint32_t kSampleLen = 256;   // src/cudaWarpCompress.cu:1425-1426
constexpr uint32_t kSampleBlocks = 1024;
constexpr uint32_t kMinMaxBlocks = 1024;

template <typename T> struct Prec;
template <> struct Prec<float> {
  using Q = uint32_t;
  static constexpr uint32_t kHeader = 24, kMeta = 36, kPad = 4;
  __device__ static float lo() { return 3.402823466e+38f; }
};
template <> struct Prec<double> {
  using Q = uint64_t;
  static constexpr uint32_t kHeader = 32, kMeta = 64, kPad = 8;
  __device__ static double lo() { return 1.7976931348623157e+308; }
};

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ uint32_t sat_u32(float f) {
  if (!(f >= 1.0f)) return 0u;
  if (f >= 4294967296.0f) return 0xFFFFFFFFu;
  return static_cast<uint32_t>(f);
}
__device__ __forceinline__ uint32_t sat_u32(double f) {
  if (!(f >= 1.0)) return 0u;
  if (f >= 4294967296.0) return 0xFFFFFFFFu;
  return static_cast<uint32_t>(f);
}
__device__ __forceinline__ uint64_t sat_u64(double f) {
  if (!(f >= 1.0)) return 0ull;
  if (f >= 18446744073709551616.0) return ~0ull;
  return static_cast<uint64_t>(f);
}
__device__ __forceinline__ uint32_t eff_bits(uint32_t x) {
  return x ? 32u - __builtin_clz(x) : 0u;
}
__device__ __forceinline__ uint32_t eff_bits(uint64_t x) {
  return x ? 64u - __builtin_clzll(x) : 0u;
}
__device__ __forceinline__ uint32_t shr(uint32_t v, uint32_t s) {
  return s >= 32 ? 0u : (v >> s);
}
__device__ __forceinline__ uint64_t shr(uint64_t v, uint32_t s) {
  return s >= 64 ? 0ull : (v >> s);
}
template <typename Q> __device__ __forceinline__ Q udiv(Q a, Q b) {
  return b ? a / b : ~Q(0);
}
template <typename Q> __device__ __forceinline__ Q urem(Q a, Q b) {
  return b ? a % b : a;
}

__device__ __forceinline__ float dmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ float dmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double dmin(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ double dmax(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float dfma(float a, float b, float c) {
  return fmaf(a, b, c);
}
__device__ __forceinline__ double dfma(double a, double b, double c) {
  return fma(a, b, c);
}
__device__ __forceinline__ float dceil(float a) { return ceilf(a); }
__device__ __forceinline__ double dceil(double a) { return ceil(a); }

// LDS ordering inside a single-wave workgroup.  A wave's LDS instructions
// execute in order, so a wavefront-scope fence plus wave_barrier (a code
// motion barrier) is enough; unlike __syncthreads it does not wait for
// outstanding global loads, which keeps prefetches in flight.  Only valid
// for kernels launched with 64-thread workgroups.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// wave64 collectives (ds_bpermute based shuffles)
template <typename V> __device__ __forceinline__ V wave_min(V v) {
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) v = dmin(v, __shfl_xor(v, m));
  return v;
}
template <typename V> __device__ __forceinline__ V wave_max(V v) {
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) v = dmax(v, __shfl_xor(v, m));
  return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) {
    const uint32_t o = static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), m));
    v = o > v ? o : v;
  }
  return v;
}
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t t = static_cast<uint32_t>(__shfl_up(static_cast<int>(v), d));
    if (lane >= static_cast<uint32_t>(d)) v += t;
  }
  return v;
}
__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t v, uint32_t lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t t = static_cast<uint64_t>(
        __shfl_up(static_cast<long long>(v), d));
    if (lane >= static_cast<uint32_t>(d)) v += t;
  }
  return v;
}

// DPP collectives.  update_dpp moves a value between lanes inside the
// VALU instruction that consumes it (no LDS round trip); the controls are
// the GCN/CDNA DPP encodings.  Every reduction ends in lane 63 and is read
// back as a uniform value.
constexpr int kDppQuad1032 = 0xB1;      // quad_perm:[1,0,3,2]
constexpr int kDppQuad2301 = 0x4E;      // quad_perm:[2,3,0,1]
constexpr int kDppRowShr = 0x110;       // row_shr:n is kDppRowShr + n
constexpr int kDppWaveShr1 = 0x138;     // wave_shr:1
constexpr int kDppRowMirror = 0x140;
constexpr int kDppRowHalfMirror = 0x141;
constexpr int kDppRowBcast15 = 0x142;   // lane 15 of each row -> next row
constexpr int kDppRowBcast31 = 0x143;   // lane 31 -> rows 2 and 3

// Lanes outside kRowMask, or reading from outside their row with
// kBound set, receive `old`.
template <int kCtrl, int kRowMask = 0xF, bool kBound = false>
__device__ __forceinline__ uint32_t dpp(uint32_t old, uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(
      static_cast<int>(old), static_cast<int>(v), kCtrl, kRowMask, 0xF, kBound));
}
__device__ __forceinline__ uint32_t lane63(uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}

__device__ __forceinline__ uint32_t wave_or(uint32_t v) {
  v |= dpp<kDppQuad1032>(0u, v);
  v |= dpp<kDppQuad2301>(0u, v);
  v |= dpp<kDppRowHalfMirror>(0u, v);
  v |= dpp<kDppRowMirror>(0u, v);
  v |= dpp<kDppRowBcast15, 0xA>(0u, v);
  v |= dpp<kDppRowBcast31, 0xC>(0u, v);
  return lane63(v);
}

// Inclusive prefix sum over lanes (Hillis-Steele inside each row of 16,
// then the row totals).
__device__ __forceinline__ uint32_t wave_scan(uint32_t v) {
  v += dpp<kDppRowShr + 1, 0xF, true>(0u, v);
  v += dpp<kDppRowShr + 2, 0xF, true>(0u, v);
  v += dpp<kDppRowShr + 4, 0xF, true>(0u, v);
  v += dpp<kDppRowShr + 8, 0xF, true>(0u, v);
  v += dpp<kDppRowBcast15, 0xA>(0u, v);
  v += dpp<kDppRowBcast31, 0xC>(0u, v);
  return v;
}

// IEEE-mode v_min/v_max on operands that are already canonical (results
// of earlier min/max): NaN operands are quiet there and are ignored, as
// fminf/fmaxf ignore them.
__device__ __forceinline__ float vmin(float a, float b) {
  float r;
  asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ float vmax(float a, float b) {
  float r;
  asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
template <bool kMax>
__device__ __forceinline__ float wave_ext(float v) {
  auto step = [&](uint32_t o) {
    v = kMax ? vmax(v, __uint_as_float(o)) : vmin(v, __uint_as_float(o));
  };
  uint32_t u = __float_as_uint(v);
  step(dpp<kDppQuad1032>(u, u)); u = __float_as_uint(v);
  step(dpp<kDppQuad2301>(u, u)); u = __float_as_uint(v);
  step(dpp<kDppRowHalfMirror>(u, u)); u = __float_as_uint(v);
  step(dpp<kDppRowMirror>(u, u)); u = __float_as_uint(v);
  step(dpp<kDppRowBcast15, 0xA>(u, u)); u = __float_as_uint(v);
  step(dpp<kDppRowBcast31, 0xC>(u, u));
  return __uint_as_float(lane63(__float_as_uint(v)));
}

// A double that is equal in every lane, read from the first active lane
// (so that the compiler can keep it in scalar registers).
template <typename V>
__device__ __forceinline__ V uniform_f64(V v) {
  if constexpr (sizeof(V) == 8) {
    const uint64_t u = static_cast<uint64_t>(__double_as_longlong(static_cast<double>(v)));
    const uint32_t lo = static_cast<uint32_t>(
        __builtin_amdgcn_readfirstlane(static_cast<int>(static_cast<uint32_t>(u))));
    const uint32_t hi = static_cast<uint32_t>(
        __builtin_amdgcn_readfirstlane(static_cast<int>(static_cast<uint32_t>(u >> 32))));
    return static_cast<V>(__longlong_as_double(static_cast<long long>(
        (static_cast<uint64_t>(hi) << 32) | lo)));
  } else {
    return v;
  }
}

// Bit-matrix transposes.  One swap stage exchanges the m-masked bits of
// a >> j with those of b (two shifts and two bitfield inserts).
__device__ __forceinline__ void swap_bits(uint32_t& a, uint32_t& b, int j,
                                          uint32_t m) {
  const uint32_t na = ((b << j) & (m << j)) | (a & ~(m << j));
  b = ((a >> j) & m) | (b & ~m);
  a = na;
}

// Transpose of 16 words, two 16x16 blocks at once: on entry a[i] holds
// item i (bits 0..31); on exit a[b] bits 0..15 hold bit b of items 0..15
// and bits 16..31 hold bit b+16.  Its own inverse.
__device__ __forceinline__ void transpose16x2(uint32_t (&a)[16]) {
#pragma unroll
  for (int st = 0; st < 4; ++st) {
    const int j = 8 >> st;
    const uint32_t m = st == 0 ? 0x00FF00FFu : st == 1 ? 0x0F0F0F0Fu
                     : st == 2 ? 0x33333333u : 0x55555555u;
#pragma unroll
    for (int k = 0; k < 16; ++k)
      if (!(k & j)) swap_bits(a[k], a[k + j], j, m);
  }
}

// 16x16 transpose of 16-bit values held two per register.  pack_pair(lo,
// hi) lays out items lo and hi (low 16 bits each) as bytes [lo.0, hi.0,
// lo.1, hi.1]; transpose16_bytes then transposes the four 8x8 blocks
// formed by each byte lane of r[0..7].  Starting from r[i] =
// pack_pair(item i, item i+8) it leaves r[b] bits 0..15 = bit b of items
// 0..15 and bits 16..31 = bit b+8; starting from r[b] = pack_pair(plane b,
// plane b+8) it leaves r[i] = item i | item i+8 << 16.
__device__ __forceinline__ uint32_t pack_pair(uint32_t lo, uint32_t hi) {
  return __builtin_amdgcn_perm(hi, lo, 0x05010400u);
}
__device__ __forceinline__ void transpose16_bytes(uint32_t (&r)[8]) {
#pragma unroll
  for (int st = 0; st < 3; ++st) {
    const int j = 4 >> st;
    const uint32_t m = st == 0 ? 0x0F0F0F0Fu : st == 1 ? 0x33333333u : 0x55555555u;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (!(k & j)) swap_bits(r[k], r[k + j], j, m);
  }
}
// Same transpose from the natural layout r[i] = lo | hi << 16.
__device__ __forceinline__ void transpose16_packed(uint32_t (&r)[8]) {
#pragma unroll
  for (int i = 0; i < 8; ++i) r[i] = __builtin_amdgcn_perm(r[i], r[i], 0x03010200u);
  transpose16_bytes(r);
}
__device__ __forceinline__ uint32_t plane_word16(const uint32_t (&r)[8], int b) {
  return b < 8 ? (r[b] & 0xFFFFu) : (r[b - 8] >> 16);
}

// Word of bit plane b (0..31) after transpose16x2.
__device__ __forceinline__ uint32_t plane_word(const uint32_t (&a)[16], int b) {
  return b < 16 ? (a[b] & 0xFFFFu) : (a[b - 16] >> 16);
}

// Exact unsigned 32-bit division by a per-group divisor d >= 1
// (Granlund & Montgomery, "Division by invariant integers using
// multiplication", 1994, fig. 4.1): q = (t + ((n - t) >> s1)) >> s2 with
// t = mulhi(m, n).
struct Div32 {
  uint32_t m, s1, s2;
};
__device__ __forceinline__ Div32 div32_make(uint32_t d) {
  const uint32_t l = d > 1 ? 32u - __builtin_clz(d - 1u) : 0u;  // ceil(log2 d)
  const uint64_t num = (static_cast<uint64_t>(1) << 32) *
                       ((static_cast<uint64_t>(1) << l) - d);
  Div32 r;
  r.m = static_cast<uint32_t>(num / d + 1u);
  r.s1 = l < 1 ? l : 1u;
  r.s2 = l > 0 ? l - 1u : 0u;
  return r;
}
__device__ __forceinline__ uint32_t div32(uint32_t n, Div32 v) {
  const uint32_t t = __umulhi(v.m, n);
  return (t + ((n - t) >> v.s1)) >> v.s2;
}
// Quotient for the decoder: divisor 0 follows udiv; wide operands (FP64
// streams only) take the slow path.
template <typename Q>
__device__ __forceinline__ Q qdiv(Q n, Q d, Div32 v) {
  if (d == 0) return ~Q(0);
  if (sizeof(Q) == 8 && ((n >> 32) != 0 || (d >> 32) != 0)) return n / d;
  return static_cast<Q>(div32(static_cast<uint32_t>(n), v));
}

// Section sizes (src/cudaWarpCompress.cu:324-329; FP64 3D
// src/cudaWarpCompressDouble.cu:325-329, FP64 2D :686-692).
template <typename T, int DIM>
__device__ __host__ __forceinline__ void section_sizes(
    uint32_t U, uint32_t db, uint32_t cb, uint32_t rb, uint32_t* D,
    uint32_t* C, uint32_t* R) {
  const uint32_t B = (U + 7u) >> 3;
  if (sizeof(T) == 4) {
    *D = (B * db + 3u) & ~3u;
    *C = (B * cb + 3u) & ~3u;
    *R = 128u * rb;
  } else {
    *D = (B * db + 7u) & ~7u;
    *C = DIM == 3 ? ((B * cb + 7u) & ~7u) : ((B * cb + 3u) & ~3u);
    *R = 128u * rb;
  }
}

// Block counts per axis from a group's extrema
// (src/cudaWarpCompress.cu:150-154; FP64 3D: u32 cast, 2D encoder: u64 cast,
// src/cudaWarpCompressDouble.cu:151-154,525-526).
template <typename T, int DIM>
__device__ __forceinline__ void block_counts(T mn0, T mx0, T mn1, T mx1,
                                             T seg2eb,
                                             typename Prec<T>::Q* nbx,
                                             typename Prec<T>::Q* nby) {
  using Q = typename Prec<T>::Q;
  const T fx = dceil((mx0 - mn0) / seg2eb);
  const T fy = dceil((mx1 - mn1) / seg2eb);
  if (sizeof(T) == 8 && DIM == 2) {
    *nbx = static_cast<Q>(sat_u64(static_cast<double>(fx)) + 1ull);
  } else {
    *nbx = static_cast<Q>(static_cast<uint32_t>(sat_u32(fx) + 1u));
  }
  *nby = static_cast<Q>(static_cast<uint32_t>(sat_u32(fy) + 1u));
}

// ------------------------------------------------------- workspace layout
struct WsLayout {
  uint64_t info, win, loc, off, bsum, slots, total;
};
__host__ __device__ inline uint64_t align256(uint64_t v) {
  return (v + 255) & ~255ull;
}
inline WsLayout ws_layout(uint32_t bits, uint64_t G) {
  (void)bits;
  WsLayout w;
  w.info = 0;
  w.win = 256;
  w.loc = align256(w.win + 6ull * kSampleBlocks * 8);
  w.off = align256(w.loc + 8 * G);
  w.bsum = align256(w.off + 8 * G);
  w.slots = align256(w.bsum + 8 * ((G + 1023) / 1024));
  w.total = align256(w.slots + static_cast<uint64_t>(kSlot) * G);
  return w;
}

struct DevInfo {  // == gpz_result, lives at workspace offset 0
  uint64_t stream_bytes;
  uint32_t seg_size;
  uint32_t adjusted;
  int32_t status;
  uint32_t reserved;
};

// ------------------------------------------------------ seg-size kernels
// Window b covers points [b*stride, b*stride + 256); windows that would run
// past N are skipped by the reference kernel (src/cudaWarpCompress.cu:1230)
// and leave the zeroed scratch in place, so they contribute 0.0.
template <typename T, int DIM>
__device__ __forceinline__ void seg_sample_window(
    const T* __restrict__ x, const T* __restrict__ y, const T* __restrict__ z,
    uint64_t n, uint64_t stride, T* __restrict__ win, uint32_t b) {
  __shared__ T red[4][6];
  const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
  const uint64_t s = static_cast<uint64_t>(b) * stride;
  if (s + kSampleLen > n) {
    if (t < 6) win[b * 6 + t] = T(0);
    return;
  }
  T v[3];
  v[0] = x[s + t];
  v[1] = y[s + t];
  v[2] = DIM == 3 ? z[s + t] : T(0);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const T lo = wave_min(v[a]);
    const T hi = wave_max(v[a]);
    if (lane == 0) {
      red[w][2 * a] = lo;
      red[w][2 * a + 1] = hi;
    }
  }
  __syncthreads();
  if (t < 6) {
    T r = red[0][t];
    for (int i = 1; i < 4; ++i)
      r = (t & 1) ? dmax(r, red[i][t]) : dmin(r, red[i][t]);
    win[b * 6 + t] = r;
  }
}
template <typename T, int DIM>
__global__ __launch_bounds__(256) void seg_sample_kernel(
    const T* __restrict__ x, const T* __restrict__ y, const T* __restrict__ z,
    uint64_t n, uint64_t stride, T* __restrict__ win) {
  seg_sample_window<T, DIM>(x, y, z, n, stride, win, blockIdx.x);
}

struct SegArgs {
  uint32_t seg_in;
  int32_t is_rel;
  int32_t rel_small;  // (double)rel_eb <= 1e-4, src/cudaWarpCompress.cu:1517
  int32_t sampled;    // window extrema were computed
  double eb;
};

// src/cudaWarpCompress.cu:1488-1556 (FP32) and
// src/cudaWarpCompressDouble.cu:1487-1553 (FP64), evaluated in T with the
// same operation order as the reference host code.
template <typename T, int DIM>
__device__ __forceinline__ void seg_decide_one(const T* __restrict__ win,
                                               const SegArgs& a, DevInfo* info) {
  __shared__ T red[4][6];
  const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
  if (!a.sampled) {
    if (t == 0) {
      info->seg_size = a.seg_in;
      info->adjusted = 0;
      info->status = a.seg_in ? GPZ_OK : GPZ_ERR_INVALID_ARG;
      info->stream_bytes = 0;
    }
    return;
  }
  T e[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) e[i] = (i & 1) ? -Prec<T>::lo() : Prec<T>::lo();
  for (uint32_t b = t; b < kSampleBlocks; b += 256)
#pragma unroll
    for (int i = 0; i < 6; ++i)
      e[i] = (i & 1) ? dmax(e[i], win[b * 6 + i]) : dmin(e[i], win[b * 6 + i]);
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    e[i] = (i & 1) ? wave_max(e[i]) : wave_min(e[i]);
    if (lane == 0) red[w][i] = e[i];
  }
  __syncthreads();
  if (t != 0) return;
  T g[6];
  for (int i = 0; i < 6; ++i) {
    g[i] = red[0][i];
    for (int k = 1; k < 4; ++k)
      g[i] = (i & 1) ? dmax(g[i], red[k][i]) : dmin(g[i], red[k][i]);
  }
  if (DIM == 2) { g[4] = T(0); g[5] = T(0); }
  const T dx = g[1] - g[0], dy = g[3] - g[2], dz = g[5] - g[4];
  const T mean = (dx + dy + dz) / T(3.0f);
  T var;
  if (sizeof(T) == 4 && DIM == 2)
    var = ((dx - mean) * (dx - mean) + (dy - mean) * (dy - mean) +
           (dz - mean) * (dz - mean)) / T(2.0f);
  else
    var = ((dx - mean) * (dx - mean) + (dy - mean) * (dy - mean) +
           (dz - mean) * (dz - mean)) / T(3.0f);
  const T ratio = sqrt(var) / mean;
  T max_range = dmax(dx, dy);
  max_range = dmax(max_range, dz);
  uint32_t seg = a.seg_in;
  if (seg == 0 && a.is_rel) {
    if (ratio < T(0.1f)) seg = a.rel_small ? 32u : 2u;
    else seg = a.rel_small ? 4u : 1u;
  }
  int32_t status = GPZ_OK;
  uint32_t adjusted = 0;
  if (sizeof(T) == 4) {
    const T eb = static_cast<T>(a.eb);
    auto blocks = [&](uint32_t s) -> uint64_t {
      const uint64_t mb =
          sat_u64(static_cast<double>(dceil(max_range / (static_cast<T>(2u * s) * eb)))) + 1ull;
      return DIM == 3 ? mb * mb * mb : mb * mb;
    };
    if (blocks(seg) > 0xFFFFFFFFull) {
      uint32_t tmp = seg;
      for (int i = 0; i < 11; ++i) {
        tmp *= 2;
        if (blocks(tmp) <= 0xFFFFFFFFull) break;
      }
      if (tmp > 1024) status = GPZ_ERR_SEG_TOO_LARGE;
      seg = tmp;
      adjusted = 1;
    }
  }
  if (seg == 0 && status == GPZ_OK) status = GPZ_ERR_INVALID_ARG;
  info->seg_size = seg;
  info->adjusted = adjusted;
  info->status = status;
  info->stream_bytes = 0;
}
template <typename T, int DIM>
__global__ __launch_bounds__(256) void seg_decide_kernel(const T* __restrict__ win,
                                                         SegArgs a,
                                                         DevInfo* info) {
  seg_decide_one<T, DIM>(win, a, info);
}

// ------------------------------------------------------------- encoder
struct EncodeArgs {
  const void* x;
  const void* y;
  const void* z;
  uint64_t n, G;
  uint32_t last_n;  // points in group G-1 (1..1024)
  double eb;
  const DevInfo* info;
  uint8_t* meta;    // stream meta table [G * kMeta]
  uint64_t* loc;    // [G]
  uint8_t* slots;   // [G * kSlot]
};

// LDS of one encode wave (bytes).  During the sort: padded keys and the
// digit counters (the pair sort parks its values in the counter region
// between the counter reads and the next pass).  After it: values by
// position for the gather, then per-run deltas and start ranks, the
// per-lane dump slots that non-start ranks write to, and the small-U plane
// stage.  8448 bytes let 16 waves share a CU.
constexpr uint32_t kLdsKeyB = 0;        // u32 [1088]: 16 keys + 1 pad word per row
constexpr uint32_t kLdsCntB = 4352;     // u32 [16 digits][64 lanes]
constexpr uint32_t kLdsBytes = 8448;
constexpr uint32_t kGatherB = 0;        // u32 [1024] values by position
constexpr uint32_t kRunDeltaB = 0;      // u32 [1024]
constexpr uint32_t kRunStartB = 4096;   // u16 [1040]
constexpr uint32_t kRunDump = 1600;     // + lane -> bytes 6400.. (delta), 7296.. (start)
constexpr uint32_t kStageB = 7424;      // <= 384 bytes
static_assert(kRunStartB + 2 * 1040 <= 4 * kRunDump &&
              4 * (kRunDump + 64) <= kRunStartB + 2 * kRunDump &&
              kRunStartB + 2 * (kRunDump + 64) <= kStageB &&
              kStageB + 384 <= kLdsBytes && kLdsCntB >= 4 * 1088,
              "LDS regions overlap");

// Sort key buffer: position p at word p + p/16.  Rows of 17 words keep
// both the blocked read-back (lane l reads 17l + k) and an identity
// scatter free of bank conflicts on 64 banks.
__device__ __forceinline__ uint32_t pad_word(uint32_t p) {
  return p + (p >> 4);
}
// Byte offset of pad_word(p), 4 * (p + p/16), in two VALU operations (the
// compiler's own form of the expression takes three).
__device__ __forceinline__ uint32_t pad_byte(uint32_t p) {
  uint32_t a;
  asm("v_add_lshl_u32 %0, %1, %2, 2" : "=v"(a) : "v"(p), "v"(p >> 4));
  return a;
}

// Global-memory pointers.  The group bodies take their buffers from an
// argument struct; in a batch that struct is read from a device array, and
// a pointer read from memory has an unknown address space, so every access
// through it would be a flat access (which also makes every LDS wait wait
// for global loads).  The casts state what the buffers are.
#define GPZ_GLOBAL __attribute__((address_space(1)))
template <typename P>
__device__ __forceinline__ GPZ_GLOBAL P* gbl(P* p) {
  return (GPZ_GLOBAL P*)p;
}

// Quantised code of one coordinate: sat_u32 / sat_u64 of fma(x, 1/(2eb),
// -min/(2eb)) (src/cudaWarpCompress.cu:179-184,
// src/cudaWarpCompressDouble.cu:186-193).  v_cvt_u32_f32 truncates,
// clamps to [0, 2^32-1] and maps NaN to 0, the same result as CUDA's
// cvt.rzi.u32.f32 and as sat_u32 above.
__device__ __forceinline__ uint32_t cvt_u32_sat(float f) {
  uint32_t r;
  asm("v_cvt_u32_f32 %0, %1" : "=v"(r) : "v"(f));
  return r;
}
template <typename T>
__device__ __forceinline__ typename Prec<T>::Q quant_q(T x, T inv, T nmi) {
  if constexpr (sizeof(T) == 4) return cvt_u32_sat(dfma(x, inv, nmi));
  else return sat_u64(dfma(x, inv, nmi));
}

// One pass of the stable LSD radix sort over a 4-bit digit.  Items are in
// blocked order (lane l holds positions 16l..16l+15).  Each lane counts its
// items in its own counter column (word 64*digit + lane, i.e. bank = lane)
// with LDS atomics: a column is never shared, so the atomics neither
// contend nor conflict and their return values are the lane-local ranks in
// item order.  An exclusive scan over (digit, lane) order turns the
// counters into bases: lane L scans the 16 counters [16L, 16L+16), held as
// four 16-byte chunks that it visits in the order rot[q] = (q + a) & 3,
// a = (L >> 2) & 3, so that the 16 lanes served in one LDS cycle hit 16
// different bank groups.  The items are then scattered into the padded key
// buffer and read back blocked.  With kPair the values travel along
// (unpadded).
struct ScanLane {
  uint32_t chunk[4];  // uint4 index of the chunk visited at step q
  bool wrap1, wrap2, wrap3;  // step q holds a chunk before rot[0] (a + q >= 4)
  bool a1, a2, a3;
};
__device__ __forceinline__ ScanLane scan_lane(uint32_t lane) {
  ScanLane s;
  const uint32_t a = (lane >> 2) & 3u;
#pragma unroll
  for (int q = 0; q < 4; ++q) s.chunk[q] = 4u * lane + ((q + a) & 3u);
  s.wrap1 = a >= 3; s.wrap2 = a >= 2; s.wrap3 = a >= 1;
  s.a1 = a == 1; s.a2 = a == 2; s.a3 = a == 3;
  return s;
}

template <bool kPair>
__device__ __forceinline__ void radix_pass(uint32_t (&key)[16], uint32_t (&val)[16],
                                           uint32_t shift, uint32_t lane,
                                           const ScanLane& sl, uint8_t* lds) {
  uint32_t* s_key = reinterpret_cast<uint32_t*>(lds + kLdsKeyB);
  uint32_t* s_cnt = reinterpret_cast<uint32_t*>(lds + kLdsCntB);
  uint32_t* s_val = s_cnt;  // free once every counter has been read
  uint4* c4 = reinterpret_cast<uint4*>(s_cnt);
#pragma unroll
  for (int q = 0; q < 4; ++q) c4[sl.chunk[q]] = make_uint4(0, 0, 0, 0);
  wave_sync();
  uint32_t idx[16], rnk[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    idx[k] = (__builtin_amdgcn_ubfe(key[k], shift, 4) << 6) + lane;
    rnk[k] = __hip_atomic_fetch_add(s_cnt + idx[k], 1u, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  wave_sync();
  {
    uint32_t cv[16], S[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint4 w = c4[sl.chunk[q]];
      cv[4 * q] = w.x; cv[4 * q + 1] = w.y; cv[4 * q + 2] = w.z; cv[4 * q + 3] = w.w;
      S[q] = w.x + w.y + w.z + w.w;
    }
    // P[q]: sum of the chunks visited before step q; the chunk at step q
    // starts at ex + P[q] - P[4-a] (+ T when it lies after chunk a).
    const uint32_t P1 = S[0], P2 = S[0] + S[1], P3 = P2 + S[2], T = P3 + S[3];
    const uint32_t Pa = sl.a1 ? P3 : sl.a2 ? P2 : sl.a3 ? P1 : T;
    const uint32_t X = wave_scan(T) - T - Pa;
    const uint32_t Y = X + T;
    uint32_t run[4] = {Y, (sl.wrap1 ? X : Y) + P1, (sl.wrap2 ? X : Y) + P2,
                       (sl.wrap3 ? X : Y) + P3};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const uint32_t t = cv[4 * q + e];
        cv[4 * q + e] = run[q];
        run[q] += t;
      }
      c4[sl.chunk[q]] = make_uint4(cv[4 * q], cv[4 * q + 1], cv[4 * q + 2], cv[4 * q + 3]);
    }
  }
  wave_sync();
  // all counter reads first: they must precede the scatter writes (the
  // values overwrite the counters) and are then not ordered behind them
#pragma unroll
  for (int k = 0; k < 16; ++k) rnk[k] += s_cnt[idx[k]];
  if (kPair) wave_sync();
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    *reinterpret_cast<uint32_t*>(lds + kLdsKeyB + pad_byte(rnk[k])) = key[k];
    if (kPair) s_val[rnk[k]] = val[k];
  }
  wave_sync();
#pragma unroll
  for (int k = 0; k < 16; ++k) key[k] = s_key[17 * lane + k];
#pragma unroll
  for (int q4 = 0; q4 < 4; ++q4) {
    if (kPair) {
      const uint4 vv = reinterpret_cast<const uint4*>(s_val)[4 * lane + q4];
      val[4 * q4 + 0] = vv.x; val[4 * q4 + 1] = vv.y;
      val[4 * q4 + 2] = vv.z; val[4 * q4 + 3] = vv.w;
    }
  }
  wave_sync();
}

// MSB-first bit planes of 16 per-lane values (runs 16l..16l+15) stored to
// dst: plane p (bit nb-1-p) is B bytes at dst + p*B, lane l owns bytes 2l,
// 2l+1.  Feeding the transpose item t^7 in slot t makes bit s of a plane
// word hold run 16l + (s^7), i.e. both bytes come out MSB-first.
template <int kMaxBits>
__device__ __forceinline__ void store_msb_planes(const uint32_t (&v)[16], uint32_t nb,
                                                 uint32_t B, uint32_t padded,
                                                 GPZ_GLOBAL uint8_t* dst, uint32_t lane) {
  const bool w0 = 2u * lane < B, w1 = 2u * lane + 1u < B;
  if (kMaxBits <= 16 || nb <= 16) {
    uint32_t r[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) r[i] = pack_pair(v[7 - i], v[15 - i]);
    transpose16_bytes(r);
#pragma unroll
    for (int b = 0; b < 16; ++b) {
      if (b < static_cast<int>(nb)) {
        const uint32_t w = plane_word16(r, b);
        GPZ_GLOBAL uint8_t* d = dst + (nb - 1u - b) * B + 2u * lane;
        if (w0) d[0] = static_cast<uint8_t>(w);
        if (w1) d[1] = static_cast<uint8_t>(w >> 8);
      }
    }
  } else {
    uint32_t t16[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) t16[t] = v[t ^ 7];
    transpose16x2(t16);
#pragma unroll
    for (int b = 0; b < kMaxBits; ++b) {
      if (b < static_cast<int>(nb)) {
        const uint32_t w = plane_word(t16, b);
        GPZ_GLOBAL uint8_t* d = dst + (nb - 1u - b) * B + 2u * lane;
        if (w0) d[0] = static_cast<uint8_t>(w);
        if (w1) d[1] = static_cast<uint8_t>(w >> 8);
      }
    }
  }
  if (lane < padded - B * nb) dst[B * nb + lane] = 0;  // zero pad bytes
}

// Stage 1-2 of one group: extrema, then block id and relative position of
// every point (src/cudaWarpCompress.cu:120-200).  kFast: power-of-two seg
// with 3*log2(seg) <= 30 and every shifted code, block count and block id
// small enough that 24-bit multiplies are exact and no 31-bit mask bites.
template <typename T, int DIM, bool kFull, bool kFast>
__device__ __forceinline__ void quantise(const T (&px)[16], const T (&py)[16],
                                         const T (&pz)[16], T inv, const T (&nmi)[3],
                                         uint32_t seg, uint32_t nbx32, uint32_t nbxy32,
                                         uint32_t jbase, uint32_t nvalid,
                                         uint32_t (&key)[16], uint32_t (&val)[16],
                                         uint32_t* bor_out, uint32_t* ror_out) {
  const uint32_t sh = static_cast<uint32_t>(__builtin_ctz(seg));
  const uint32_t m = seg - 1u;
  const bool pow2 = (seg & m) == 0u;
  // z term of the power-of-two relative position; 0 when 2*sh >= 32 (the
  // product by seg*seg wraps to 0 in 32 bits)
  const uint32_t zsh = (2u * sh) & 31u;
  const uint32_t mz = 2u * sh >= 32 ? 0u : m;
  uint32_t bor = 0, ror = 0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const auto qx = quant_q<T>(px[k], inv, nmi[0]);
    const auto qy = quant_q<T>(py[k], inv, nmi[1]);
    const auto qz = DIM == 3 ? quant_q<T>(pz[k], inv, nmi[2]) : decltype(qx)(0);
    const uint32_t sx = static_cast<uint32_t>(qx >> sh);
    const uint32_t sy = static_cast<uint32_t>(qy >> sh);
    const uint32_t sz = static_cast<uint32_t>(qz >> sh);
    const uint32_t lx = static_cast<uint32_t>(qx), ly = static_cast<uint32_t>(qy),
                   lz = static_cast<uint32_t>(qz);
    uint32_t b, r;
    if (kFast) {
      // b = sx + sy*nbx + sz*nbxy with 24-bit multiply-adds, and
      // r = ((lz&m) << 2sh) | ((ly&m) << sh) | (lx&m) as shift-or chains
      // (five operations each; the compiler's own forms take six)
      if (DIM == 3) {
        uint32_t t;
        asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(t) : "v"(sz), "s"(nbxy32), "v"(sx));
        asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(b) : "v"(sy), "s"(nbx32), "v"(t));
        uint32_t u;
        asm("v_lshl_or_b32 %0, %1, %2, %3" : "=v"(u) : "v"(lz & m), "s"(sh), "v"(ly & m));
        asm("v_lshl_or_b32 %0, %1, %2, %3" : "=v"(r) : "v"(u), "s"(sh), "v"(lx & m));
      } else {
        asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(b) : "v"(sy), "s"(nbx32), "v"(sx));
        asm("v_lshl_or_b32 %0, %1, %2, %3" : "=v"(r) : "v"(ly & m), "s"(sh), "v"(lx & m));
      }
    } else {
      b = sx + sy * nbx32;
      if (DIM == 3) b += sz * nbxy32;
      if (pow2) {
        r = (lx & m) | ((ly & m) << sh);
        if (DIM == 3) r |= (lz & mz) << zsh;
      } else {
        r = (lx & m) + (ly & m) * seg;
        if (DIM == 3) r += (lz & m) * seg * seg;
      }
      b &= 0x7FFFFFFFu;
      r &= 0x7FFFFFFFu;
    }
    if (kFull) {
      key[k] = b;
      val[k] = r;
      bor |= b;
      ror |= r;
    } else {
      const bool v = jbase + 32u * k < nvalid;
      key[k] = v ? b : 0xFFFFFFFFu;
      val[k] = v ? r : 0xFFFFFFFFu;
      bor |= v ? b : 1u;
      ror |= v ? r : 1u;
    }
  }
  *bor_out = bor;
  *ror_out = ror;
}

// The same stage for FP64, one axis at a time: the coordinates are read
// again (from L2: the group was read for its extrema just before) and each
// axis's codes are folded into the block ids and relative positions before
// the next axis is read, so that 32 rather than 96 registers hold doubles.
// Sums and products are taken in the same 32-bit wrap-around arithmetic as
// quantise(), whose result they equal.
template <typename T, int DIM, bool kFull, bool kFast>
__device__ __forceinline__ void quantise_axes(const GPZ_GLOBAL T* X, const GPZ_GLOBAL T* Y,
                                              const GPZ_GLOBAL T* Z,
                                              T inv, const T (&nmi)[3],
                                              uint32_t seg, uint32_t nbx32, uint32_t nbxy32,
                                              uint32_t jbase, uint32_t nvalid,
                                              uint32_t (&key)[16], uint32_t (&val)[16],
                                              uint32_t* bor_out, uint32_t* ror_out) {
  const uint32_t sh = static_cast<uint32_t>(__builtin_ctz(seg));
  const uint32_t m = seg - 1u;
  const bool pow2 = (seg & m) == 0u;
  const uint32_t zsh = (2u * sh) & 31u;
  const uint32_t mz = 2u * sh >= 32 ? 0u : m;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const GPZ_GLOBAL T* P = d == 0 ? X : d == 1 ? Y : Z;
    uint32_t jb = jbase;
    asm volatile("" : "+v"(jb));  // a fresh read, not the registers of the first
    T v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      uint32_t j = jb + 32u * k;
      if (!kFull) j = j < nvalid ? j : nvalid - 1u;
      v[k] = P[j];
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const auto q = quant_q<T>(v[k], inv, nmi[d]);
      const uint32_t s = static_cast<uint32_t>(q >> sh);
      const uint32_t l = static_cast<uint32_t>(q);
      if (d == 0) {
        key[k] = s;
        val[k] = l & m;
      } else if (d == 1) {
        if (kFast) {
          asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(key[k]) : "v"(s), "s"(nbx32), "v"(key[k]));
        } else {
          key[k] += s * nbx32;
        }
        val[k] = (kFast || pow2) ? val[k] | ((l & m) << sh) : val[k] + (l & m) * seg;
      } else {
        if (kFast) {
          asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(key[k]) : "v"(s), "s"(nbxy32), "v"(key[k]));
          val[k] |= (l & m) << (2u * sh);
        } else {
          key[k] += s * nbxy32;
          val[k] = pow2 ? val[k] | ((l & mz) << zsh) : val[k] + (l & m) * seg * seg;
        }
      }
    }
  }
  uint32_t bor = 0, ror = 0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    uint32_t b = key[k], r = val[k];
    if (!kFast) {
      b &= 0x7FFFFFFFu;
      r &= 0x7FFFFFFFu;
    }
    if (kFull) {
      key[k] = b;
      val[k] = r;
      bor |= b;
      ror |= r;
    } else {
      const bool v = jbase + 32u * k < nvalid;
      key[k] = v ? b : 0xFFFFFFFFu;
      val[k] = v ? r : 0xFFFFFFFFu;
      bor |= v ? b : 1u;
      ror |= v ? r : 1u;
    }
  }
  *bor_out = bor;
  *ror_out = ror;
}

template <typename T, int DIM, bool kFull>
__device__ __forceinline__ void encode_group(const EncodeArgs& a, uint64_t c,
                                             uint32_t lane, uint8_t* lds) {
  using Q = typename Prec<T>::Q;
  const uint32_t seg = gbl(a.info)->seg_size;
  const T eb = static_cast<T>(a.eb);
  const uint64_t base = c * kGroup;
  const uint32_t nvalid = kFull ? kGroup : a.last_n;
  const bool last = (c == a.G - 1);
  const uint32_t jbase = 512u * (lane & 1u) + (lane >> 1);

  // ---- stage 0: coordinates in blocked sort order (lane l, item k = slot
  // 32*(16*(l&1) + k) + (l>>1)).  In the partial last group, slots past N
  // re-read the group's last point: the extrema do not change and their
  // keys are replaced below.
  T px[16], py[16], pz[16];
  {
    const GPZ_GLOBAL T* X = gbl(static_cast<const T*>(a.x)) + base;
    const GPZ_GLOBAL T* Y = gbl(static_cast<const T*>(a.y)) + base;
    const GPZ_GLOBAL T* Z = gbl(static_cast<const T*>(a.z)) + base;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      uint32_t j = jbase + 32u * k;
      if (!kFull) j = j < nvalid ? j : nvalid - 1u;
      px[k] = X[j];
      py[k] = Y[j];
      pz[k] = DIM == 3 ? Z[j] : T(0);
    }
  }

  // ---- stage 1: group extrema
  T mn[3], mx[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) { mn[d] = Prec<T>::lo(); mx[d] = -Prec<T>::lo(); }
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    mn[0] = dmin(mn[0], px[k]); mx[0] = dmax(mx[0], px[k]);
    mn[1] = dmin(mn[1], py[k]); mx[1] = dmax(mx[1], py[k]);
    if (DIM == 3) { mn[2] = dmin(mn[2], pz[k]); mx[2] = dmax(mx[2], pz[k]); }
  }
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    if constexpr (sizeof(T) == 4) {
      mn[d] = wave_ext<false>(mn[d]);
      mx[d] = wave_ext<true>(mx[d]);
    } else {
      // wave-uniform: kept in scalar registers through the sort
      mn[d] = uniform_f64(wave_min(mn[d]));
      mx[d] = uniform_f64(wave_max(mx[d]));
    }
  }
  if (DIM == 2) { mn[2] = T(0); mx[2] = T(0); }

  // ---- stage 2: quantise, block id, relative position
  const T inv = T(1) / (T(2) * eb);
  const T seg2eb = static_cast<T>(2u * seg) * eb;
  Q nbx, nby;
  block_counts<T, DIM>(mn[0], mx[0], mn[1], mx[1], seg2eb, &nbx, &nby);
  const uint32_t nbx32 = static_cast<uint32_t>(nbx);
  const uint32_t nbxy32 = static_cast<uint32_t>(nbx) * static_cast<uint32_t>(nby);
  const T nmi[3] = {-mn[0] * inv, -mn[1] * inv, -mn[2] * inv};
  const uint32_t sh = static_cast<uint32_t>(__builtin_ctz(seg));
  // The fast path needs every point's shifted code below the code of the
  // axis maximum, which holds while x -> fma(x, inv, nmi) is monotone
  // (inv finite and positive).
  bool fast = ((seg & (seg - 1u)) == 0u) && sh <= (DIM == 3 ? 10u : 15u) &&
              inv > T(0) && inv <= Prec<T>::lo();
  if (fast) {
    const uint64_t s0 = static_cast<uint64_t>(quant_q<T>(mx[0], inv, nmi[0]) >> sh);
    const uint64_t s1 = static_cast<uint64_t>(quant_q<T>(mx[1], inv, nmi[1]) >> sh);
    const uint64_t s2 = DIM == 3 ? static_cast<uint64_t>(quant_q<T>(mx[2], inv, nmi[2]) >> sh) : 0;
    const uint64_t bx = static_cast<uint64_t>(nbx);
    const uint64_t bxy = DIM == 3 ? bx * static_cast<uint64_t>(nby) : 0;
    fast = ((s0 | s1 | s2 | bx | bxy) >> 24) == 0 &&
           s0 + s1 * bx + s2 * bxy <= 0x7FFFFFFFull;
  }
  uint32_t key[16], val[16], bor, ror;
  if constexpr (sizeof(T) == 8) {
    const GPZ_GLOBAL T* X = gbl(static_cast<const T*>(a.x)) + base;
    const GPZ_GLOBAL T* Y = gbl(static_cast<const T*>(a.y)) + base;
    const GPZ_GLOBAL T* Z = DIM == 3 ? gbl(static_cast<const T*>(a.z)) + base : nullptr;
    if (fast)
      quantise_axes<T, DIM, kFull, true>(X, Y, Z, inv, nmi, seg, nbx32, nbxy32, jbase,
                                         nvalid, key, val, &bor, &ror);
    else
      quantise_axes<T, DIM, kFull, false>(X, Y, Z, inv, nmi, seg, nbx32, nbxy32, jbase,
                                          nvalid, key, val, &bor, &ror);
  } else {
    if (fast)
      quantise<T, DIM, kFull, true>(px, py, pz, inv, nmi, seg, nbx32, nbxy32, jbase,
                                    nvalid, key, val, &bor, &ror);
    else
      quantise<T, DIM, kFull, false>(px, py, pz, inv, nmi, seg, nbx32, nbxy32, jbase,
                                     nvalid, key, val, &bor, &ror);
  }
  const uint32_t bidbits = eff_bits(wave_or(bor));
  const uint32_t rb = eff_bits(wave_or(ror));

  // ---- stage 3: stable sort by block id over [0, sortbits), ties in
  // blocked position order.  With <= 22 key bits the position rides in the
  // low 10 bits of one word and the values stay put, indexed by position.
  const uint32_t sortbits = bidbits + (last ? 1u : 0u);
  const ScanLane sl = scan_lane(lane);
  if (sortbits == 0) {
    // one block id: the blocked order is already sorted
  } else if (sortbits <= 22) {
#pragma unroll
    for (int k = 0; k < 16; ++k) key[k] = (key[k] << 10) | (16u * lane + k);
    for (uint32_t shift = 10; shift < 10 + sortbits; shift += 4)
      radix_pass<false>(key, val, shift, lane, sl, lds);
    wave_sync();
    // values by position in 17-word rows: the stores (lane l writes row l)
    // and a gather of nearly sorted positions are free of bank conflicts
    uint32_t* s_val = reinterpret_cast<uint32_t*>(lds + kGatherB);
#pragma unroll
    for (int k = 0; k < 16; ++k) s_val[17 * lane + k] = val[k];
    wave_sync();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      val[k] = *reinterpret_cast<const uint32_t*>(lds + kGatherB + pad_byte(key[k] & 1023u));
      key[k] >>= 10;
    }
    // invalid ranks keep all-ones block ids (never compared below)
  } else if (((sortbits - 19u) & ~3u) + rb <= 32u) {
    // 23..32 key bits: pair passes (keys and values move) over the lowest
    // `lo` bits only, until <= 22 bits are left; then the packed passes
    // with the rank after the pair passes in the low 10 bits.  The low
    // block-id bits ride above the rb value bits meanwhile (rb + lo <= 32),
    // and one gather by that rank brings both back.
    const uint32_t lo = (sortbits - 19u) & ~3u;   // 4, 8 or 12
    for (uint32_t shift = 0; shift < lo; shift += 4)
      radix_pass<true>(key, val, shift, lane, sl, lds);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      val[k] |= (key[k] & ((1u << lo) - 1u)) << rb;
      key[k] = ((key[k] >> lo) << 10) | (16u * lane + k);
    }
    for (uint32_t shift = 10; shift < 10 + sortbits - lo; shift += 4)
      radix_pass<false>(key, val, shift, lane, sl, lds);
    wave_sync();
    uint32_t* s_val = reinterpret_cast<uint32_t*>(lds + kGatherB);
#pragma unroll
    for (int k = 0; k < 16; ++k) s_val[17 * lane + k] = val[k];
    wave_sync();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const uint32_t v =
          *reinterpret_cast<const uint32_t*>(lds + kGatherB + pad_byte(key[k] & 1023u));
      key[k] = ((key[k] >> 10) << lo) | (v >> rb);
      val[k] = v & ((1u << rb) - 1u);
    }
    // invalid ranks: all-ones values in the rb bits the planes use; their
    // block ids are never compared below
  } else {
    for (uint32_t shift = 0; shift < sortbits; shift += 4)
      radix_pass<true>(key, val, shift, lane, sl, lds);
  }
  wave_sync();

  // ---- stage 4a: runs of equal block id (lane l holds ranks 16l..16l+15).
  // Run u's delta and start rank go to u32 / u16 tables; ranks that start
  // no run write to a per-lane dump slot instead of branching.
  uint32_t* s_rdelta = reinterpret_cast<uint32_t*>(lds + kRunDeltaB);
  uint16_t* s_rstart = reinterpret_cast<uint16_t*>(lds + kRunStartB);
  const uint32_t up = dpp<kDppWaveShr1>(0u, key[15]);
  const uint32_t prev0 = lane == 0 ? key[0] : up;
  bool f[16];
  uint32_t nf = 0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    f[k] = key[k] != (k ? key[k - 1] : prev0);
    if (k == 0) f[0] = f[0] || lane == 0;
    if (!kFull) f[k] = f[k] && (16u * lane + k < nvalid);
    nf += f[k] ? 1u : 0u;
  }
  const uint32_t fincl = wave_scan(nf);
  const uint32_t U = lane63(fincl);
  const bool small = U <= 64;
  // Every valid rank starts its own run: deltas and counts (all 1) stay in
  // registers and the run tables are not built.
  const bool distinct = !small && U == nvalid;
  if (!distinct) {
    // deltas past U read 0 and starts past U read nvalid, so that counts
    // past U come out 0
    const uint32_t nv2 = nvalid | (nvalid << 16);
    uint4* d4 = reinterpret_cast<uint4*>(s_rdelta);
    uint4* s4 = reinterpret_cast<uint4*>(s_rstart);
    const uint4 z4 = make_uint4(0, 0, 0, 0), n4 = make_uint4(nv2, nv2, nv2, nv2);
    if (small) {
      if (lane < 16) d4[lane] = z4;
      if (lane < 9) s4[lane] = n4;
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) d4[64 * q + lane] = z4;
      s4[lane] = n4;
      s4[64 + lane] = n4;
      if (lane < 2) s4[128 + lane] = n4;
    }
    uint32_t ridx = fincl - nf;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const uint32_t w = f[k] ? ridx : kRunDump + lane;
      s_rdelta[w] = key[k] - (k ? key[k - 1] : prev0);  // run 0: prev0 = key[0]
      s_rstart[w] = static_cast<uint16_t>(16u * lane + k);
      ridx += f[k] ? 1u : 0u;
    }
  }
  const uint32_t outlier = static_cast<uint32_t>(
      __builtin_amdgcn_readfirstlane(static_cast<int>(key[0])));
  wave_sync();

  // ---- stage 4b: delta and count planes, MSB-first over run index
  GPZ_GLOBAL uint8_t* slot = gbl(a.slots) + c * static_cast<uint64_t>(kSlot);
  uint32_t db, cb, D, C, R;
  if (small) {
    // One run per lane, lane l holding run l^7: a ballot of bit b is then
    // plane b with every byte already MSB-first.  Lane j < B writes byte j
    // of each plane into the stage, which is copied out as dwords.
    const uint32_t u = lane ^ 7u;
    const uint32_t d = s_rdelta[u];
    const uint32_t cn = static_cast<uint32_t>(s_rstart[u + 1]) - s_rstart[u];
    db = eff_bits(wave_or(d));
    cb = eff_bits(wave_or(cn));
    section_sizes<T, DIM>(U, db, cb, rb, &D, &C, &R);
    const uint32_t B = (U + 7u) >> 3;
    uint8_t* st = lds + kStageB;
    uint32_t* st32 = reinterpret_cast<uint32_t*>(st);
    const uint32_t nw = (D + C) >> 2;
    if (lane < nw) st32[lane] = 0;
    if (lane + 64 < nw) st32[lane + 64] = 0;
    const uint32_t sb = 8u * lane;
    uint32_t addr = (db - 1u) * B + lane;  // plane db-1-b, byte `lane`
#pragma unroll
    for (int b = 0; b < 32; ++b) {
      if (b < static_cast<int>(db)) {
        const uint64_t pm = __builtin_amdgcn_ballot_w64(((d >> b) & 1u) != 0u);
        if (lane < B) st[addr] = static_cast<uint8_t>(pm >> sb);
        addr -= B;
      }
    }
    addr = D + (cb - 1u) * B + lane;
#pragma unroll
    for (int b = 0; b < 16; ++b) {
      if (b < static_cast<int>(cb)) {
        const uint64_t pm = __builtin_amdgcn_ballot_w64(((cn >> b) & 1u) != 0u);
        if (lane < B) st[addr] = static_cast<uint8_t>(pm >> sb);
        addr -= B;
      }
    }
    wave_sync();
    GPZ_GLOBAL uint32_t* sl32 = reinterpret_cast<GPZ_GLOBAL uint32_t*>(slot);
    if (lane < nw) sl32[lane] = st32[lane];
    if (lane + 64 < nw) sl32[lane + 64] = st32[lane + 64];
  } else {
    uint32_t del[16], cnt[16];
    if (distinct) {
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const bool v = kFull || 16u * lane + k < nvalid;
        del[k] = v ? key[k] - (k ? key[k - 1] : prev0) : 0u;
        cnt[k] = v ? 1u : 0u;
      }
    } else {
      const uint4* d4 = reinterpret_cast<const uint4*>(s_rdelta) + 4 * lane;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint4 w = d4[q];
        del[4 * q] = w.x; del[4 * q + 1] = w.y; del[4 * q + 2] = w.z; del[4 * q + 3] = w.w;
      }
      const uint4* s4 = reinterpret_cast<const uint4*>(s_rstart) + 2 * lane;
      const uint4 a0 = s4[0], a1 = s4[1];
      const uint32_t sw[9] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w,
                              static_cast<uint32_t>(s_rstart[16 * lane + 16])};
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        cnt[2 * i] = (sw[i] >> 16) - (sw[i] & 0xFFFFu);
        cnt[2 * i + 1] = (sw[i + 1] & 0xFFFFu) - (sw[i] >> 16);
      }
    }
    uint32_t dor = 0, cor = 0;
#pragma unroll
    for (int t = 0; t < 16; ++t) { dor |= del[t]; cor |= cnt[t]; }
    db = eff_bits(wave_or(dor));
    cb = eff_bits(wave_or(cor));
    section_sizes<T, DIM>(U, db, cb, rb, &D, &C, &R);
    const uint32_t B = (U + 7u) >> 3;
    store_msb_planes<32>(del, db, B, D, slot, lane);
    store_msb_planes<16>(cnt, cb, B, C, slot + D, lane);
  }

  // ---- stage 4c: relative-position planes, LSB-first over sorted rank
  {
    GPZ_GLOBAL uint8_t* pr = slot + D + C;
    if (rb <= 16) {
      uint32_t r[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) r[i] = pack_pair(val[i], val[i + 8]);
      transpose16_bytes(r);
#pragma unroll
      for (int b = 0; b < 16; ++b) {
        if (b < static_cast<int>(rb))
          *reinterpret_cast<GPZ_GLOBAL uint16_t*>(pr + (rb - 1u - b) * 128u + 2u * lane) =
              static_cast<uint16_t>(plane_word16(r, b));
      }
    } else {
      uint32_t t16[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) t16[k] = val[k];
      transpose16x2(t16);
#pragma unroll
      for (int b = 0; b < 32; ++b) {
        if (b < static_cast<int>(rb))
          *reinterpret_cast<GPZ_GLOBAL uint16_t*>(pr + (rb - 1u - b) * 128u + 2u * lane) =
              static_cast<uint16_t>(plane_word(t16, b));
      }
    }
  }

  // ---- meta and size
  {
    constexpr uint32_t M = Prec<T>::kMeta;
    uint32_t word = 0;
    if (sizeof(T) == 4) {
      if (lane < 6) {
        const T v = lane == 0 ? mn[0] : lane == 1 ? mn[1] : lane == 2 ? mn[2]
                  : lane == 3 ? mx[0] : lane == 4 ? mx[1] : mx[2];
        word = __float_as_uint(static_cast<float>(v));
      } else if (lane == 6) {
        word = outlier;
      } else if (lane == 7) {
        word = db | (cb << 16);
      } else if (lane == 8) {
        word = rb | (U << 16);
      }
    } else {
      if (lane < 12) {
        const uint32_t e = lane >> 1;
        const T v = e == 0 ? mn[0] : e == 1 ? mn[1] : e == 2 ? mn[2]
                  : e == 3 ? mx[0] : e == 4 ? mx[1] : mx[2];
        const uint64_t u = static_cast<uint64_t>(__double_as_longlong(static_cast<double>(v)));
        word = (lane & 1) ? static_cast<uint32_t>(u >> 32) : static_cast<uint32_t>(u);
      } else if (lane == 12) {
        word = outlier;
      } else if (lane == 14) {
        word = db | (cb << 16);
      } else if (lane == 15) {
        word = rb | (U << 16);
      }
    }
    if (lane < M / 4)
      reinterpret_cast<GPZ_GLOBAL uint32_t*>(gbl(a.meta) + c * M)[lane] = word;
  }
  if (lane == 0) gbl(a.loc)[c] = D + C + R;
}

// One wave per group.  (A variant that prefetched the next group's
// coordinates into registers measured slower: the doubled register set
// halves occupancy.)
// Group c of one snapshot by one wave; lds is the wave's own kLdsBytes.
// The group size comes from 32-bit values: the last group's size is
// precomputed on the host.
template <typename T, int DIM>
__device__ __forceinline__ void encode_one(const EncodeArgs& a, uint32_t c,
                                           uint32_t lane, uint8_t* lds) {
  if (c + 1u < static_cast<uint32_t>(a.G) || a.last_n == kGroup)
    encode_group<T, DIM, true>(a, c, lane, lds);
  else
    encode_group<T, DIM, false>(a, c, lane, lds);
}

// FP32 3D at 4 waves per SIMD (its wide-key paths spill; the common
// <= 22-bit path does not); FP32 2D, whose uniform groups mostly take the
// wide-key path, and FP64 at 3 without spills.
template <typename T, int DIM>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && DIM == 3 ? 4 : 3)))
void encode_kernel(EncodeArgs a) {
  __shared__ __align__(16) uint8_t lds[kLdsBytes];
  if (a.info->status != GPZ_OK) return;
  encode_one<T, DIM>(a, blockIdx.x, threadIdx.x, lds);
}

// ------------------------------------------------- offsets scan + header
// Exclusive scan of the G payload sizes (the reference's CUB DeviceScan,
// src/cudaWarpCompress.cu:1658-1666) in three steps over blocks of 1024
// groups: block sums, a scan of the block sums (which also writes the
// header and the total size), and a down-sweep writing every offset.
constexpr uint32_t kScanBlock = 1024;

struct ScanArgs {
  const uint64_t* loc;
  uint64_t* off;
  uint64_t* bsum;
  uint8_t* stream;
  DevInfo* info;
  uint64_t n, G;
  uint32_t dim;
  double eb;
};

// Block-wide (256 threads) exclusive scan; *total receives the block sum.
__device__ __forceinline__ uint64_t block_excl_scan256(uint64_t v,
                                                      uint64_t* total,
                                                      uint64_t* sh) {
  const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
  const uint64_t incl = wave_incl_scan(v, lane);
  if (lane == 63) sh[w] = incl;
  __syncthreads();
  uint64_t base = 0, tot = 0;
#pragma unroll
  for (uint32_t i = 0; i < 4; ++i) {
    base += i < w ? sh[i] : 0;
    tot += sh[i];
  }
  *total = tot;
  return base + incl - v;
}

__global__ __launch_bounds__(256) void scan_reduce_kernel(ScanArgs a) {
  __shared__ uint64_t sh[4];
  if (a.info->status != GPZ_OK) return;
  const uint64_t i0 = static_cast<uint64_t>(blockIdx.x) * kScanBlock + 4 * threadIdx.x;
  uint64_t v = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) v += i0 + i < a.G ? a.loc[i0 + i] : 0;
  uint64_t tot;
  block_excl_scan256(v, &tot, sh);
  if (threadIdx.x == 0) a.bsum[blockIdx.x] = tot;
}

// Stream header and total size (src/cudaWarpCompress.cu:1668-1683, FP64
// src/cudaWarpCompressDouble.cu:1667-1683); `payload` = sum of all loc.
template <typename T>
__device__ __forceinline__ void write_header(const ScanArgs& a, uint64_t payload) {
  constexpr uint32_t P = Prec<T>::kHeader, M = Prec<T>::kMeta;
  a.info->stream_bytes = P + (8ull + M) * a.G + payload;
  uint32_t* h = reinterpret_cast<uint32_t*>(a.stream);
  h[0] = a.dim;
  h[1] = static_cast<uint32_t>(a.n);
  h[2] = static_cast<uint32_t>(a.n >> 32);
  h[3] = sizeof(T) * 8;
  if (sizeof(T) == 4) {
    h[4] = __float_as_uint(static_cast<float>(a.eb));
    h[5] = a.info->seg_size;
  } else {
    const uint64_t u = static_cast<uint64_t>(__double_as_longlong(a.eb));
    h[4] = static_cast<uint32_t>(u);
    h[5] = static_cast<uint32_t>(u >> 32);
    h[6] = a.info->seg_size;
    h[7] = 32;   // parameter_size, src/cudaWarpCompressDouble.cu:1674-1680
  }
}

template <typename T>
__global__ __launch_bounds__(256) void scan_top_kernel(ScanArgs a, uint32_t nb) {
  __shared__ uint64_t sh[4];
  if (a.info->status != GPZ_OK) return;
  uint64_t carry = 0;
  for (uint32_t base = 0; base < nb; base += 256) {
    const uint32_t i = base + threadIdx.x;
    const uint64_t v = i < nb ? a.bsum[i] : 0;
    uint64_t tot;
    const uint64_t ex = block_excl_scan256(v, &tot, sh);
    if (i < nb) a.bsum[i] = carry + ex;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) write_header<T>(a, carry);
}

template <typename T>
__global__ __launch_bounds__(256) void scan_down_kernel(ScanArgs a) {
  __shared__ uint64_t sh[4];
  constexpr uint32_t P = Prec<T>::kHeader;
  if (a.info->status != GPZ_OK) return;
  const uint64_t i0 = static_cast<uint64_t>(blockIdx.x) * kScanBlock + 4 * threadIdx.x;
  uint64_t v[4], s = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    v[i] = i0 + i < a.G ? a.loc[i0 + i] : 0;
    s += v[i];
  }
  uint64_t tot;
  uint64_t run = a.bsum[blockIdx.x] + block_excl_scan256(s, &tot, sh);
  uint64_t* soff = reinterpret_cast<uint64_t*>(a.stream + P);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (i0 + i < a.G) {
      a.off[i0 + i] = run;
      soff[i0 + i] = run;
    }
    run += v[i];
  }
}

struct CompactArgs {
  const uint64_t* loc;
  const uint64_t* off;
  const uint8_t* slots;
  uint8_t* stream;
  const DevInfo* info;
  uint64_t G;
};

// Payload compaction (src/cudaWarpCompress.cu:32-56): every wave copies
// whole groups from their fixed slots to their packed offsets, grid-stride,
// eight dword loads in flight per lane.
template <typename T>
__device__ __forceinline__ void compact_group(const CompactArgs& a, uint64_t c,
                                              uint32_t lane) {
  constexpr uint32_t P = Prec<T>::kHeader, M = Prec<T>::kMeta;
  GPZ_GLOBAL uint8_t* pay = gbl(a.stream) + P + (8ull + M) * a.G;
  const uint32_t words = static_cast<uint32_t>(gbl(a.loc)[c]) >> 2;
  const GPZ_GLOBAL uint32_t* src =
      reinterpret_cast<const GPZ_GLOBAL uint32_t*>(gbl(a.slots) + c * static_cast<uint64_t>(kSlot));
  GPZ_GLOBAL uint32_t* dst = reinterpret_cast<GPZ_GLOBAL uint32_t*>(pay + gbl(a.off)[c]);
  for (uint32_t i0 = 0; i0 < words; i0 += 512) {
    uint32_t v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t i = i0 + 64u * k + lane;
      v[k] = i < words ? src[i] : 0u;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t i = i0 + 64u * k + lane;
      if (i < words) dst[i] = v[k];
    }
  }
}
template <typename T>
__global__ __launch_bounds__(256) void compact_kernel(CompactArgs a) {
  if (a.info->status != GPZ_OK) return;
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wave = (static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x) >> 6;
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * 4;
  for (uint64_t c = wave; c < a.G; c += nwaves) compact_group<T>(a, c, lane);
}

// ------------------------------------------------------------- decoder
struct DecodeArgs {
  const uint8_t* stream;
  uint64_t stream_bytes;
  uint64_t n, G;
  uint32_t last_n;  // points in group G-1 (1..1024)
  double eb;
  uint32_t seg;
  void* x;
  void* y;
  void* z;
};

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Point coordinate from its doubled code q2 = (q << 1) | 1
// (src/cudaWarpCompress.cu:1003-1010): fma(q2, eb, min).
template <typename T, typename Q>
__device__ __forceinline__ T dequant(Q q2, T eb, T mn) {
  return dfma(static_cast<T>(q2), eb, mn);
}

// One wave per group, reading the group's planes straight from the
// stream.  Groups of at most 64 runs (the common case for clustered data)
// hold one run per lane: each lane picks its bit out of every MSB-first
// plane byte, block ids and start ranks are lane prefix sums, the block
// coordinates of every run go to a 64-entry LDS table and every rank finds
// its run by counting run starts in a 1024-bit map.  Larger groups undo
// the encoder's 16x16 bit transposes (lane l: runs 16l..16l+15), keep one
// block id per run in LDS and split it per point with per-group magic
// divisors.  Either way lane l writes the 16 consecutive points of ranks
// 16l..16l+15.
// LDS of one decode wave: [0, 1024*sizeof(Q)) block id per run, or the
// small-group table (four Q per run); then 1024 run-start bytes, or the
// small-group bitmap.  The output stage reuses all of it.
template <typename T> constexpr uint32_t dec_lds_bytes() {
  return kGroup * sizeof(typename Prec<T>::Q) + kGroup;
}
// The small-group delta/count bytes (256) sit after the 64-run table.
template <typename T> constexpr uint32_t kDecSmallDcB() {
  return 4u * 64u * sizeof(typename Prec<T>::Q);
}
static_assert(4u * 64u * 4u + 256u <= 1024u * 4u && 4u * 64u * 8u + 256u <= 1024u * 8u,
              "small-group delta/count bytes stay below the run-start map");

// Values of reads issued under separate branches, pinned where all of them
// have started: the compiler cannot sink their first use (and with it a
// wait for the read) into each branch.
template <int N>
__device__ __forceinline__ void pin_reads(uint32_t (&v)[N]) {
#pragma unroll
  for (int i = 0; i < N; ++i) asm volatile("" : "+v"(v[i]));
}

// Lane l's bytes 2l and 2l+1 (b0 | b1 << 8) of the MSB-first planes
// b0 + j, j < 16, of a section with nb planes of B bytes: plane b sits at
// (nb-1-b)*B; planes at or past nb read 0.  Each read starts under its own
// uniform `b < nb` and nothing uses it before the last one has started, so
// the reads are in flight together (a use inside the branch makes every
// plane wait for the one before).  With B even the two bytes are one
// aligned 16-bit read; lanes past the plane (2l >= B) re-read byte 0.
__device__ __forceinline__ void read_plane_pairs(const GPZ_GLOBAL uint8_t* sec, uint32_t nb, uint32_t B,
                                                 uint32_t lane, int b0, uint32_t (&h)[16]) {
  const uint32_t o0 = 2u * lane < B ? 2u * lane : 0u;
  const uint32_t o1 = 2u * lane + 1u < B ? 2u * lane + 1u : 0u;
#pragma unroll
  for (int j = 0; j < 16; ++j) h[j] = 0;
  if ((B & 1u) == 0u) {
#pragma unroll
    for (int j = 0; j < 16; ++j)
      if (b0 + j < static_cast<int>(nb))
        h[j] = *reinterpret_cast<const GPZ_GLOBAL uint16_t*>(sec + (nb - 1u - (b0 + j)) * B + o0);
    pin_reads(h);
  } else {
    uint32_t hi[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) hi[j] = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j)
      if (b0 + j < static_cast<int>(nb)) {
        h[j] = sec[(nb - 1u - (b0 + j)) * B + o0];
        hi[j] = sec[(nb - 1u - (b0 + j)) * B + o1];
      }
    pin_reads(h);
    pin_reads(hi);
#pragma unroll
    for (int j = 0; j < 16; ++j) h[j] |= hi[j] << 8;
  }
}

template <typename T, int DIM>
__device__ __forceinline__ void decode_group(const DecodeArgs& a, uint32_t c32,
                                             uint32_t lane, uint8_t* lds) {
  using Q = typename Prec<T>::Q;
  constexpr uint32_t P = Prec<T>::kHeader, M = Prec<T>::kMeta;
  constexpr uint32_t kStartB = kGroup * sizeof(Q);
  Q* s_bid = reinterpret_cast<Q*>(lds);
  Q* s_tab = reinterpret_cast<Q*>(lds);
  uint8_t* s_start = lds + kStartB;
  uint32_t* s_map = reinterpret_cast<uint32_t*>(lds + kStartB);

  const uint64_t c = c32;
  const GPZ_GLOBAL uint8_t* stream = gbl(a.stream);
  const GPZ_GLOBAL uint8_t* meta = stream + P + 8 * a.G + M * c;
  T mm[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) mm[i] = reinterpret_cast<const GPZ_GLOBAL T*>(meta)[i];
  const Q outlier = *reinterpret_cast<const GPZ_GLOBAL Q*>(meta + 6 * sizeof(T));
  const GPZ_GLOBAL uint16_t* u16 =
      reinterpret_cast<const GPZ_GLOBAL uint16_t*>(meta + 6 * sizeof(T) + sizeof(Q));
  const uint32_t db = u16[0], cb = u16[1], rb = u16[2], U = u16[3];
  const uint64_t off = reinterpret_cast<const GPZ_GLOBAL uint64_t*>(stream + P)[c];
  uint32_t D, C, R;
  // FP64-2D streams carry a 4-byte-padded count section (encoder layout).
  section_sizes<T, DIM>(U, db, cb, rb, &D, &C, &R);
  // A group whose sections do not fit the stream or the format's limits
  // (corrupt input) is left undecoded rather than read out of bounds.
  const uint64_t pay0 = P + (8ull + M) * a.G + off;
  if (U == 0 || U > kGroup || db > 32 || cb > 16 || rb > 32 ||
      pay0 + D + C + R > a.stream_bytes)
    return;
  const GPZ_GLOBAL uint8_t* pay = stream + pay0;
  const uint32_t B = (U + 7u) >> 3;

  // Groups of <= 64 runs with <= 16 delta bits hold one run per lane.
  // Their delta and count sections (D + C <= 216 bytes) are read here as
  // one word per lane, before the relative-position planes, so that both
  // reads are in flight at once; the bytes are picked out through LDS
  // below.  Lanes past the sections re-read word 0 (C >= 4: every group
  // has a count plane).
  const bool small = U <= 64 && db <= 16;
  uint32_t dcw = 0;
  if (small)
    dcw = *reinterpret_cast<const GPZ_GLOBAL uint32_t*>(pay + (4u * lane < D + C ? 4u * lane : 0u));

  // ---- relative positions of ranks 16l..16l+15 (LSB-first planes, one
  // 16-bit word per lane and plane)
  uint32_t rel[16];
  {
    const GPZ_GLOBAL uint8_t* pr = pay + D + C + 2u * lane;
    auto plane16 = [&](int b) -> uint32_t {
      return *reinterpret_cast<const GPZ_GLOBAL uint16_t*>(pr + (rb - 1u - b) * 128u);
    };
    if (rb <= 16) {
      uint32_t w[16];
#pragma unroll
      for (int b = 0; b < 16; ++b) w[b] = b < static_cast<int>(rb) ? plane16(b) : 0u;
      uint32_t r[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) r[i] = pack_pair(w[i], w[i + 8]);
      transpose16_bytes(r);
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        rel[i] = r[i] & 0xFFFFu;
        rel[i + 8] = r[i] >> 16;
      }
    } else {
      // reads under `b < rb`, first used after the last one (see
      // read_plane_pairs)
#pragma unroll
      for (int k = 0; k < 16; ++k) rel[k] = 0;
      uint32_t hw[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) hw[k] = 0;
#pragma unroll
      for (int b = 0; b < 16; ++b) rel[b] = plane16(b);
#pragma unroll
      for (int b = 16; b < 32; ++b)
        if (b < static_cast<int>(rb)) hw[b - 16] = plane16(b);
      pin_reads(hw);
#pragma unroll
      for (int k = 0; k < 16; ++k) rel[k] |= hw[k] << 16;
      transpose16x2(rel);
    }
  }

  const T eb = static_cast<T>(a.eb);
  const uint32_t seg = a.seg;
  const T seg2eb = static_cast<T>(2u * seg) * eb;
  Q nbx, nby;
  block_counts<T, DIM>(mm[0], mm[3], mm[1], mm[4], seg2eb, &nbx, &nby);
  const uint64_t base = c * kGroup;
  const uint32_t nvalid =
      c32 + 1u < static_cast<uint32_t>(a.G) ? kGroup : a.last_n;
  const uint32_t sh = static_cast<uint32_t>(__builtin_ctz(seg));
  const Q m = static_cast<Q>(seg - 1u);
  const Q sq = static_cast<Q>(seg);
  const Q nxy = static_cast<Q>(nbx * nby);
  const Div32 dvx = div32_make(static_cast<uint32_t>(nbx ? nbx : 1));
  const Div32 dvy = div32_make(static_cast<uint32_t>(nby ? nby : 1));
  const Div32 dvxy = div32_make(static_cast<uint32_t>(nxy ? nxy : 1));
  // block coordinates of block id b, each times seg
  auto split = [&](Q b, Q* bx, Q* by, Q* bz) {
    const Q qb = qdiv<Q>(b, nbx, dvx);            // b / nbx
    *bx = (nbx ? b - qb * nbx : b) * sq;          // b % nbx
    *by = (nby ? qb - qdiv<Q>(qb, nby, dvy) * nby : qb) * sq;
    *bz = DIM == 3 ? qdiv<Q>(b, nxy, dvxy) * sq : Q(0);
  };
  auto inc_scan = [&](Q v) -> Q {
    if constexpr (sizeof(Q) == 4) return wave_scan(v);
    else return static_cast<Q>(wave_incl_scan(static_cast<uint64_t>(v), lane));
  };

  T ox[16], oy[16], oz[16];
  if (small) {
    // ---- one run per lane.  Plane bytes past the plane (u >= 8B) only
    // feed runs >= U; they re-read byte 0 of the plane.
    const uint32_t u = lane;
    const bool valid = u < U;
    const uint32_t boff = u < 8u * B ? (u >> 3) : 0u;
    const uint32_t bsh = 7u - (u & 7u);
    uint8_t* s_dc = lds + kDecSmallDcB<T>();
    reinterpret_cast<uint32_t*>(s_dc)[lane] = dcw;
    wave_sync();
    uint32_t pd[16], pc[16];
#pragma unroll
    for (int b = 0; b < 16; ++b) {
      pd[b] = b < static_cast<int>(db) ? s_dc[(db - 1u - b) * B + boff] : 0u;
      pc[b] = b < static_cast<int>(cb) ? s_dc[D + (cb - 1u - b) * B + boff] : 0u;
    }
    uint32_t dv = 0, cv = 0;
#pragma unroll
    for (int b = 0; b < 16; ++b) {
      dv |= ((pd[b] >> bsh) & 1u) << b;
      cv |= ((pc[b] >> bsh) & 1u) << b;
    }
    // block id = outlier + deltas of runs 1..u (delta of run 0 ignored,
    // src/cudaWarpCompress.cu:930-935); start rank = counts of runs < u
    const Q dq = u == 0 ? outlier : (valid ? static_cast<Q>(dv) : Q(0));
    const Q bid = inc_scan(dq);
    const uint32_t cn = valid ? cv : 0u;
    const uint32_t st = wave_scan(cn) - cn;
    Q bx, by, bz;
    split(bid, &bx, &by, &bz);
    if (lane < 32) s_map[lane] = 0u;
    if (valid) {
      s_tab[4 * u + 0] = (bx << 1) | 1u;
      s_tab[4 * u + 1] = (by << 1) | 1u;
      s_tab[4 * u + 2] = (bz << 1) | 1u;
    }
    wave_sync();
    if (valid && st < kGroup)
      __hip_atomic_fetch_or(s_map + (st >> 5), 1u << (st & 31u), __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_WORKGROUP);
    wave_sync();
    // ---- rank -> run: starts at or before the rank, minus one
    const uint32_t w = (s_map[lane >> 1] >> (16u * (lane & 1u))) & 0xFFFFu;
    const uint32_t nw = __popc(w);
    const uint32_t ex = wave_scan(nw) - nw - 1u;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const uint32_t run = ex + __popc(w & ((2u << k) - 1u));
      const Q rl = static_cast<Q>(rel[k]);
      const Q* tb = s_tab + 4 * run;
      ox[k] = dequant<T>(tb[0] + ((rl & m) << 1), eb, mm[0]);
      oy[k] = dequant<T>(tb[1] + ((shr(rl, sh) & m) << 1), eb, mm[1]);
      if (DIM == 3) oz[k] = dequant<T>(tb[2] + (shr(rl, 2u * sh) << 1), eb, mm[2]);
    }
  } else {
    // ---- runs 16l..16l+15: lane l's two bytes of each MSB-first plane
    // (b0 | b1 << 16) through the encoder's transpose in reverse; slot s
    // then holds run 16l + (s^7).  Bytes past the plane (2l >= B) only feed
    // runs >= U; they re-read byte 0 to stay inside the stream.
    uint32_t del[16], cnt[16];
    if (db <= 16) {
      uint32_t w[16];
      read_plane_pairs(pay, db, B, lane, 0, w);
      uint32_t r[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) r[i] = __builtin_amdgcn_perm(w[i + 8], w[i], 0x05010400u);
      transpose16_bytes(r);
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        del[7 - i] = r[i] & 0xFFFFu;
        del[15 - i] = r[i] >> 16;
      }
    } else {
      uint32_t t16[16], hw[16];
      read_plane_pairs(pay, db, B, lane, 0, t16);
      read_plane_pairs(pay, db, B, lane, 16, hw);
#pragma unroll
      for (int t = 0; t < 16; ++t) t16[t] |= hw[t] << 16;
      transpose16x2(t16);
#pragma unroll
      for (int t = 0; t < 16; ++t) del[t ^ 7] = t16[t];
    }
    {
      uint32_t w[16];
      read_plane_pairs(pay + D, cb, B, lane, 0, w);
      uint32_t r[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) r[i] = __builtin_amdgcn_perm(w[i + 8], w[i], 0x05010400u);
      transpose16_bytes(r);
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        cnt[7 - i] = r[i] & 0xFFFFu;
        cnt[15 - i] = r[i] >> 16;
      }
    }
    // block id per run = outlier + prefix of deltas; run end = prefix of
    // counts
    reinterpret_cast<uint4*>(s_start)[lane] = make_uint4(0, 0, 0, 0);
    Q bsum = 0;
    uint32_t csum = 0;
    Q bid[16];
    uint32_t cend[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const uint32_t i = 16u * lane + t;
      const Q d = (i == 0) ? outlier : (i < U ? static_cast<Q>(del[t]) : Q(0));
      const uint32_t cn = i < U ? cnt[t] : 0u;
      bsum += d;
      csum += cn;
      bid[t] = bsum;
      cend[t] = csum;
    }
    const Q bpre = inc_scan(bsum) - bsum;
    const uint32_t cpre = wave_scan(csum) - csum;
    if (U == nvalid) {
      // every rank is its own run (all counts are 1): rank 16l+t has block
      // id bid[t] + bpre, no run lookup needed
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const Q rl = static_cast<Q>(rel[k]);
        Q bx, by, bz;
        split(static_cast<Q>(bid[k] + bpre), &bx, &by, &bz);
        ox[k] = dequant<T>((((bx + (rl & m)) << 1) | 1u), eb, mm[0]);
        oy[k] = dequant<T>((((by + (shr(rl, sh) & m)) << 1) | 1u), eb, mm[1]);
        if (DIM == 3) oz[k] = dequant<T>((((bz + shr(rl, 2u * sh)) << 1) | 1u), eb, mm[2]);
      }
    } else {
      wave_sync();
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        const uint32_t i = 16u * lane + t;
        if (i < U) {
          s_bid[i] = static_cast<Q>(bid[t] + bpre);
          const uint32_t st = cend[t] + cpre - cnt[t];
          if (st < kGroup) s_start[st] = 1;
        }
      }
      wave_sync();
      // ---- rank -> run index by a prefix count of run starts
      const uint4 fv = reinterpret_cast<const uint4*>(s_start)[lane];
      const uint32_t fw[4] = {fv.x, fv.y, fv.z, fv.w};
      uint32_t rs = 0;
      uint32_t runof[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        rs += (fw[k >> 2] >> (8 * (k & 3))) & 0xFFu;
        runof[k] = rs;
      }
      const uint32_t rpre = wave_scan(rs) - rs;
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        uint32_t run = runof[k] + rpre;
        run = run ? run - 1u : 0u;
        if (run >= U) run = U - 1u;
        const Q rl = static_cast<Q>(rel[k]);
        Q bx, by, bz;
        split(s_bid[run], &bx, &by, &bz);
        ox[k] = dequant<T>((((bx + (rl & m)) << 1) | 1u), eb, mm[0]);
        oy[k] = dequant<T>((((by + (shr(rl, sh) & m)) << 1) | 1u), eb, mm[1]);
        if (DIM == 3) oz[k] = dequant<T>((((bz + shr(rl, 2u * sh)) << 1) | 1u), eb, mm[2]);
      }
    }
  }

  // ---- store ranks 16l..16l+15 of each plane.  A full group goes
  // through LDS one axis at a time so that every store instruction writes
  // 64 consecutive 16-byte vectors (lane l: vectors l + 64j); with lanes
  // holding 16 consecutive points directly, each store would touch 32
  // cache lines.  Vector i lives at i + i/16 (no bank conflicts on either
  // side).
  if (nvalid == kGroup) {
    constexpr int V = 16 / sizeof(T);   // elements per 16-byte vector
    constexpr int NV = 16 / V;          // vectors per lane and axis
    static_assert(64 * NV * 17 / 16 * 16 <= kStartB + kGroup, "output stage fits");
    uint4* sv = reinterpret_cast<uint4*>(lds);
    auto put = [&](const T (&o)[16], void* dst) {
      wave_sync();
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        const uint32_t i = NV * lane + j;
        T e[V];
#pragma unroll
        for (int t = 0; t < V; ++t) e[t] = o[V * j + t];
        uint4 w;
        memcpy(&w, e, 16);
        sv[i + (i >> 4)] = w;
      }
      wave_sync();
      GPZ_GLOBAL uint4* d4 = reinterpret_cast<GPZ_GLOBAL uint4*>(gbl(static_cast<T*>(dst)) + base);
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        const uint32_t i = lane + 64u * j;
        // streaming stores: the output is written once and not read back
        // by this kernel, and cached copies would only be written back
        // while the next kernel runs
        const uint4 v = sv[i + (i >> 4)];
        const u32x4 nv = {v.x, v.y, v.z, v.w};
        __builtin_nontemporal_store(nv, reinterpret_cast<GPZ_GLOBAL u32x4*>(d4 + i));
      }
    };
    put(ox, a.x);
    put(oy, a.y);
    if (DIM == 3) put(oz, a.z);
  } else {
    GPZ_GLOBAL T* X = gbl(static_cast<T*>(a.x)) + base + 16u * lane;
    GPZ_GLOBAL T* Y = gbl(static_cast<T*>(a.y)) + base + 16u * lane;
    GPZ_GLOBAL T* Z = gbl(static_cast<T*>(a.z)) + base + 16u * lane;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (16u * lane + k < nvalid) {
        X[k] = ox[k];
        Y[k] = oy[k];
        if (DIM == 3) Z[k] = oz[k];
      }
    }
  }
}

template <typename T, int DIM>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 5 : 3)))
void decode_kernel(DecodeArgs a) {
  __shared__ __align__(16) uint8_t lds[dec_lds_bytes<T>()];
  decode_group<T, DIM>(a, blockIdx.x, threadIdx.x, lds);
}

// ------------------------------------------------------------- batches
// Many snapshots of uneven size, dimension and precision per call.  For
// each (precision, dimension) class the groups of all its snapshots form a
// device-side queue of segments (snapshot, first group, group count);
// persistent waves take segments w, w + waves, w + 2*waves, ... so that a
// launch covers every snapshot of the class.  The per-group work is the
// per-call path's own (encode_one, compact_group, decode_group), so a
// snapshot's stream is the same whether it is compressed alone or in a
// batch.

struct BatchItem {    // device descriptor of one snapshot to compress
  EncodeArgs e;
  ScanArgs sc;
  CompactArgs ca;
  SegArgs sa;
  DevInfo* info;
  void* win;
  uint64_t stride;    // sample window stride
};

struct Segment {
  uint32_t item, g0, ng, pad;
};

// Classes: 0 = f32 2D, 1 = f32 3D, 2 = f64 2D, 3 = f64 3D.
constexpr int kClasses = 4;
__host__ __device__ inline int class_of(uint32_t bits, uint32_t dim) {
  return (bits == 64 ? 2 : 0) + (dim == 3 ? 1 : 0);
}

__device__ __forceinline__ uint32_t wave_in_block() {
  return static_cast<uint32_t>(
      __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6)));
}

// One wave per sample window (4 points per lane), four windows per
// workgroup; same window extrema as seg_sample_window.
template <typename T, int DIM>
__global__ __launch_bounds__(256) void batch_seg_sample_kernel(const BatchItem* items,
                                                               const uint32_t* cls) {
  const BatchItem& it = items[cls[blockIdx.y]];
  if (!it.sa.sampled) return;
  const uint32_t lane = threadIdx.x & 63u, b = blockIdx.x * 4 + wave_in_block();
  T* win = static_cast<T*>(it.win);
  const uint64_t st = static_cast<uint64_t>(b) * it.stride;
  if (st + kSampleLen > it.e.n) {
    if (lane < 6) win[b * 6 + lane] = T(0);
    return;
  }
  const T* P3[3] = {static_cast<const T*>(it.e.x), static_cast<const T*>(it.e.y),
                    static_cast<const T*>(it.e.z)};
  T r = T(0);
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    T lo = Prec<T>::lo(), hi = -Prec<T>::lo();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const T v = P3[a][st + 64u * k + lane];
      lo = dmin(lo, v);
      hi = dmax(hi, v);
    }
    if constexpr (sizeof(T) == 4) {
      lo = wave_ext<false>(lo);
      hi = wave_ext<true>(hi);
    } else {
      lo = wave_min(lo);
      hi = wave_max(hi);
    }
    if (lane == 2u * a) r = lo;
    if (lane == 2u * a + 1u) r = hi;
  }
  if (lane < 6) win[b * 6 + lane] = r;  // 2D: z extrema stay 0
}

template <typename T, int DIM>
__global__ __launch_bounds__(256) void batch_seg_decide_kernel(const BatchItem* items,
                                                               const uint32_t* cls) {
  const BatchItem& it = items[cls[blockIdx.x]];
  seg_decide_one<T, DIM>(static_cast<const T*>(it.win), it.sa, it.info);
}

// The lane index, the LDS slice and the descriptor pointer are made opaque
// in every iteration of a persistent loop: otherwise the compiler hoists
// every lane-dependent value of the group body out of the loop and spills.
// The slice and the descriptor are made opaque as an LDS byte offset and a
// snapshot index, not as pointers: an opaque pointer loses its address
// space, and every LDS access of the group body, and every global access
// through the pointers of the descriptor, becomes a flat access.
#define GPZ_OPAQUE_ITER(item, ln, loff) asm volatile("" : "+s"(item), "+v"(ln), "+s"(loff))

// Up to 8 waves per workgroup, each with its own kLdsBytes of dynamic LDS.
// FP32 at four waves per SIMD, as the single-group kernel: the loop state
// spills some registers at 128, and the mix still ran 4 % faster than at
// three waves without spills.
template <typename T, int DIM>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1)))
void batch_encode_kernel(const BatchItem* __restrict__ items,
                         const Segment* __restrict__ segs, uint32_t nsegs) {
  extern __shared__ __align__(16) uint8_t dyn_lds[];
  const uint32_t lane = threadIdx.x & 63u, w = wave_in_block();
  const uint32_t nw = blockDim.x >> 6;
  const uint32_t lds_off = w * kLdsBytes;
  for (uint32_t s = blockIdx.x * nw + w; s < nsegs; s += gridDim.x * nw) {
    const Segment sg = segs[s];
    for (uint32_t g = sg.g0; g < sg.g0 + sg.ng; ++g) {
      uint32_t item = sg.item, ln = lane, loff = lds_off;
      GPZ_OPAQUE_ITER(item, ln, loff);
      const EncodeArgs* ap = &items[item].e;
      uint8_t* lp = dyn_lds + loff;
      if (ap->info->status != GPZ_OK) break;
      encode_one<T, DIM>(*ap, g, ln, lp);
    }
  }
}

// One workgroup per snapshot: offsets by a blocked exclusive scan of its
// payload sizes, then the header.
template <typename T>
__global__ __launch_bounds__(256) void batch_scan_kernel(const BatchItem* items,
                                                         const uint32_t* cls) {
  __shared__ uint64_t sh[4];
  constexpr uint32_t P = Prec<T>::kHeader;
  const ScanArgs& a = items[cls[blockIdx.x]].sc;
  if (a.info->status != GPZ_OK) return;
  uint64_t* soff = reinterpret_cast<uint64_t*>(a.stream + P);
  uint64_t carry = 0;
  for (uint64_t base = 0; base < a.G; base += kScanBlock) {
    const uint64_t i0 = base + 4 * threadIdx.x;
    uint64_t v[4], sum = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      v[i] = i0 + i < a.G ? a.loc[i0 + i] : 0;
      sum += v[i];
    }
    uint64_t tot;
    uint64_t run = carry + block_excl_scan256(sum, &tot, sh);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i0 + i < a.G) {
        a.off[i0 + i] = run;
        soff[i0 + i] = run;
      }
      run += v[i];
    }
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) write_header<T>(a, carry);
}

template <typename T>
__global__ __launch_bounds__(256) void batch_compact_kernel(const BatchItem* items,
                                                            const Segment* segs,
                                                            uint32_t nsegs) {
  const uint32_t lane = threadIdx.x & 63u, w = wave_in_block();
  for (uint32_t s = blockIdx.x * 4 + w; s < nsegs; s += gridDim.x * 4) {
    const Segment sg = segs[s];
    const CompactArgs& a = items[sg.item].ca;
    if (a.info->status != GPZ_OK) continue;
    for (uint32_t g = sg.g0; g < sg.g0 + sg.ng; ++g) compact_group<T>(a, g, lane);
  }
}

template <typename T, int DIM>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1)))
void batch_decode_kernel(const DecodeArgs* __restrict__ items,
                         const Segment* __restrict__ segs, uint32_t nsegs) {
  extern __shared__ __align__(16) uint8_t dyn_lds[];
  const uint32_t lane = threadIdx.x & 63u, w = wave_in_block();
  const uint32_t nw = blockDim.x >> 6;
  const uint32_t lds_off = w * dec_lds_bytes<T>();
  for (uint32_t s = blockIdx.x * nw + w; s < nsegs; s += gridDim.x * nw) {
    const Segment sg = segs[s];
    for (uint32_t g = sg.g0; g < sg.g0 + sg.ng; ++g) {
      uint32_t item = sg.item, ln = lane, loff = lds_off;
      GPZ_OPAQUE_ITER(item, ln, loff);
      const DecodeArgs* ap = &items[item];
      uint8_t* lp = dyn_lds + loff;
      decode_group<T, DIM>(*ap, g, ln, lp);
    }
  }
}

// ---------------------------------------------------------- min / max
template <typename T, int DIM>
__global__ __launch_bounds__(256) void minmax_kernel(const T* __restrict__ x,
                                                     const T* __restrict__ y,
                                                     const T* __restrict__ z,
                                                     uint64_t n, T* part) {
  __shared__ T red[4][6];
  const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
  T e[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) e[i] = (i & 1) ? -Prec<T>::lo() : Prec<T>::lo();
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * 256;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + t; i < n; i += stride) {
    const T vx = x[i], vy = y[i];
    e[0] = dmin(e[0], vx); e[1] = dmax(e[1], vx);
    e[2] = dmin(e[2], vy); e[3] = dmax(e[3], vy);
    if (DIM == 3) {
      const T vz = z[i];
      e[4] = dmin(e[4], vz); e[5] = dmax(e[5], vz);
    }
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    e[i] = (i & 1) ? wave_max(e[i]) : wave_min(e[i]);
    if (lane == 0) red[w][i] = e[i];
  }
  __syncthreads();
  if (t < 6) {
    T r = red[0][t];
    for (int k = 1; k < 4; ++k) r = (t & 1) ? dmax(r, red[k][t]) : dmin(r, red[k][t]);
    part[blockIdx.x * 6 + t] = r;
  }
}

// --------------------------------------------------------------- host
template <typename T> struct HostPrec;
template <> struct HostPrec<float> { static constexpr uint32_t P = 24, M = 36; };
template <> struct HostPrec<double> { static constexpr uint32_t P = 32, M = 64; };

inline uint64_t groups_of(uint64_t n) { return (n + kGroup - 1) / kGroup; }


bool params_ok(const gpz_params* p) {
  if (!p) return false;
  if (p->dim != 2 && p->dim != 3) return false;
  if (p->float_bits != 32 && p->float_bits != 64) return false;
  if (p->n == 0 || groups_of(p->n) > 0x7FFFFFFFull) return false;
  if (p->mode != GPZ_MODE_ABS && p->mode != GPZ_MODE_REL) return false;
  if (!(p->eb >= 0.0) || p->eb > 1.0e300) return false;
  // abs mode without an explicit segment size leaves seg = 0 in the
  // reference (src/cudaWarpCompress.cu:1513-1531), which breaks the bound;
  // it is rejected here.
  if (p->mode == GPZ_MODE_ABS && p->seg_size == 0) return false;
  return true;
}

template <typename T, int DIM>
int compress_impl(const gpz_params* p, const void* dx, const void* dy,
                  const void* dz, void* dstream, void* dws,
                  hipStream_t s) {
  const uint64_t n = p->n, G = groups_of(n);
  const WsLayout L = ws_layout(sizeof(T) * 8, G);
  uint8_t* ws = static_cast<uint8_t*>(dws);
  DevInfo* info = reinterpret_cast<DevInfo*>(ws + L.info);
  T* win = reinterpret_cast<T*>(ws + L.win);
  // FP32 always runs the sampler (its block-count guard can double any seg,
  // src/cudaWarpCompress.cu:1532-1554); FP64 only needs it to choose one.
  const bool sample = sizeof(T) == 4 || (p->seg_size == 0 && p->mode == GPZ_MODE_REL);
  if (sample) {
    const uint64_t stride = n - kSampleLen;  // u64 wrap for n < 256, as the reference
    uint64_t st = stride / kSampleBlocks;
    if (st < 1) st = 1;
    LaunchTimer lt("seg_sample", s);
    hipLaunchKernelGGL((seg_sample_kernel<T, DIM>), dim3(kSampleBlocks), dim3(256), 0, s,
                       static_cast<const T*>(dx), static_cast<const T*>(dy),
                       static_cast<const T*>(dz), n, st, win);
  }
  SegArgs sa;
  sa.seg_in = p->seg_size;
  sa.is_rel = p->mode == GPZ_MODE_REL;
  const double rel = sizeof(T) == 4 ? static_cast<double>(static_cast<float>(p->rel_eb))
                                    : p->rel_eb;
  sa.rel_small = rel <= 1e-4;
  sa.sampled = sample;
  sa.eb = sizeof(T) == 4 ? static_cast<double>(static_cast<float>(p->eb)) : p->eb;
  {
    LaunchTimer lt("seg_decide", s);
    hipLaunchKernelGGL((seg_decide_kernel<T, DIM>), dim3(1), dim3(256), 0, s, win, sa, info);
  }

  EncodeArgs ea;
  ea.x = dx; ea.y = dy; ea.z = dz;
  ea.n = n; ea.G = G; ea.eb = sa.eb;
  ea.last_n = static_cast<uint32_t>(n - (G - 1) * kGroup);
  ea.info = info;
  ea.meta = static_cast<uint8_t*>(dstream) + HostPrec<T>::P + 8 * G;
  ea.loc = reinterpret_cast<uint64_t*>(ws + L.loc);
  ea.slots = ws + L.slots;
  {
    LaunchTimer lt("encode", s);
    hipLaunchKernelGGL((encode_kernel<T, DIM>), dim3(static_cast<uint32_t>(G)),
                       dim3(64), 0, s, ea);
  }

  ScanArgs sc;
  sc.loc = ea.loc;
  sc.off = reinterpret_cast<uint64_t*>(ws + L.off);
  sc.bsum = reinterpret_cast<uint64_t*>(ws + L.bsum);
  sc.stream = static_cast<uint8_t*>(dstream);
  sc.info = info;
  sc.n = n; sc.G = G; sc.dim = DIM; sc.eb = sa.eb;
  const uint32_t nb = static_cast<uint32_t>((G + kScanBlock - 1) / kScanBlock);
  {
    LaunchTimer lt("scan", s);
    hipLaunchKernelGGL(scan_reduce_kernel, dim3(nb), dim3(256), 0, s, sc);
    hipLaunchKernelGGL((scan_top_kernel<T>), dim3(1), dim3(256), 0, s, sc, nb);
    hipLaunchKernelGGL((scan_down_kernel<T>), dim3(nb), dim3(256), 0, s, sc);
  }

  CompactArgs ca;
  ca.loc = ea.loc;
  ca.off = sc.off;
  ca.slots = ea.slots;
  ca.stream = sc.stream;
  ca.info = info;
  ca.G = G;
  {
    LaunchTimer lt("compact", s);
    const uint64_t wgs = (G + 3) / 4 < 4096 ? (G + 3) / 4 : 4096;
    hipLaunchKernelGGL((compact_kernel<T>), dim3(static_cast<uint32_t>(wgs)), dim3(256), 0, s, ca);
  }
  return hipGetLastError() == hipSuccess ? GPZ_OK : GPZ_ERR_HIP;
}

template <typename T, int DIM>
int decompress_impl(const gpz_header* h, const void* dstream,
                    uint64_t stream_bytes, void* dx, void* dy, void* dz,
                    hipStream_t s) {
  DecodeArgs da;
  da.stream = static_cast<const uint8_t*>(dstream);
  da.stream_bytes = stream_bytes;
  da.n = h->n;
  da.last_n = static_cast<uint32_t>(h->n - (h->groups - 1) * kGroup);
  da.G = h->groups;
  da.eb = h->eb;
  da.seg = h->seg_size;
  da.x = dx; da.y = dy; da.z = dz;
  {
    LaunchTimer lt("decode", s);
    hipLaunchKernelGGL((decode_kernel<T, DIM>), dim3(static_cast<uint32_t>(h->groups)),
                       dim3(64), 0, s, da);
  }
  return hipGetLastError() == hipSuccess ? GPZ_OK : GPZ_ERR_HIP;
}

template <typename T, int DIM>
int minmax_impl(uint64_t n, const void* dx, const void* dy, const void* dz,
                void* dws, void* mm, hipStream_t s) {
  T* part = static_cast<T*>(dws);
  hipLaunchKernelGGL((minmax_kernel<T, DIM>), dim3(kMinMaxBlocks), dim3(256), 0, s,
                     static_cast<const T*>(dx), static_cast<const T*>(dy),
                     static_cast<const T*>(dz), n, part);
  if (hipGetLastError() != hipSuccess) return GPZ_ERR_HIP;
  T h[kMinMaxBlocks * 6];
  if (hipMemcpyAsync(h, part, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return GPZ_ERR_HIP;
  // Same result as the reference's serial std::min/std::max sweep for
  // finite data (min/max are exact); the z extent is 0 in 2D.
  T r[6];
  for (int i = 0; i < 6; ++i) r[i] = h[i];
  for (uint32_t b = 1; b < kMinMaxBlocks; ++b)
    for (int i = 0; i < 6; ++i)
      r[i] = (i & 1) ? (h[b * 6 + i] > r[i] ? h[b * 6 + i] : r[i])
                     : (h[b * 6 + i] < r[i] ? h[b * 6 + i] : r[i]);
  const T out[6] = {r[0], r[2], DIM == 3 ? r[4] : T(0),
                    r[1], r[3], DIM == 3 ? r[5] : T(0)};
  memcpy(mm, out, sizeof(out));
  return GPZ_OK;
}

// ---- host side of batches
int device_cus() {
  static std::mutex mu;
  static std::map<int, int> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  std::lock_guard<std::mutex> g(mu);
  auto it = cache.find(dev);
  if (it != cache.end()) return it->second;
  int n = 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) n = 0;
  cache[dev] = n;
  return n;
}

// Workgroups of `threads` that fill the device for kernel f.
uint32_t fill_blocks(const void* f, uint32_t threads, size_t lds) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, static_cast<int>(threads), lds) !=
          hipSuccess || per_cu < 1)
    per_cu = 1;
  const int cus = device_cus();
  return static_cast<uint32_t>(per_cu * (cus > 0 ? cus : 1));
}

struct BatchLayout {
  uint64_t items, cls, segs, infos, regions, total;
};
uint64_t align256h(uint64_t v) { return (v + 255) & ~255ull; }

template <typename Item>
BatchLayout batch_layout(uint32_t count, uint64_t total_groups, uint64_t region_bytes) {
  BatchLayout L;
  L.items = 0;
  L.cls = align256h(sizeof(Item) * count);
  L.segs = align256h(L.cls + 4ull * count);
  L.infos = align256h(L.segs + sizeof(Segment) * total_groups);
  L.regions = align256h(L.infos + sizeof(DevInfo) * count);
  L.total = L.regions + region_bytes;
  return L;
}

bool tuning_ok(const gpz_batch_tuning* t) {
  if (!t) return true;
  const uint32_t s = t->seg_groups, w = t->waves_per_wg;
  return (s == 1 || s == 2 || s == 4 || s == 8) && (w == 1 || w == 2 || w == 4 || w == 8);
}

gpz_batch_tuning default_tuning() {
  gpz_batch_tuning t;
  t.seg_groups = 1;
  t.waves_per_wg = 1;
  return t;
}

// Class lists and segment queues shared by both directions.
template <typename GetClass, typename GetGroups>
void build_queues(uint32_t count, uint32_t seg_groups, GetClass cls_of, GetGroups groups_of_item,
                  gpz_batch_plan* plan, std::vector<uint32_t>* cls,
                  std::vector<Segment>* segs) {
  cls->clear();
  segs->clear();
  for (int k = 0; k < kClasses; ++k) {
    plan->cls_first[k] = static_cast<uint32_t>(cls->size());
    plan->cls_seg_first[k] = static_cast<uint32_t>(segs->size());
    for (uint32_t i = 0; i < count; ++i) {
      if (cls_of(i) != k) continue;
      cls->push_back(i);
      const uint64_t G = groups_of_item(i);
      for (uint64_t g = 0; g < G; g += seg_groups) {
        Segment sg;
        sg.item = i;
        sg.g0 = static_cast<uint32_t>(g);
        sg.ng = static_cast<uint32_t>(G - g < seg_groups ? G - g : seg_groups);
        sg.pad = 0;
        segs->push_back(sg);
      }
    }
    plan->cls_items[k] = static_cast<uint32_t>(cls->size()) - plan->cls_first[k];
    plan->cls_segs[k] = static_cast<uint32_t>(segs->size()) - plan->cls_seg_first[k];
  }
}

int upload(void* d_ws, const std::vector<uint8_t>& host, hipStream_t s) {
  if (hipMemcpyAsync(d_ws, host.data(), host.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return GPZ_ERR_HIP;
  return GPZ_OK;
}

// Timer names per class (f32 2D, f32 3D, f64 2D, f64 3D) and stage.
const char* batch_timer(int stage, int k) {
  static const char* names[5][kClasses] = {
      {"batch_seg/f32_2d", "batch_seg/f32_3d", "batch_seg/f64_2d", "batch_seg/f64_3d"},
      {"batch_encode/f32_2d", "batch_encode/f32_3d", "batch_encode/f64_2d", "batch_encode/f64_3d"},
      {"batch_scan/f32_2d", "batch_scan/f32_3d", "batch_scan/f64_2d", "batch_scan/f64_3d"},
      {"batch_compact/f32_2d", "batch_compact/f32_3d", "batch_compact/f64_2d",
       "batch_compact/f64_3d"},
      {"batch_decode/f32_2d", "batch_decode/f32_3d", "batch_decode/f64_2d", "batch_decode/f64_3d"}};
  return names[stage][k];
}

template <typename T, int DIM>
void batch_compress_class(const gpz_batch_plan* plan, int k, uint8_t* ws, hipStream_t s,
                          bool timed) {
  const uint32_t ni = plan->cls_items[k], ns = plan->cls_segs[k];
  if (!ni) return;
  const BatchLayout L = batch_layout<BatchItem>(plan->count, plan->total_groups, 0);
  const BatchItem* items = reinterpret_cast<const BatchItem*>(ws + L.items);
  const uint32_t* cls = reinterpret_cast<const uint32_t*>(ws + L.cls) + plan->cls_first[k];
  const Segment* segs = reinterpret_cast<const Segment*>(ws + L.segs) + plan->cls_seg_first[k];
  const uint32_t W = plan->waves_per_wg;
  {
    LaunchTimer lt(timed ? batch_timer(0, k) : nullptr, s);
    hipLaunchKernelGGL((batch_seg_sample_kernel<T, DIM>), dim3(kSampleBlocks / 4, ni), dim3(256),
                       0, s, items, cls);
    hipLaunchKernelGGL((batch_seg_decide_kernel<T, DIM>), dim3(ni), dim3(256), 0, s, items, cls);
  }
  {
    LaunchTimer lt(timed ? batch_timer(1, k) : nullptr, s);
    const size_t lds = static_cast<size_t>(W) * kLdsBytes;
    hipLaunchKernelGGL((batch_encode_kernel<T, DIM>), dim3(plan->grid[k]), dim3(64 * W), lds, s,
                       items, segs, ns);
  }
  {
    LaunchTimer lt(timed ? batch_timer(2, k) : nullptr, s);
    hipLaunchKernelGGL((batch_scan_kernel<T>), dim3(ni), dim3(256), 0, s, items, cls);
  }
  {
    LaunchTimer lt(timed ? batch_timer(3, k) : nullptr, s);
    const uint32_t need = (ns + 3) / 4, fill = 4096;
    hipLaunchKernelGGL((batch_compact_kernel<T>), dim3(need < fill ? need : fill), dim3(256), 0,
                       s, items, segs, ns);
  }
}

template <typename T, int DIM>
void batch_decompress_class(const gpz_batch_plan* plan, int k, uint8_t* ws, hipStream_t s,
                            bool timed) {
  const uint32_t ns = plan->cls_segs[k];
  if (!plan->cls_items[k]) return;
  const BatchLayout L = batch_layout<DecodeArgs>(plan->count, plan->total_groups, 0);
  const DecodeArgs* items = reinterpret_cast<const DecodeArgs*>(ws + L.items);
  const Segment* segs = reinterpret_cast<const Segment*>(ws + L.segs) + plan->cls_seg_first[k];
  const uint32_t W = plan->waves_per_wg;
  LaunchTimer lt(timed ? batch_timer(4, k) : nullptr, s);
  const size_t lds = static_cast<size_t>(W) * dec_lds_bytes<T>();
  hipLaunchKernelGGL((batch_decode_kernel<T, DIM>), dim3(plan->grid[k]), dim3(64 * W), lds, s,
                     items, segs, ns);
}

// Persistent grid of class k: enough workgroups for every segment, at
// most what fills the device.
template <typename T, int DIM>
uint32_t plan_grid(uint32_t kind, uint32_t W, uint32_t ns) {
  const uint32_t fill = kind == 0
      ? fill_blocks(reinterpret_cast<const void*>(batch_encode_kernel<T, DIM>), 64 * W,
                    static_cast<size_t>(W) * kLdsBytes)
      : fill_blocks(reinterpret_cast<const void*>(batch_decode_kernel<T, DIM>), 64 * W,
                    static_cast<size_t>(W) * dec_lds_bytes<T>());
  const uint32_t need = (ns + W - 1) / W;
  return need < fill ? (need ? need : 1u) : fill;
}
void plan_grids(gpz_batch_plan* plan) {
  const uint32_t W = plan->waves_per_wg, kind = plan->kind;
  plan->grid[0] = plan_grid<float, 2>(kind, W, plan->cls_segs[0]);
  plan->grid[1] = plan_grid<float, 3>(kind, W, plan->cls_segs[1]);
  plan->grid[2] = plan_grid<double, 2>(kind, W, plan->cls_segs[2]);
  plan->grid[3] = plan_grid<double, 3>(kind, W, plan->cls_segs[3]);
}

// Per device: three auxiliary streams (default priority) and events, so
// that the classes of a batch run concurrently; under capture the fork and
// join become parallel graph branches.
struct ForkStreams {
  hipStream_t aux[kClasses - 1] = {};
  hipEvent_t fork = nullptr, join[kClasses - 1] = {};
  bool ok = false;
};
ForkStreams* fork_streams() {
  static std::mutex mu;
  static std::map<int, ForkStreams> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> g(mu);
  ForkStreams& f = cache[dev];
  if (!f.ok) {
    bool ok = hipEventCreateWithFlags(&f.fork, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; i < kClasses - 1 && ok; ++i)
      ok = hipStreamCreateWithFlags(&f.aux[i], hipStreamNonBlocking) == hipSuccess &&
           hipEventCreateWithFlags(&f.join[i], hipEventDisableTiming) == hipSuccess;
    if (!ok) return nullptr;
    f.ok = true;
  }
  return &f;
}

int batch_run_impl(const gpz_batch_plan* plan, void* d_ws, hipStream_t s, bool timed) {
  uint8_t* ws = static_cast<uint8_t*>(d_ws);
  ForkStreams* fs = fork_streams();
  if (!fs) return GPZ_ERR_HIP;
  // classes with work, largest first on the caller's stream
  int order[kClasses], n = 0;
  for (int k = 0; k < kClasses; ++k)
    if (plan->cls_items[k]) order[n++] = k;
  for (int a = 0; a < n; ++a)
    for (int b = a + 1; b < n; ++b)
      if (plan->cls_segs[order[b]] * static_cast<uint64_t>(plan->seg_groups) >
          plan->cls_segs[order[a]] * static_cast<uint64_t>(plan->seg_groups)) {
        const int t = order[a]; order[a] = order[b]; order[b] = t;
      }
  if (n > 1 && (hipEventRecord(fs->fork, s) != hipSuccess)) return GPZ_ERR_HIP;
  for (int i = 0; i < n; ++i) {
    hipStream_t cs = i == 0 ? s : fs->aux[i - 1];
    if (i > 0 && hipStreamWaitEvent(cs, fs->fork, 0) != hipSuccess) return GPZ_ERR_HIP;
    const int k = order[i];
    if (plan->kind == 0) {
      switch (k) {
        case 0: batch_compress_class<float, 2>(plan, 0, ws, cs, timed); break;
        case 1: batch_compress_class<float, 3>(plan, 1, ws, cs, timed); break;
        case 2: batch_compress_class<double, 2>(plan, 2, ws, cs, timed); break;
        default: batch_compress_class<double, 3>(plan, 3, ws, cs, timed); break;
      }
    } else {
      switch (k) {
        case 0: batch_decompress_class<float, 2>(plan, 0, ws, cs, timed); break;
        case 1: batch_decompress_class<float, 3>(plan, 1, ws, cs, timed); break;
        case 2: batch_decompress_class<double, 2>(plan, 2, ws, cs, timed); break;
        default: batch_decompress_class<double, 3>(plan, 3, ws, cs, timed); break;
      }
    }
    if (i > 0 && (hipEventRecord(fs->join[i - 1], cs) != hipSuccess ||
                  hipStreamWaitEvent(s, fs->join[i - 1], 0) != hipSuccess))
      return GPZ_ERR_HIP;
  }
  return hipGetLastError() == hipSuccess ? GPZ_OK : GPZ_ERR_HIP;
}

}  // namespace

extern "C" {

int gpz_version(void) { return 100; }

uint64_t gpz_stream_bound(const gpz_params* p) {
  if (!p || p->n == 0) return 0;
  const uint64_t G = groups_of(p->n);
  const uint64_t P = p->float_bits == 64 ? 32 : 24;
  const uint64_t M = p->float_bits == 64 ? 64 : 36;
  return P + (8 + M) * G + static_cast<uint64_t>(kSlot) * G;
}

uint64_t gpz_compress_workspace_bytes(const gpz_params* p) {
  if (!p || p->n == 0) return 0;
  return ws_layout(p->float_bits == 64 ? 64 : 32, groups_of(p->n)).total;
}

int gpz_compress(const gpz_params* p, const void* d_x, const void* d_y,
                 const void* d_z, void* d_stream, uint64_t cap, void* d_ws,
                 uint64_t ws_bytes, gpz_stream_t stream) {
  if (!params_ok(p) || !d_x || !d_y || (p->dim == 3 && !d_z) || !d_stream)
    return GPZ_ERR_INVALID_ARG;
  if (!d_ws || ws_bytes < gpz_compress_workspace_bytes(p)) return GPZ_ERR_WORKSPACE;
  if (cap < gpz_stream_bound(p)) return GPZ_ERR_CAPACITY;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (p->float_bits == 32)
    return p->dim == 3 ? compress_impl<float, 3>(p, d_x, d_y, d_z, d_stream, d_ws, s)
                       : compress_impl<float, 2>(p, d_x, d_y, d_z, d_stream, d_ws, s);
  return p->dim == 3 ? compress_impl<double, 3>(p, d_x, d_y, d_z, d_stream, d_ws, s)
                     : compress_impl<double, 2>(p, d_x, d_y, d_z, d_stream, d_ws, s);
}

int gpz_compress_result(const void* d_ws, gpz_result* out, gpz_stream_t stream) {
  if (!d_ws || !out) return GPZ_ERR_INVALID_ARG;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (hipMemcpyAsync(out, d_ws, sizeof(gpz_result), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return GPZ_ERR_HIP;
  return GPZ_OK;
}

int gpz_parse_header(const void* h_stream, uint64_t bytes, gpz_header* out) {
  if (!h_stream || !out || bytes < 24) return GPZ_ERR_INVALID_ARG;
  const uint8_t* b = static_cast<const uint8_t*>(h_stream);
  gpz_header h;
  memset(&h, 0, sizeof(h));
  memcpy(&h.dim, b, 4);
  memcpy(&h.n, b + 4, 8);
  memcpy(&h.float_bits, b + 12, 4);
  if (h.float_bits == 32) {
    float e;
    memcpy(&e, b + 16, 4);
    h.eb = e;
    memcpy(&h.seg_size, b + 20, 4);
    h.header_bytes = 24;
  } else if (h.float_bits == 64) {
    if (bytes < 32) return GPZ_ERR_CORRUPT;
    memcpy(&h.eb, b + 16, 8);
    memcpy(&h.seg_size, b + 24, 4);
    h.header_bytes = 32;
  } else {
    return GPZ_ERR_CORRUPT;
  }
  if ((h.dim != 2 && h.dim != 3) || h.n == 0 || h.seg_size == 0 ||
      groups_of(h.n) > 0x7FFFFFFFull)
    return GPZ_ERR_CORRUPT;
  h.groups = groups_of(h.n);
  h.min_stream_bytes = h.header_bytes + (8ull + (h.float_bits == 64 ? 64 : 36)) * h.groups;
  *out = h;
  return GPZ_OK;
}

int gpz_read_header(const void* d_stream, gpz_header* out, gpz_stream_t stream) {
  if (!d_stream || !out) return GPZ_ERR_INVALID_ARG;
  uint8_t h[32];
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (hipMemcpyAsync(h, d_stream, 32, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return GPZ_ERR_HIP;
  return gpz_parse_header(h, 32, out);
}

int gpz_stream_bytes(const gpz_header* h, const void* d_stream,
                     uint64_t* bytes, gpz_stream_t stream) {
  if (!h || !d_stream || !bytes || h->groups == 0) return GPZ_ERR_INVALID_ARG;
  const uint64_t M = h->float_bits == 64 ? 64 : 36;
  const uint64_t G = h->groups;
  const uint8_t* b = static_cast<const uint8_t*>(d_stream);
  uint64_t off = 0;
  uint8_t meta[64];
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (hipMemcpyAsync(&off, b + h->header_bytes + 8 * (G - 1), 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(meta, b + h->header_bytes + 8 * G + M * (G - 1), M, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return GPZ_ERR_HIP;
  uint16_t u[4];
  memcpy(u, meta + (M == 64 ? 56 : 28), 8);
  uint32_t D, C, R;
  if (h->float_bits == 64) {
    if (h->dim == 3) section_sizes<double, 3>(u[3], u[0], u[1], u[2], &D, &C, &R);
    else section_sizes<double, 2>(u[3], u[0], u[1], u[2], &D, &C, &R);
  } else {
    section_sizes<float, 3>(u[3], u[0], u[1], u[2], &D, &C, &R);
  }
  *bytes = h->min_stream_bytes + off + D + C + R;
  return GPZ_OK;
}

int gpz_decompress(const gpz_header* h, const void* d_stream,
                   uint64_t stream_bytes, void* d_x, void* d_y, void* d_z,
                   gpz_stream_t stream) {
  if (!h || !d_stream || !d_x || !d_y || (h->dim == 3 && !d_z))
    return GPZ_ERR_INVALID_ARG;
  if ((h->dim != 2 && h->dim != 3) || h->n == 0 || h->seg_size == 0 ||
      h->groups != groups_of(h->n) || h->groups > 0x7FFFFFFFull ||
      stream_bytes < h->min_stream_bytes)
    return GPZ_ERR_CORRUPT;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (h->float_bits == 32)
    return h->dim == 3 ? decompress_impl<float, 3>(h, d_stream, stream_bytes, d_x, d_y, d_z, s)
                       : decompress_impl<float, 2>(h, d_stream, stream_bytes, d_x, d_y, d_z, s);
  if (h->float_bits == 64)
    return h->dim == 3 ? decompress_impl<double, 3>(h, d_stream, stream_bytes, d_x, d_y, d_z, s)
                       : decompress_impl<double, 2>(h, d_stream, stream_bytes, d_x, d_y, d_z, s);
  return GPZ_ERR_CORRUPT;
}

uint64_t gpz_min_max_workspace_bytes(void) { return 6ull * kMinMaxBlocks * 8; }

int gpz_min_max(uint32_t dim, uint32_t float_bits, uint64_t n,
                const void* d_x, const void* d_y, const void* d_z, void* d_ws,
                void* mm, gpz_stream_t stream) {
  if ((dim != 2 && dim != 3) || n == 0 || !d_x || !d_y || (dim == 3 && !d_z) ||
      !d_ws || !mm)
    return GPZ_ERR_INVALID_ARG;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (float_bits == 32)
    return dim == 3 ? minmax_impl<float, 3>(n, d_x, d_y, d_z, d_ws, mm, s)
                    : minmax_impl<float, 2>(n, d_x, d_y, d_z, d_ws, mm, s);
  if (float_bits == 64)
    return dim == 3 ? minmax_impl<double, 3>(n, d_x, d_y, d_z, d_ws, mm, s)
                    : minmax_impl<double, 2>(n, d_x, d_y, d_z, d_ws, mm, s);
  return GPZ_ERR_INVALID_ARG;
}

int gpz_timing_enable(int on) {
  std::lock_guard<std::mutex> g(timing().mu);
  timing().on = on != 0;
  return GPZ_OK;
}

int gpz_timing_reset(void) {
  std::lock_guard<std::mutex> g(timing().mu);
  for (auto& r : timing().pending) {
    (void)hipEventSynchronize(r.b);
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  timing().pending.clear();
  timing().totals.clear();
  return GPZ_OK;
}

int gpz_timing_read(const char* kernel, double* total_ms, uint64_t* launches) {
  if (!kernel || !total_ms || !launches) return GPZ_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> g(timing().mu);
  for (auto& r : timing().pending) {
    float ms = 0.f;
    if (hipEventSynchronize(r.b) != hipSuccess ||
        hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess)
      return GPZ_ERR_HIP;
    auto& t = timing().totals[r.name];
    t.first += ms;
    t.second += 1;
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  timing().pending.clear();
  auto it = timing().totals.find(kernel);
  *total_ms = it == timing().totals.end() ? 0.0 : it->second.first;
  *launches = it == timing().totals.end() ? 0 : it->second.second;
  return GPZ_OK;
}

double gpz_rel_to_abs(uint32_t float_bits, double e, const void* mm) {
  if (!mm) return 0.0;
  if (float_bits == 32) {
    const float* v = static_cast<const float*>(mm);
    float r = fmaxf(v[3] - v[0], v[4] - v[1]);
    r = fmaxf(r, v[5] - v[2]);
    return static_cast<double>(r * static_cast<float>(e));
  }
  const double* v = static_cast<const double*>(mm);
  double r = fmax(v[3] - v[0], v[4] - v[1]);
  r = fmax(r, v[5] - v[2]);
  return r * e;
}

}  // extern "C"

// ------------------------------------------------------------- batches
extern "C" {

uint64_t gpz_batch_compress_workspace_bytes(const gpz_batch_item* items, uint32_t count) {
  if (!items || count == 0) return 0;
  uint64_t groups = 0, regions = 0;
  for (uint32_t i = 0; i < count; ++i) {
    const gpz_params& p = items[i].p;
    if (p.n == 0) return 0;
    const uint64_t G = groups_of(p.n);
    groups += G;
    regions += align256h(ws_layout(p.float_bits == 64 ? 64 : 32, G).total);
  }
  return batch_layout<BatchItem>(count, groups, regions).total;
}

int gpz_batch_plan_compress(const gpz_batch_item* items, uint32_t count,
                            const gpz_batch_tuning* tune, void* d_ws, uint64_t ws_bytes,
                            gpz_batch_plan* plan, gpz_stream_t stream) {
  if (!items || count == 0 || count > 65535 || !plan || !tuning_ok(tune))
    return GPZ_ERR_INVALID_ARG;
  uint64_t groups = 0;
  for (uint32_t i = 0; i < count; ++i) {
    const gpz_batch_item& it = items[i];
    if (!params_ok(&it.p) || !it.x || !it.y || (it.p.dim == 3 && !it.z) || !it.stream)
      return GPZ_ERR_INVALID_ARG;
    if (it.cap < gpz_stream_bound(&it.p)) return GPZ_ERR_CAPACITY;
    groups += groups_of(it.p.n);
  }
  if (groups > 0xFFFFFFFFull) return GPZ_ERR_INVALID_ARG;
  const uint64_t need = gpz_batch_compress_workspace_bytes(items, count);
  if (!d_ws || ws_bytes < need) return GPZ_ERR_WORKSPACE;
  const gpz_batch_tuning t = tune ? *tune : default_tuning();
  memset(plan, 0, sizeof(*plan));
  plan->count = count;
  plan->kind = 0;
  plan->seg_groups = t.seg_groups;
  plan->waves_per_wg = t.waves_per_wg;
  plan->total_groups = groups;
  plan->ws_bytes = need;

  std::vector<uint32_t> cls;
  std::vector<Segment> segs;
  build_queues(
      count, t.seg_groups,
      [&](uint32_t i) { return class_of(items[i].p.float_bits, items[i].p.dim); },
      [&](uint32_t i) { return groups_of(items[i].p.n); }, plan, &cls, &segs);
  const BatchLayout L = batch_layout<BatchItem>(count, groups, 0);
  std::vector<uint8_t> host(L.regions, 0);
  uint8_t* ws = static_cast<uint8_t*>(d_ws);
  DevInfo* infos = reinterpret_cast<DevInfo*>(ws + L.infos);
  uint64_t r = L.regions;
  for (uint32_t i = 0; i < count; ++i) {
    const gpz_batch_item& it = items[i];
    const gpz_params& p = it.p;
    const bool f64 = p.float_bits == 64;
    const uint64_t n = p.n, G = groups_of(n);
    const WsLayout W = ws_layout(f64 ? 64 : 32, G);
    uint8_t* region = ws + r;
    r += align256h(W.total);
    BatchItem b;
    memset(&b, 0, sizeof(b));
    const double eb = f64 ? p.eb : static_cast<double>(static_cast<float>(p.eb));
    const double rel = f64 ? p.rel_eb : static_cast<double>(static_cast<float>(p.rel_eb));
    // FP32 always samples (its block-count guard can double any seg); FP64
    // only to choose one (as compress_impl).
    const bool sample = !f64 || (p.seg_size == 0 && p.mode == GPZ_MODE_REL);
    uint8_t* stream8 = static_cast<uint8_t*>(it.stream);
    const uint32_t P = f64 ? 32 : 24;
    b.info = infos + i;
    b.win = region + W.win;
    {
      uint64_t st = (n - kSampleLen) / kSampleBlocks;  // u64 wrap for n < 256
      b.stride = st < 1 ? 1 : st;
    }
    b.sa.seg_in = p.seg_size;
    b.sa.is_rel = p.mode == GPZ_MODE_REL;
    b.sa.rel_small = rel <= 1e-4;
    b.sa.sampled = sample;
    b.sa.eb = eb;
    b.e.x = it.x; b.e.y = it.y; b.e.z = it.z;
    b.e.n = n; b.e.G = G; b.e.eb = eb;
    b.e.last_n = static_cast<uint32_t>(n - (G - 1) * kGroup);
    b.e.info = b.info;
    b.e.meta = stream8 + P + 8 * G;
    b.e.loc = reinterpret_cast<uint64_t*>(region + W.loc);
    b.e.slots = region + W.slots;
    b.sc.loc = b.e.loc;
    b.sc.off = reinterpret_cast<uint64_t*>(region + W.off);
    b.sc.bsum = reinterpret_cast<uint64_t*>(region + W.bsum);
    b.sc.stream = stream8;
    b.sc.info = b.info;
    b.sc.n = n; b.sc.G = G; b.sc.dim = p.dim; b.sc.eb = eb;
    b.ca.loc = b.e.loc;
    b.ca.off = b.sc.off;
    b.ca.slots = b.e.slots;
    b.ca.stream = stream8;
    b.ca.info = b.info;
    b.ca.G = G;
    memcpy(host.data() + L.items + sizeof(BatchItem) * i, &b, sizeof(b));
  }
  memcpy(host.data() + L.cls, cls.data(), 4 * cls.size());
  memcpy(host.data() + L.segs, segs.data(), sizeof(Segment) * segs.size());
  plan_grids(plan);
  return upload(d_ws, host, reinterpret_cast<hipStream_t>(stream));
}

uint64_t gpz_batch_decompress_workspace_bytes(const gpz_batch_ditem* items, uint32_t count) {
  if (!items || count == 0) return 0;
  uint64_t groups = 0;
  for (uint32_t i = 0; i < count; ++i) groups += items[i].h.groups;
  return batch_layout<DecodeArgs>(count, groups, 0).total;
}

int gpz_batch_plan_decompress(const gpz_batch_ditem* items, uint32_t count,
                              const gpz_batch_tuning* tune, void* d_ws, uint64_t ws_bytes,
                              gpz_batch_plan* plan, gpz_stream_t stream) {
  if (!items || count == 0 || count > 65535 || !plan || !tuning_ok(tune))
    return GPZ_ERR_INVALID_ARG;
  uint64_t groups = 0;
  for (uint32_t i = 0; i < count; ++i) {
    const gpz_batch_ditem& it = items[i];
    const gpz_header* h = &it.h;
    if (!it.stream || !it.x || !it.y || (h->dim == 3 && !it.z)) return GPZ_ERR_INVALID_ARG;
    if ((h->dim != 2 && h->dim != 3) || (h->float_bits != 32 && h->float_bits != 64) ||
        h->n == 0 || h->seg_size == 0 || h->groups != groups_of(h->n) ||
        h->groups > 0x7FFFFFFFull || it.bytes < h->min_stream_bytes)
      return GPZ_ERR_CORRUPT;
    groups += h->groups;
  }
  if (groups > 0xFFFFFFFFull) return GPZ_ERR_INVALID_ARG;
  const uint64_t need = gpz_batch_decompress_workspace_bytes(items, count);
  if (!d_ws || ws_bytes < need) return GPZ_ERR_WORKSPACE;
  const gpz_batch_tuning t = tune ? *tune : default_tuning();
  memset(plan, 0, sizeof(*plan));
  plan->count = count;
  plan->kind = 1;
  plan->seg_groups = t.seg_groups;
  plan->waves_per_wg = t.waves_per_wg;
  plan->total_groups = groups;
  plan->ws_bytes = need;
  std::vector<uint32_t> cls;
  std::vector<Segment> segs;
  build_queues(
      count, t.seg_groups,
      [&](uint32_t i) { return class_of(items[i].h.float_bits, items[i].h.dim); },
      [&](uint32_t i) { return items[i].h.groups; }, plan, &cls, &segs);
  const BatchLayout L = batch_layout<DecodeArgs>(count, groups, 0);
  std::vector<uint8_t> host(L.regions, 0);
  for (uint32_t i = 0; i < count; ++i) {
    const gpz_batch_ditem& it = items[i];
    DecodeArgs d;
    memset(&d, 0, sizeof(d));
    d.stream = static_cast<const uint8_t*>(it.stream);
    d.stream_bytes = it.bytes;
    d.n = it.h.n;
    d.G = it.h.groups;
    d.last_n = static_cast<uint32_t>(it.h.n - (it.h.groups - 1) * kGroup);
    d.eb = it.h.eb;
    d.seg = it.h.seg_size;
    d.x = it.x; d.y = it.y; d.z = it.z;
    memcpy(host.data() + L.items + sizeof(DecodeArgs) * i, &d, sizeof(d));
  }
  memcpy(host.data() + L.cls, cls.data(), 4 * cls.size());
  memcpy(host.data() + L.segs, segs.data(), sizeof(Segment) * segs.size());
  plan_grids(plan);
  return upload(d_ws, host, reinterpret_cast<hipStream_t>(stream));
}

int gpz_batch_run(const gpz_batch_plan* plan, void* d_ws, gpz_stream_t stream) {
  if (!plan || !d_ws || plan->count == 0) return GPZ_ERR_INVALID_ARG;
  return batch_run_impl(plan, d_ws, reinterpret_cast<hipStream_t>(stream), true);
}

int gpz_batch_compress_results(const gpz_batch_plan* plan, const void* d_ws, gpz_result* out,
                               gpz_stream_t stream) {
  if (!plan || plan->kind != 0 || !d_ws || !out) return GPZ_ERR_INVALID_ARG;
  const BatchLayout L = batch_layout<BatchItem>(plan->count, plan->total_groups, 0);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (hipMemcpyAsync(out, static_cast<const uint8_t*>(d_ws) + L.infos,
                     sizeof(gpz_result) * plan->count, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return GPZ_ERR_HIP;
  return GPZ_OK;
}

int gpz_batch_graph_create(const gpz_batch_plan* plan, void* d_ws, void** graph_exec) {
  if (!plan || !d_ws || !graph_exec || plan->count == 0) return GPZ_ERR_INVALID_ARG;
  *graph_exec = nullptr;
  hipStream_t cs;
  if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return GPZ_ERR_HIP;
  int rc = GPZ_OK;
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    rc = GPZ_ERR_HIP;
  } else {
    const int r1 = batch_run_impl(plan, d_ws, cs, false);
    if (hipStreamEndCapture(cs, &g) != hipSuccess || r1 != GPZ_OK) rc = GPZ_ERR_HIP;
  }
  hipGraphExec_t e = nullptr;
  if (rc == GPZ_OK && hipGraphInstantiate(&e, g, nullptr, nullptr, 0) != hipSuccess)
    rc = GPZ_ERR_HIP;
  if (g) (void)hipGraphDestroy(g);
  (void)hipStreamDestroy(cs);
  if (rc == GPZ_OK) *graph_exec = e;
  return rc;
}

int gpz_batch_graph_launch(void* graph_exec, gpz_stream_t stream) {
  if (!graph_exec) return GPZ_ERR_INVALID_ARG;
  return hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec),
                        reinterpret_cast<hipStream_t>(stream)) == hipSuccess
             ? GPZ_OK
             : GPZ_ERR_HIP;
}

int gpz_batch_graph_destroy(void* graph_exec) {
  if (!graph_exec) return GPZ_ERR_INVALID_ARG;
  return hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec)) == hipSuccess
             ? GPZ_OK
             : GPZ_ERR_HIP;
}

// Sweeps groups per segment {1, 2, 4, 8} x waves per workgroup {1, 2, 4,
// 8}: plans, runs once to warm up, then times `reps` runs with hipEvents.
// table[4*i + j] receives the ms per run of seg_groups 2^i, waves 2^j.
static int autotune_common(bool compress, const void* items, uint32_t count, void* d_ws,
                           uint64_t ws_bytes, gpz_stream_t stream, uint32_t reps,
                           gpz_batch_tuning* best, float* table) {
  if (!best || reps == 0) return GPZ_ERR_INVALID_ARG;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return GPZ_ERR_HIP;
  float best_ms = -1.0f;
  int rc = GPZ_OK;
  for (uint32_t i = 0; i < 4 && rc == GPZ_OK; ++i) {
    for (uint32_t j = 0; j < 4 && rc == GPZ_OK; ++j) {
      gpz_batch_tuning t;
      t.seg_groups = 1u << i;
      t.waves_per_wg = 1u << j;
      gpz_batch_plan plan;
      rc = compress ? gpz_batch_plan_compress(static_cast<const gpz_batch_item*>(items), count,
                                              &t, d_ws, ws_bytes, &plan, stream)
                    : gpz_batch_plan_decompress(static_cast<const gpz_batch_ditem*>(items),
                                                count, &t, d_ws, ws_bytes, &plan, stream);
      if (rc != GPZ_OK) break;
      rc = batch_run_impl(&plan, d_ws, s, false);
      if (rc != GPZ_OK) break;
      (void)hipEventRecord(e0, s);
      for (uint32_t k = 0; k < reps && rc == GPZ_OK; ++k) rc = batch_run_impl(&plan, d_ws, s, false);
      (void)hipEventRecord(e1, s);
      float ms = 0.0f;
      if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
        rc = GPZ_ERR_HIP;
      ms /= static_cast<float>(reps);
      if (table) table[4 * i + j] = ms;
      if (rc == GPZ_OK && (best_ms < 0.0f || ms < best_ms)) {
        best_ms = ms;
        *best = t;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

int gpz_batch_autotune_compress(const gpz_batch_item* items, uint32_t count, void* d_ws,
                                uint64_t ws_bytes, gpz_stream_t stream, uint32_t reps,
                                gpz_batch_tuning* best, float* table) {
  return autotune_common(true, items, count, d_ws, ws_bytes, stream, reps, best, table);
}

int gpz_batch_autotune_decompress(const gpz_batch_ditem* items, uint32_t count, void* d_ws,
                                  uint64_t ws_bytes, gpz_stream_t stream, uint32_t reps,
                                  gpz_batch_tuning* best, float* table) {
  return autotune_common(false, items, count, d_ws, ws_bytes, stream, reps, best, table);
}

}  // extern "C"
