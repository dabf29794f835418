// gpz_oracle.cpp -- CPU restatement of GPZ's four-stage particle compressor.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle and the CPU
// baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it, and only as the checker.  The product path (libgpz_amd.so)
// never links, loads or falls back to it.
//
// It restates, in scalar C++ (OpenMP over 1024-point groups), the algorithm of
//   src/cudaWarpCompress.cu        (FP32: encode :58-435/:437-796, decode
//                                    :798-1008/:1010-1214, seg size
//                                    :1216-1557, stream assembly :1573-1741,
//                                    decode host :1743-1856)
//   src/cudaWarpCompressDouble.cu  (FP64 counterparts, same line ranges +-3)
//   include/cudaUtils.cuh:69-75    (effective bits)
//   include/utils.cuh:44-65        (findMinimumValues)
//   src/main.cpp:35-55,246-276     (rel -> abs error bound)
// Each function below cites the reference lines it follows.
//
// Where the reference relies on undefined or unspecified behaviour the
// oracle fixes one definition (and the HIP kernels follow the same one):
//   * float -> unsigned casts saturate (NaN -> 0), as CUDA cvt.rzi does;
//   * shifts by >= the word width give 0, as PTX shr does;
//   * the decoder's FMA contraction of  min + float(2q+1)*eb  is made
//     explicit with fmaf / fma;
//   * pad bytes are zero;
//   * sample windows that the seg-size sampler skips (start+256 > N) read
//     0.0 for all six extrema (the value a freshly zeroed scratch buffer
//     holds in the reference CLI, src/main.cpp:318);
//   * intra-warp races in the reference (src/cudaWarpCompress.cu:269-311,
//     :316-344) take their intended, barrier-ordered outcome;
//   * the decoder accepts U == 1024 (reference prefix loop writes one
//     element past its array, src/cudaWarpCompress.cu:932);
//   * FP64-2D streams are decoded with the encoder's count-section padding
//     (4 B, src/cudaWarpCompressDouble.cu:691) and the encoder's u64 block
//     count (:525-526), so that every stream the encoder writes round-trips.
//
// Parity status of each function is recorded in the header of
// tests/test_oracle.py (golden vectors, closed forms, brute force).

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr uint32_t kGroup = 1024;  // include/cudaCompress.cuh:9-11 (32 x 32)

// ---- CUDA conversion semantics (cvt.rzi.{u32,u64}.{f32,f64}: truncate,
//      saturate, NaN -> 0) ------------------------------------------------
inline uint32_t sat_u32(float f) {
  if (!(f >= 1.0f)) return 0u;            // NaN, negatives, (-1,1)
  if (f >= 4294967296.0f) return 0xFFFFFFFFu;
  return static_cast<uint32_t>(f);
}
inline uint32_t sat_u32(double f) {
  if (!(f >= 1.0)) return 0u;
  if (f >= 4294967296.0) return 0xFFFFFFFFu;
  return static_cast<uint32_t>(f);
}
inline uint64_t sat_u64(double f) {
  if (!(f >= 1.0)) return 0u;
  if (f >= 18446744073709551616.0) return ~0ull;
  return static_cast<uint64_t>(f);
}
inline uint64_t sat_u64(float f) { return sat_u64(static_cast<double>(f)); }

// include/cudaUtils.cuh:69-75  (32 - clz, 64 - clzll)
inline int eff_bits(uint32_t x) { return x ? 32 - __builtin_clz(x) : 0; }
inline int eff_bits(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }

// PTX shr clamps the shift amount at the register width.
inline uint32_t shr(uint32_t v, uint32_t s) { return s >= 32 ? 0u : (v >> s); }
inline uint64_t shr(uint64_t v, uint32_t s) { return s >= 64 ? 0ull : (v >> s); }

// Integer division by zero is undefined in both languages; PTX div.u32 by 0
// yields all-ones and rem by 0 yields the dividend.  Only groups whose block
// count wrapped to 0 (ceil() saturated) can divide by zero.
template <typename Q> inline Q udiv(Q a, Q b) { return b ? a / b : ~Q(0); }
template <typename Q> inline Q urem(Q a, Q b) { return b ? a % b : a; }

inline uint32_t ctz_seg(uint32_t seg) {  // __ffs(seg) - 1, seg >= 1
  return static_cast<uint32_t>(__builtin_ctz(seg));
}

template <typename T> struct Traits;
template <> struct Traits<float> {
  using Q = uint32_t;
  static constexpr uint32_t kHeader = 24;  // src/cudaWarpCompress.cu:1668
  static constexpr uint32_t kMeta = 36;    // src/cudaWarpCompress.cu:20
  static constexpr uint32_t kBits = 32;
  static float lo() { return FLT_MAX; }
  static float fmin_(float a, float b) { return fminf(a, b); }
  static float fmax_(float a, float b) { return fmaxf(a, b); }
  static float fma_(float a, float b, float c) { return fmaf(a, b, c); }
  static float ceil_(float a) { return ceilf(a); }
};
template <> struct Traits<double> {
  using Q = uint64_t;
  static constexpr uint32_t kHeader = 32;  // src/cudaWarpCompressDouble.cu:1668
  static constexpr uint32_t kMeta = 64;    // src/cudaWarpCompressDouble.cu:20
  static constexpr uint32_t kBits = 64;
  static double lo() { return DBL_MAX; }
  static double fmin_(double a, double b) { return fmin(a, b); }
  static double fmax_(double a, double b) { return fmax(a, b); }
  static double fma_(double a, double b, double c) { return fma(a, b, c); }
  static double ceil_(double a) { return ceil(a); }
};

// Section sizes, src/cudaWarpCompress.cu:324-329 (FP32, 4-byte pad) and
// src/cudaWarpCompressDouble.cu:325-329 (FP64 3D, 8-byte pad) / :686-692
// (FP64 2D: count section padded to 4 bytes).
template <typename T>
inline void section_sizes(int dim, uint32_t U, uint32_t db, uint32_t cb,
                          uint32_t rb, uint32_t* D, uint32_t* C, uint32_t* R) {
  const uint32_t B = ((U + 7u) & ~7u) >> 3;
  if (sizeof(T) == 4) {
    *D = (B * db + 3u) & ~3u;
    *C = (B * cb + 3u) & ~3u;
    *R = ((kGroup * rb >> 3) + 3u) & ~3u;
  } else {
    *D = (B * db + 7u) & ~7u;
    *C = dim == 3 ? ((B * cb + 7u) & ~7u) : ((B * cb + 3u) & ~3u);
    *R = ((kGroup * rb >> 3) + 7u) & ~7u;
  }
}

template <typename T>
struct GroupOut {
  T mn[3], mx[3];
  typename Traits<T>::Q outlier;
  uint16_t db, cb, rb, U;
  std::vector<uint8_t> payload;
};

struct StageDump {   // optional per-group stage outputs for tests
  uint64_t* q;       // [3][1024] quantised codes (0 for invalid slots)
  uint64_t* bid;     // [1024] block id per slot (all-ones if invalid)
  uint64_t* rel;     // [1024] relative position per slot
  uint16_t* perm;    // [1024] sorted rank -> slot index j
  uint32_t* info;    // nbx, nby, shift, bidbits, rb, U, db, cb, nvalid
};

// Encode one 1024-point group.
//   quantise  : src/cudaWarpCompress.cu:96-184   (FP64 :96-184)
//   block id  : src/cudaWarpCompress.cu:141-166,186-208
//   sort      : src/cudaWarpCompress.cu:210-220  (CUB BlockRadixSort: stable,
//               blocked input, thread t item i = slot 32i+t)
//   runs      : src/cudaWarpCompress.cu:222-322
//   meta/sizes: src/cudaWarpCompress.cu:313-348
//   bit planes: src/cudaWarpCompress.cu:352-434
//   2D        : src/cudaWarpCompress.cu:437-796 (z dropped, min/max z = 0)
template <typename T>
void encode_group(int dim, const T* x, const T* y, const T* z, uint64_t n,
                  uint64_t c, uint64_t G, T eb, uint32_t seg, GroupOut<T>& out,
                  StageDump* dump) {
  using Q = typename Traits<T>::Q;
  using Tr = Traits<T>;
  const uint64_t base = c * kGroup;
  const uint32_t nvalid =
      static_cast<uint32_t>(std::min<uint64_t>(kGroup, n - base));
  const bool last = (c == G - 1);

  T mn[3] = {Tr::lo(), Tr::lo(), Tr::lo()};
  T mx[3] = {-Tr::lo(), -Tr::lo(), -Tr::lo()};
  const T* P[3] = {x, y, z};
  for (uint32_t j = 0; j < nvalid; ++j)
    for (int a = 0; a < dim; ++a) {
      mn[a] = Tr::fmin_(mn[a], P[a][base + j]);
      mx[a] = Tr::fmax_(mx[a], P[a][base + j]);
    }
  if (dim == 2) { mn[2] = 0; mx[2] = 0; }

  const T inv = T(1) / (T(2) * eb);
  const T seg2eb = static_cast<T>(2u * seg) * eb;
  Q nbx, nby;
  if (sizeof(T) == 4) {
    nbx = static_cast<Q>(sat_u32(Tr::ceil_((mx[0] - mn[0]) / seg2eb)) + 1u);
    nby = static_cast<Q>(sat_u32(Tr::ceil_((mx[1] - mn[1]) / seg2eb)) + 1u);
  } else if (dim == 3) {  // cudaWarpCompressDouble.cu:151-154 (u32 cast)
    nbx = static_cast<Q>(static_cast<uint32_t>(
        sat_u32(Tr::ceil_((mx[0] - mn[0]) / seg2eb)) + 1u));
    nby = static_cast<Q>(static_cast<uint32_t>(
        sat_u32(Tr::ceil_((mx[1] - mn[1]) / seg2eb)) + 1u));
  } else {                // cudaWarpCompressDouble.cu:525-526 (u64 cast)
    nbx = static_cast<Q>(sat_u64(Tr::ceil_((mx[0] - mn[0]) / seg2eb)) + 1u);
    nby = 0;
  }
  const uint32_t sh = ctz_seg(seg);
  const Q m = static_cast<Q>(seg - 1u);
  const Q sq = static_cast<Q>(seg);

  Q bid[kGroup], rel[kGroup];
  int bidbits = 0, rb = 0;
  const T nmi[3] = {-mn[0] * inv, -mn[1] * inv, -mn[2] * inv};
  for (uint32_t j = 0; j < kGroup; ++j) {
    if (j < nvalid) {
      Q q[3] = {0, 0, 0};
      for (int a = 0; a < dim; ++a) {
        const T f = Tr::fma_(P[a][base + j], inv, nmi[a]);
        q[a] = sizeof(T) == 4 ? static_cast<Q>(sat_u32(f))
                              : static_cast<Q>(sat_u64(f));
      }
      Q b = shr(q[0], sh) + shr(q[1], sh) * nbx;
      Q r = (q[0] & m) + (q[1] & m) * sq;
      if (dim == 3) {
        b += shr(q[2], sh) * nbx * nby;
        r += (q[2] & m) * sq * sq;
      }
      bid[j] = b & Q(0x7FFFFFFF);
      rel[j] = r & Q(0x7FFFFFFF);
      bidbits = std::max(bidbits, eff_bits(bid[j]));
      rb = std::max(rb, eff_bits(rel[j]));
      if (dump)
        for (int a = 0; a < 3; ++a) dump->q[a * kGroup + j] = q[a];
    } else {
      bid[j] = ~Q(0);
      rel[j] = ~Q(0);
      bidbits = std::max(bidbits, 1);
      rb = std::max(rb, 1);
      if (dump)
        for (int a = 0; a < 3; ++a) dump->q[a * kGroup + j] = 0;
    }
    if (dump) { dump->bid[j] = bid[j]; dump->rel[j] = rel[j]; }
  }

  // Stage 3: stable sort by the low `sortbits` bits of the key, starting
  // from CUB's blocked arrangement: logical position L = 32*(j%32) + j/32.
  const int sortbits = bidbits + (last ? 1 : 0);
  const Q kmask = sortbits >= static_cast<int>(Traits<T>::kBits)
                      ? ~Q(0)
                      : ((Q(1) << sortbits) - 1);
  uint16_t order[kGroup];
  for (uint32_t L = 0; L < kGroup; ++L) order[L] = (L % 32) * 32 + L / 32;
  std::stable_sort(order, order + kGroup, [&](uint16_t a, uint16_t b) {
    return (bid[a] & kmask) < (bid[b] & kmask);
  });
  Q sbid[kGroup], srel[kGroup];
  for (uint32_t r = 0; r < kGroup; ++r) {
    sbid[r] = bid[order[r]];
    srel[r] = rel[order[r]];
  }

  // Stage 4a: runs of equal block id among the nvalid valid ranks.
  uint32_t U = 0;
  uint16_t start[kGroup];
  for (uint32_t r = 0; r < nvalid; ++r)
    if (r == 0 || sbid[r] != sbid[r - 1]) start[U++] = static_cast<uint16_t>(r);
  Q delta[kGroup];
  uint32_t cnt[kGroup];
  int db = 0, cb = 0;
  for (uint32_t k = 0; k < U; ++k) {
    const uint32_t e = (k + 1 < U) ? start[k + 1] : nvalid;
    cnt[k] = e - start[k];
    delta[k] = k == 0 ? Q(0) : Q(sbid[start[k]] - sbid[start[k] - 1]);
    db = std::max(db, eff_bits(delta[k]));
    cb = std::max(cb, eff_bits(cnt[k]));
  }

  uint32_t D, C, R;
  section_sizes<T>(dim, U, db, cb, rb, &D, &C, &R);
  const uint32_t B = (U + 7u) >> 3;
  out.payload.assign(D + C + R, 0);
  uint8_t* pl = out.payload.data();
  // delta planes, MSB-first bit order, plane 0 = highest bit
  for (int p = 0; p < db; ++p) {
    const int b = db - 1 - p;
    for (uint32_t k = 0; k < U; ++k)
      if ((delta[k] >> b) & 1) pl[p * B + k / 8] |= uint8_t(0x80u >> (k % 8));
  }
  uint8_t* pc = pl + D;
  for (int p = 0; p < cb; ++p) {
    const int b = cb - 1 - p;
    for (uint32_t k = 0; k < U; ++k)
      if ((cnt[k] >> b) & 1) pc[p * B + k / 8] |= uint8_t(0x80u >> (k % 8));
  }
  // rel planes: LSB-first over sorted rank, 128 bytes per plane
  uint8_t* pr = pl + D + C;
  for (int p = 0; p < rb; ++p) {
    const int b = rb - 1 - p;
    for (uint32_t r = 0; r < kGroup; ++r)
      if ((srel[r] >> b) & 1) pr[p * 128 + r / 8] |= uint8_t(1u << (r % 8));
  }

  for (int a = 0; a < 3; ++a) { out.mn[a] = mn[a]; out.mx[a] = mx[a]; }
  out.outlier = sbid[0];
  out.db = static_cast<uint16_t>(db);
  out.cb = static_cast<uint16_t>(cb);
  out.rb = static_cast<uint16_t>(rb);
  out.U = static_cast<uint16_t>(U);

  if (dump) {
    for (uint32_t r = 0; r < kGroup; ++r) dump->perm[r] = order[r];
    const uint32_t info[9] = {static_cast<uint32_t>(nbx),
                              static_cast<uint32_t>(nby), sh,
                              static_cast<uint32_t>(bidbits),
                              static_cast<uint32_t>(rb), U,
                              static_cast<uint32_t>(db),
                              static_cast<uint32_t>(cb), nvalid};
    memcpy(dump->info, info, sizeof(info));
  }
}

template <typename T>
void write_meta(uint8_t* dst, const GroupOut<T>& g) {
  // src/cudaWarpCompress.cu:332-344 ; src/cudaWarpCompressDouble.cu:333-344
  T mm[6] = {g.mn[0], g.mn[1], g.mn[2], g.mx[0], g.mx[1], g.mx[2]};
  memcpy(dst, mm, sizeof(mm));
  memcpy(dst + 6 * sizeof(T), &g.outlier, sizeof(g.outlier));
  const uint16_t u[4] = {g.db, g.cb, g.rb, g.U};
  memcpy(dst + 6 * sizeof(T) + sizeof(g.outlier), u, sizeof(u));
}

// Seg-size selection: src/cudaWarpCompress.cu:1216-1557 (FP32) and
// src/cudaWarpCompressDouble.cu:1426-1555 (FP64).  Returns 0 or -1 when the
// doubled segment size exceeds 1024 (reference exits, :1547-1551).
template <typename T>
int seg_size(int dim, uint64_t n, const T* x, const T* y, const T* z,
             uint32_t seg_in, int is_rel, T rel_eb, T eb, uint32_t* seg_out,
             int* adjusted_out) {
  using Tr = Traits<T>;
  const int32_t sample_length = 256, num_blocks = 1024;
  const uint64_t stride =
      std::max<uint64_t>((n - sample_length) / num_blocks, 1ull);
  T g[6] = {Tr::lo(), -Tr::lo(), Tr::lo(), -Tr::lo(), Tr::lo(), -Tr::lo()};
  const T* P[3] = {x, y, z};
  for (int b = 0; b < num_blocks; ++b) {
    const uint64_t s = static_cast<uint64_t>(b) * stride;
    T w[6] = {0, 0, 0, 0, 0, 0};   // skipped window: zeroed scratch
    if (!(s + sample_length > n)) {
      for (int a = 0; a < dim; ++a) {
        T lo = P[a][s], hi = P[a][s];
        for (int t = 1; t < sample_length; ++t) {
          lo = Tr::fmin_(lo, P[a][s + t]);
          hi = Tr::fmax_(hi, P[a][s + t]);
        }
        w[2 * a] = lo;
        w[2 * a + 1] = hi;
      }
    }
    for (int a = 0; a < dim; ++a) {
      g[2 * a] = Tr::fmin_(g[2 * a], w[2 * a]);
      g[2 * a + 1] = Tr::fmax_(g[2 * a + 1], w[2 * a + 1]);
    }
  }
  if (dim == 2) { g[4] = 0; g[5] = 0; }
  const T dx = g[1] - g[0], dy = g[3] - g[2], dz = g[5] - g[4];
  const T mean = (dx + dy + dz) / T(3.0f);
  T var;
  if (sizeof(T) == 4 && dim == 2)   // FP32 2D divides by 2 (:1500-1503)
    var = ((dx - mean) * (dx - mean) + (dy - mean) * (dy - mean) +
           (dz - mean) * (dz - mean)) / T(2.0f);
  else
    var = ((dx - mean) * (dx - mean) + (dy - mean) * (dy - mean) +
           (dz - mean) * (dz - mean)) / T(3.0f);
  const T ratio = std::sqrt(var) / mean;
  T max_range = std::fmax(dx, dy);
  max_range = std::fmax(max_range, dz);
  uint32_t seg = seg_in;
  if (seg == 0 && is_rel) {
    const bool small = static_cast<double>(rel_eb) <= 1e-4;
    if (ratio < T(0.1f)) seg = small ? 32u : 2u;
    else seg = small ? 4u : 1u;
  }
  int adjusted = 0;
  if (sizeof(T) == 4) {  // FP64 compares against UINT64_MAX: never adjusts
    auto blocks = [&](uint32_t s) {
      const uint64_t mb =
          sat_u64(Tr::ceil_(max_range / (static_cast<T>(2u * s) * eb))) + 1;
      return dim == 3 ? mb * mb * mb : mb * mb;
    };
    if (blocks(seg) > 0xFFFFFFFFull) {
      uint32_t t = seg;
      for (int i = 0; i < 11; ++i) {
        t *= 2;
        if (blocks(t) <= 0xFFFFFFFFull) break;
      }
      if (t > 1024) return -1;
      seg = t;
      adjusted = 1;
    }
  }
  *seg_out = seg;
  *adjusted_out = adjusted;
  return 0;
}

template <typename T>
uint64_t stream_bytes_bound(uint64_t n) {
  const uint64_t G = (n + kGroup - 1) / kGroup;
  return Traits<T>::kHeader + G * (8 + Traits<T>::kMeta) + G * 9344ull;
}

// Stream assembly: src/cudaWarpCompress.cu:1656-1702 (header, exclusive
// scan of payload sizes, meta, compaction); FP64 :1655-1701.
template <typename T>
int compress(int dim, uint64_t n, const T* x, const T* y, const T* z, T eb,
             uint32_t seg, uint8_t* out, uint64_t cap, uint64_t* out_size) {
  using Tr = Traits<T>;
  if (n == 0 || seg == 0 || (dim != 2 && dim != 3)) return -2;
  const uint64_t G = (n + kGroup - 1) / kGroup;
  std::vector<GroupOut<T>> groups(G);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t c = 0; c < static_cast<int64_t>(G); ++c)
    encode_group<T>(dim, x, y, z, n, c, G, eb, seg, groups[c], nullptr);
  const uint64_t hdr = Tr::kHeader, offs = hdr, meta = hdr + 8 * G,
                 pay = meta + Tr::kMeta * G;
  std::vector<uint64_t> off(G);
  uint64_t acc = 0;
  for (uint64_t c = 0; c < G; ++c) {
    off[c] = acc;
    acc += groups[c].payload.size();
  }
  const uint64_t total = pay + acc;
  *out_size = total;
  if (total > cap) return -3;
  memset(out, 0, hdr);
  const uint32_t d32 = static_cast<uint32_t>(dim);
  const uint32_t bits = sizeof(T) * 8;
  memcpy(out + 0, &d32, 4);
  memcpy(out + 4, &n, 8);
  memcpy(out + 12, &bits, 4);
  memcpy(out + 16, &eb, sizeof(T));
  memcpy(out + 16 + sizeof(T), &seg, 4);
  if (sizeof(T) == 8) {
    const uint32_t ps = 32;
    memcpy(out + 28, &ps, 4);
  }
  memcpy(out + offs, off.data(), 8 * G);
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < static_cast<int64_t>(G); ++c) {
    write_meta<T>(out + meta + Tr::kMeta * c, groups[c]);
    memcpy(out + pay + off[c], groups[c].payload.data(),
           groups[c].payload.size());
  }
  return 0;
}

// Decode one group: src/cudaWarpCompress.cu:798-1008 (3D), :1010-1214 (2D);
// FP64 src/cudaWarpCompressDouble.cu:801-1217.
template <typename T>
int decode_group(int dim, const uint8_t* s, uint64_t n, uint64_t G, uint64_t c,
                 T eb, uint32_t seg, T* x, T* y, T* z, uint64_t payload_bytes) {
  using Q = typename Traits<T>::Q;
  using Tr = Traits<T>;
  const uint64_t meta_base = Tr::kHeader + 8 * G;
  const uint64_t pay_base = meta_base + Tr::kMeta * G;
  const uint8_t* meta = s + meta_base + Tr::kMeta * c;
  T mm[6];
  memcpy(mm, meta, sizeof(mm));
  Q outlier;
  memcpy(&outlier, meta + 6 * sizeof(T), sizeof(Q));
  uint16_t u[4];
  memcpy(u, meta + 6 * sizeof(T) + sizeof(Q), 8);
  const uint32_t db = u[0], cb = u[1], rb = u[2], U = u[3];
  uint64_t off;
  memcpy(&off, s + Tr::kHeader + 8 * c, 8);
  if (U == 0 || U > kGroup || db > 8 * sizeof(Q) || cb > 16 || rb > 32)
    return -4;
  uint32_t D, C, R;
  section_sizes<T>(dim, U, db, cb, rb, &D, &C, &R);
  if (off + D + C + R > payload_bytes) return -4;
  const uint8_t* pd = s + pay_base + off;
  const uint8_t* pc = pd + D;
  const uint8_t* pr = pd + D + C;
  const uint32_t B = (U + 7) >> 3;

  const T seg2eb = static_cast<T>(2u * seg) * eb;
  Q nbx, nby;
  if (sizeof(T) == 4) {
    nbx = static_cast<Q>(sat_u32(Tr::ceil_((mm[3] - mm[0]) / seg2eb)) + 1u);
    nby = static_cast<Q>(sat_u32(Tr::ceil_((mm[4] - mm[1]) / seg2eb)) + 1u);
  } else if (dim == 3) {
    nbx = static_cast<Q>(static_cast<uint32_t>(
        sat_u32(Tr::ceil_((mm[3] - mm[0]) / seg2eb)) + 1u));
    nby = static_cast<Q>(static_cast<uint32_t>(
        sat_u32(Tr::ceil_((mm[4] - mm[1]) / seg2eb)) + 1u));
  } else {  // follow the FP64-2D encoder's u64 count
    nbx = static_cast<Q>(sat_u64(Tr::ceil_((mm[3] - mm[0]) / seg2eb)) + 1u);
    nby = static_cast<Q>(static_cast<uint32_t>(
        sat_u32(Tr::ceil_((mm[4] - mm[1]) / seg2eb)) + 1u));
  }

  Q bidrun[kGroup];
  uint32_t cend[kGroup];
  Q acc = 0;
  uint32_t cacc = 0;
  for (uint32_t k = 0; k < U; ++k) {
    Q d = 0;
    for (uint32_t p = 0; p < db; ++p)
      if ((pd[p * B + k / 8] >> (7 - k % 8)) & 1) d |= Q(1) << (db - 1 - p);
    uint32_t cn = 0;
    for (uint32_t p = 0; p < cb; ++p)
      if ((pc[p * B + k / 8] >> (7 - k % 8)) & 1) cn |= 1u << (cb - 1 - p);
    acc = (k == 0) ? outlier : Q(acc + d);
    bidrun[k] = acc;
    cacc += cn;
    cend[k] = cacc;
  }
  const uint64_t base = c * kGroup;
  const uint32_t nvalid =
      static_cast<uint32_t>(std::min<uint64_t>(kGroup, n - base));
  if (cend[U - 1] < nvalid) return -4;
  const uint32_t sh = ctz_seg(seg);
  const Q m = static_cast<Q>(seg - 1u);
  uint32_t k = 0;
  for (uint32_t r = 0; r < nvalid; ++r) {
    while (r >= cend[k]) ++k;
    Q rel = 0;
    for (uint32_t p = 0; p < rb; ++p)
      if ((pr[p * 128 + r / 8] >> (r % 8)) & 1) rel |= Q(1) << (rb - 1 - p);
    const Q b = bidrun[k];
    const Q bx = urem<Q>(b, nbx);
    const Q by = urem<Q>(udiv<Q>(b, nbx), nby);
    const Q rx = rel & m;
    const Q ry = shr(rel, sh) & m;
    const Q qx = bx * static_cast<Q>(seg) + rx;
    const Q qy = by * static_cast<Q>(seg) + ry;
    x[base + r] = Tr::fma_(static_cast<T>((qx << 1) | 1), eb, mm[0]);
    y[base + r] = Tr::fma_(static_cast<T>((qy << 1) | 1), eb, mm[1]);
    if (dim == 3) {
      const Q bz = udiv<Q>(b, Q(nbx * nby));
      const Q rz = shr(rel, 2 * sh);
      const Q qz = bz * static_cast<Q>(seg) + rz;
      z[base + r] = Tr::fma_(static_cast<T>((qz << 1) | 1), eb, mm[2]);
    }
  }
  return 0;
}

template <typename T>
int decompress(const uint8_t* s, uint64_t size, T* x, T* y, T* z) {
  using Tr = Traits<T>;
  if (size < Tr::kHeader) return -4;
  uint32_t dim, bits, seg;
  uint64_t n;
  T eb;
  memcpy(&dim, s, 4);
  memcpy(&n, s + 4, 8);
  memcpy(&bits, s + 12, 4);
  memcpy(&eb, s + 16, sizeof(T));
  memcpy(&seg, s + 16 + sizeof(T), 4);
  if (bits != 8 * sizeof(T) || (dim != 2 && dim != 3) || n == 0 || seg == 0)
    return -4;
  const uint64_t G = (n + kGroup - 1) / kGroup;
  const uint64_t pay = Tr::kHeader + (8 + Tr::kMeta) * G;
  if (size < pay) return -4;
  int err = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(min : err)
  for (int64_t c = 0; c < static_cast<int64_t>(G); ++c) {
    const int e = decode_group<T>(dim, s, n, G, c, eb, seg, x, y, z,
                                  size - pay);
    err = std::min(err, e);
  }
  return err;
}

template <typename T>
void find_min_max(int dim, uint64_t n, const T* x, const T* y, const T* z,
                  T* mm) {
  // include/utils.cuh:44-65 ; the CLI zero-fills z for 2D input
  // (include/utils.cuh:141), so the z extent is 0 in 2D.
  T v[6] = {std::numeric_limits<float>::max(), std::numeric_limits<float>::max(),
            std::numeric_limits<float>::max(),
            std::numeric_limits<float>::lowest(),
            std::numeric_limits<float>::lowest(),
            std::numeric_limits<float>::lowest()};
  for (uint64_t i = 0; i < n; ++i) {
    const T zi = dim == 3 ? z[i] : T(0);
    v[0] = std::min(v[0], x[i]);
    v[1] = std::min(v[1], y[i]);
    v[2] = std::min(v[2], zi);
    v[3] = std::max(v[3], x[i]);
    v[4] = std::max(v[4], y[i]);
    v[5] = std::max(v[5], zi);
  }
  memcpy(mm, v, sizeof(v));
}

}  // namespace

extern "C" {

int oracle_version(void) { return 1; }

int oracle_num_threads(int want) {
#ifdef _OPENMP
  if (want > 0) omp_set_num_threads(want);
  return omp_get_max_threads();
#else
  (void)want;
  return 1;
#endif
}

uint64_t oracle_stream_bound(int float_bits, uint64_t n) {
  return float_bits == 64 ? stream_bytes_bound<double>(n)
                          : stream_bytes_bound<float>(n);
}

void oracle_min_max_f32(int dim, uint64_t n, const float* x, const float* y,
                        const float* z, float* mm) {
  find_min_max<float>(dim, n, x, y, z, mm);
}
void oracle_min_max_f64(int dim, uint64_t n, const double* x, const double* y,
                        const double* z, double* mm) {
  find_min_max<double>(dim, n, x, y, z, mm);
}

// src/main.cpp:35-55 (convertRelEb<T>): eb = max_range * e in T.
float oracle_rel_to_abs_f32(float e, const float* mm) {
  float r = std::fmax(mm[3] - mm[0], mm[4] - mm[1]);
  r = std::fmax(r, mm[5] - mm[2]);
  return r * e;
}
double oracle_rel_to_abs_f64(double e, const double* mm) {
  double r = std::fmax(mm[3] - mm[0], mm[4] - mm[1]);
  r = std::fmax(r, mm[5] - mm[2]);
  return r * e;
}

int oracle_seg_size_f32(int dim, uint64_t n, const float* x, const float* y,
                        const float* z, uint32_t seg_in, int is_rel,
                        float rel_eb, float eb, uint32_t* seg_out,
                        int* adjusted) {
  return seg_size<float>(dim, n, x, y, z, seg_in, is_rel, rel_eb, eb, seg_out,
                         adjusted);
}
int oracle_seg_size_f64(int dim, uint64_t n, const double* x, const double* y,
                        const double* z, uint32_t seg_in, int is_rel,
                        double rel_eb, double eb, uint32_t* seg_out,
                        int* adjusted) {
  return seg_size<double>(dim, n, x, y, z, seg_in, is_rel, rel_eb, eb, seg_out,
                          adjusted);
}

int oracle_compress_f32(int dim, uint64_t n, const float* x, const float* y,
                        const float* z, float eb, uint32_t seg, uint8_t* out,
                        uint64_t cap, uint64_t* out_size) {
  return compress<float>(dim, n, x, y, z, eb, seg, out, cap, out_size);
}
int oracle_compress_f64(int dim, uint64_t n, const double* x, const double* y,
                        const double* z, double eb, uint32_t seg, uint8_t* out,
                        uint64_t cap, uint64_t* out_size) {
  return compress<double>(dim, n, x, y, z, eb, seg, out, cap, out_size);
}

int oracle_decompress_f32(const uint8_t* s, uint64_t size, float* x, float* y,
                          float* z) {
  return decompress<float>(s, size, x, y, z);
}
int oracle_decompress_f64(const uint8_t* s, uint64_t size, double* x,
                          double* y, double* z) {
  return decompress<double>(s, size, x, y, z);
}

// Stage dump of one group (for tests): q[3*1024], bid[1024], rel[1024] as
// u64, perm[1024] (sorted rank -> slot), info[9].
int oracle_group_stages_f32(int dim, uint64_t n, const float* x,
                            const float* y, const float* z, float eb,
                            uint32_t seg, uint64_t c, uint64_t* q,
                            uint64_t* bid, uint64_t* rel, uint16_t* perm,
                            uint32_t* info) {
  const uint64_t G = (n + kGroup - 1) / kGroup;
  if (c >= G || seg == 0) return -2;
  GroupOut<float> g;
  StageDump d{q, bid, rel, perm, info};
  encode_group<float>(dim, x, y, z, n, c, G, eb, seg, g, &d);
  return 0;
}
int oracle_group_stages_f64(int dim, uint64_t n, const double* x,
                            const double* y, const double* z, double eb,
                            uint32_t seg, uint64_t c, uint64_t* q,
                            uint64_t* bid, uint64_t* rel, uint16_t* perm,
                            uint32_t* info) {
  const uint64_t G = (n + kGroup - 1) / kGroup;
  if (c >= G || seg == 0) return -2;
  GroupOut<double> g;
  StageDump d{q, bid, rel, perm, info};
  encode_group<double>(dim, x, y, z, n, c, G, eb, seg, g, &d);
  return 0;
}

}  // extern "C"
