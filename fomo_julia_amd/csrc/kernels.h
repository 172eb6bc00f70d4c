// gfx950 kernels of the elastic time step.  Header-only templates, included by
// the single translation unit fomo_mi355x.hip.
//
// Arithmetic contract (shared with the CPU oracle's FMA mode, so FP32 results
// are bit-identical): every `acc += a*b` is fma(a, b, acc); every
// `e1*f1 + e2*f2` is fma(e2, f2, e1*f1); everything else is one IEEE op.  The
// file is compiled with -ffp-contract=off so no other contraction happens.
//
// Work distribution: each kernel walks a flat list of work items (a prefix-sum
// table over the shots of a batch).  Items are per wavefront, so one launch
// can mix shots of different grid sizes; gridDim only bounds concurrency
// (items beyond it are taken in later loop trips, i.e. a persistent grid).
#pragma once
#include <hip/hip_runtime.h>
#include "fomo_device.h"

namespace fomo {

// FLAG_EDGE / FLAG_INT: a fused launch over the edge or interior tiles of
// the split step (ShotDev::xi0 ...); neither: the whole grid.
enum { FLAG_BACKUP = 1, FLAG_INJECT = 2, FLAG_EDGE = 4, FLAG_INT = 8 };
enum { MODE_FS = 1, MODE_RECORD = 2 };

__device__ __forceinline__ float fmaf_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// LDS written by one wavefront and read back by other lanes of the same
// wavefront: DS ops of a wave execute in order, so only the compiler needs to
// be kept from reordering.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Whole-wavefront lane shifts by one (DPP wave_shr:1 / wave_shl:1): a VALU
// operand modifier, no LDS round trip.  The lane with no source gets 0.
__device__ __forceinline__ float from_lower_lane(float v) {  // lane i <- lane i-1
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, false));
}
// Same shift with the hardware's bound control filling lane 0 with 0 (no
// old-value operand to initialise).
__device__ __forceinline__ float from_lower_lane_z(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ float from_upper_lane(float v) {  // lane i <- lane i+1
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, false));
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Workgroups are dealt to the 8 XCDs round-robin (block b -> XCD b % 8).  For a
// grid that is a multiple of 8 blocks, this renumbering gives each XCD one
// contiguous run of logical blocks, so neighbouring tiles (which share halo
// rows and columns) meet in the same L2.
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int xcd_block() {
  const int g = gridDim.x, b = blockIdx.x;
  if (g % NUM_XCD) return b;
  return (b % NUM_XCD) * (g / NUM_XCD) + b / NUM_XCD;
}
__device__ __forceinline__ float uniform_f(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
// A pointer read from the shot table, made wave-uniform (kept in SGPRs).
template <class T>
__device__ __forceinline__ T* uniform_p(T* p) {
  const uint64_t v = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (T*)(((uint64_t)hi << 32) | lo);
}

// Shot owning item: largest s with prefix[s] <= item (prefix[n] == total).
__device__ __forceinline__ int find_shot(const int* prefix, int n, int item) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (prefix[mid] <= item) lo = mid;
    else hi = mid;
  }
  return lo;
}

template <int N>
struct V {
  float v[N];
};
typedef float f2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));

// Field pointers come from the shot table as generic pointers; accessing
// them through the global address space keeps the loads off the LDS counter
// (flat loads would make every LDS wait drain the prefetched rows too).
typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) f2v gf2v;
typedef __attribute__((address_space(1))) f4v gf4v;

__device__ __forceinline__ float ldg(const float* p) { return *(const gfloat*)p; }
__device__ __forceinline__ void stg(float* p, float v) { *(gfloat*)p = v; }

template <int N>
__device__ __forceinline__ V<N> ldv(const float* p);
template <>
__device__ __forceinline__ V<1> ldv<1>(const float* p) {
  V<1> r;
  r.v[0] = ldg(p);
  return r;
}
template <>
__device__ __forceinline__ V<2> ldv<2>(const float* p) {
  f2v t = *(const gf2v*)p;
  V<2> r;
  r.v[0] = t.x;
  r.v[1] = t.y;
  return r;
}
template <>
__device__ __forceinline__ V<4> ldv<4>(const float* p) {
  f4v t = *(const gf4v*)p;
  V<4> r;
  r.v[0] = t.x;
  r.v[1] = t.y;
  r.v[2] = t.z;
  r.v[3] = t.w;
  return r;
}
template <int N>
__device__ __forceinline__ void stv(float* p, const V<N>& a);
template <>
__device__ __forceinline__ void stv<1>(float* p, const V<1>& a) {
  stg(p, a.v[0]);
}
template <>
__device__ __forceinline__ void stv<2>(float* p, const V<2>& a) {
  f2v t;
  t.x = a.v[0];
  t.y = a.v[1];
  *(gf2v*)p = t;
}
template <>
__device__ __forceinline__ void stv<4>(float* p, const V<4>& a) {
  f4v t;
  t.x = a.v[0];
  t.y = a.v[1];
  t.z = a.v[2];
  t.w = a.v[3];
  *(gf4v*)p = t;
}

// ---- storage traits: FP32 is the identity (no scale multiply), BF16 rounds
// to nearest even, FP16 uses the hardware conversion (round to nearest even).
// F24 keeps FP32 values rounded to a 16-bit significand (nearest, ties to
// even): bits 31..16 in the field's 16-bit plane, bits 15..8 in its byte
// plane, bits 7..0 zero.  The byte plane of a wavefield is a half-size image
// of its 16-bit planes: the byte of the element at address p is at
// (p >> 1) + loc (ShotDev::loc), so one constant per wavefield serves every
// field, buffer and row.
template <int K>
struct Store;
template <>
struct Store<ST_F32> {
  typedef float T;
};
template <>
struct Store<ST_BF16> {
  typedef uint16_t T;
};
template <>
struct Store<ST_F16> {
  typedef uint16_t T;
};
template <>
struct Store<ST_F24> {
  typedef uint16_t T;  // the 16-bit plane
};

// Element kinds of the fields (F) and of the coefficient arrays (C) per mode.
template <int K>
struct Mode {
  static constexpr int F = K, C = K;
};
template <>
struct Mode<ST_F16C> {
  static constexpr int F = ST_F32, C = ST_F16;
};
template <>
struct Mode<ST_F24> {
  static constexpr int F = ST_F24, C = ST_F32;
};

// The per-field argument of the load and store helpers ("codec"): the scale
// of a 16-bit kind (unused for FP32, a float either way), and for F24 the
// wavefield's byte-plane constant loc (pointer helpers) or a buffer
// descriptor of the byte-plane image (buffer helpers).
typedef __attribute__((address_space(1))) uint8_t gu8;
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}
// codec of a pointer helper: scale sc, or loc for F24
template <int K>
__device__ __forceinline__ auto pcodec(float sc, uint64_t loc) {
  if constexpr (K == ST_F24) return uniform_u64(loc);
  else return uniform_f(sc);
}
// F24: x rounded to a 16-bit significand (nearest, ties to even), as FP32
// bits; infinities stay, finite values past the largest one round to them
__device__ __forceinline__ uint32_t f24_bits(float x) {
  const uint32_t u = __float_as_uint(x);
  return (u + 0x7Fu + ((u >> 8) & 1u)) & 0xFFFFFF00u;
}
__device__ __forceinline__ const gu8* f24_lo(const void* p, uint64_t loc) {
  return (const gu8*)((((uint64_t)p) >> 1) + loc);
}
__device__ __forceinline__ gu8* f24_lo_w(void* p, uint64_t loc) {
  return (gu8*)((((uint64_t)p) >> 1) + loc);
}

template <int K>
__device__ __forceinline__ float from_bits(uint32_t b) {
  if constexpr (K == ST_BF16) return __uint_as_float(b << 16);
  else return (float)__builtin_bit_cast(_Float16, (uint16_t)b);
}
// FP32 -> 16-bit bits, rounded to nearest (ties to even).  Both use the
// hardware conversions (v_cvt_pk_bf16_f32, v_cvt_f16_f32); for every non-NaN
// input the BF16 result equals the integer rounding u + 0x7FFF + ((u >> 16) & 1)
// taken >> 16, which the oracle and the host encoder use.
template <int K>
__device__ __forceinline__ uint32_t to_bits(float x) {
  if constexpr (K == ST_BF16) return (uint32_t)__builtin_bit_cast(uint16_t, (__bf16)x);
  else return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)x);
}
typedef float f2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf2_t __attribute__((ext_vector_type(2)));
// Two values packed low | high << 16 (one v_cvt_pk_bf16_f32 for BF16).
template <int K>
__device__ __forceinline__ uint32_t to_bits2(float x0, float x1) {
  if constexpr (K == ST_BF16) {
    f2_t v;
    v.x = x0;
    v.y = x1;
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf2_t));
  } else {
    return to_bits<K>(x0) | (to_bits<K>(x1) << 16);
  }
}

typedef __attribute__((address_space(1))) uint16_t gu16;
typedef __attribute__((address_space(1))) uint32_t gu32;
typedef uint32_t u2v __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) u2v gu2v;

// x * s for a stored kind: BF16 arrays always carry scale 1 (they keep FP32's
// range), so the product is dropped there; x * 1 is x exactly.
template <int K>
__device__ __forceinline__ float scl(float x, float s) {
  if constexpr (K == ST_BF16) return x;
  else return x * s;
}

// N consecutive stored values at p, as floats (x stored * s).
template <int K, int N, class S>
__device__ __forceinline__ V<N> ldf(const typename Store<K>::T* p, S s) {
  if constexpr (K == ST_F32) {
    return ldv<N>(p);
  } else if constexpr (K == ST_F24) {
    static_assert(sizeof(S) == 8, "F24 pointer helpers take the byte-plane constant");
    V<N> r;
    const gu8* q = f24_lo(p, s);
#pragma unroll
    for (int c = 0; c < N; ++c)
      r.v[c] = __uint_as_float(((uint32_t)((const gu16*)p)[c] << 16) | ((uint32_t)q[c] << 8));
    return r;
  } else {
    V<N> r;
    if constexpr (N == 1) {
      r.v[0] = scl<K>(from_bits<K>(*(const gu16*)p), s);
    } else if constexpr (N == 2) {
      const uint32_t w = *(const gu32*)p;
      r.v[0] = scl<K>(from_bits<K>(w & 0xFFFFu), s);
      r.v[1] = scl<K>(from_bits<K>(w >> 16), s);
    } else {
      const u2v w = *(const gu2v*)p;
      r.v[0] = scl<K>(from_bits<K>(w.x & 0xFFFFu), s);
      r.v[1] = scl<K>(from_bits<K>(w.x >> 16), s);
      r.v[2] = scl<K>(from_bits<K>(w.y & 0xFFFFu), s);
      r.v[3] = scl<K>(from_bits<K>(w.y >> 16), s);
    }
    return r;
  }
}
template <int K, int N, class S>
__device__ __forceinline__ void stf(typename Store<K>::T* p, const V<N>& a, S s) {
  if constexpr (K == ST_F32) {
    stv<N>(p, a);
  } else if constexpr (K == ST_F24) {
    static_assert(sizeof(S) == 8, "F24 pointer helpers take the byte-plane constant");
    gu8* q = f24_lo_w(p, s);
#pragma unroll
    for (int c = 0; c < N; ++c) {
      const uint32_t b = f24_bits(a.v[c]);
      ((gu16*)p)[c] = (uint16_t)(b >> 16);
      q[c] = (uint8_t)(b >> 8);
    }
  } else if constexpr (N == 1) {
    *(gu16*)p = (uint16_t)to_bits<K>(scl<K>(a.v[0], s));
  } else if constexpr (N == 2) {
    *(gu32*)p = to_bits2<K>(scl<K>(a.v[0], s), scl<K>(a.v[1], s));
  } else {
    u2v w;
    w.x = to_bits2<K>(scl<K>(a.v[0], s), scl<K>(a.v[1], s));
    w.y = to_bits2<K>(scl<K>(a.v[2], s), scl<K>(a.v[3], s));
    *(gu2v*)p = w;
  }
}
template <int K, class S>
__device__ __forceinline__ float ld1(const typename Store<K>::T* p, S s) {
  if constexpr (K == ST_F32) return ldg(p);
  else if constexpr (K == ST_F24)
    return __uint_as_float(((uint32_t)*(const gu16*)p << 16) | ((uint32_t)*f24_lo(p, s) << 8));
  else return scl<K>(from_bits<K>(*(const gu16*)p), s);
}
// The stored bits of one element (16-bit kinds zero-extended) and their value:
// staging loops load bits unconditionally and decode afterwards, so that all
// loads of a chunk are in flight before the first wait (a conditional load
// whose value is decoded in the same branch makes the compiler wait on it
// right there).
template <int K, class S>
__device__ __forceinline__ uint32_t ld_bits(const typename Store<K>::T* p, S s) {
  if constexpr (K == ST_F32) return __float_as_uint(ldg(p));
  else if constexpr (K == ST_F24) return ((uint32_t)*(const gu16*)p << 16) | ((uint32_t)*f24_lo(p, s) << 8);
  else return (uint32_t)*(const gu16*)p;
}
template <int K, class S>
__device__ __forceinline__ float from_stored(uint32_t b, S s) {
  if constexpr (K == ST_F32 || K == ST_F24) return __uint_as_float(b);
  else return scl<K>(from_bits<K>(b), s);
}
template <int K, class S>
__device__ __forceinline__ void st1(typename Store<K>::T* p, float x, S s) {
  if constexpr (K == ST_F32) {
    stg(p, x);
  } else if constexpr (K == ST_F24) {
    const uint32_t b = f24_bits(x);
    *(gu16*)p = (uint16_t)(b >> 16);
    *f24_lo_w(p, s) = (uint8_t)(b >> 8);
  } else {
    *(gu16*)p = (uint16_t)to_bits<K>(scl<K>(x, s));
  }
}

// ---- buffer addressing: a descriptor (base + length in SGPRs), a per-lane
// byte offset and a uniform byte offset.  Loads past the length return 0 and
// stores past it are dropped, so a lane is masked by giving it offset OOB:
// every load and store stays unconditional, which keeps the compiler's
// memory-counter waits exact across rows.
typedef __amdgpu_buffer_rsrc_t Rsrc;
constexpr uint32_t OOB = 0x80000000u;  // >= every buffer length used (< 2^31)
__device__ __forceinline__ Rsrc make_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
// codec of a buffer helper: scale sc, or the byte-plane descriptor for F24
template <int K>
__device__ __forceinline__ auto bcodec(float sc, Rsrc lo) {
  if constexpr (K == ST_F24) return lo;
  else return uniform_f(sc);
}
// F24 through buffers: the codec is a descriptor of the byte-plane image of
// r's range (base (r's base >> 1) + loc), addressed with offsets halved.
template <int K, int N, class S>
__device__ __forceinline__ V<N> bload(Rsrc r, uint32_t vo, uint32_t so, S s) {
  V<N> x;
  if constexpr (K == ST_F24) {
    static_assert(!__is_same(S, float), "F24 buffer helpers take the byte-plane descriptor");
    const uint32_t lv = vo >> 1, ls = so >> 1;
    if constexpr (N == 1) {
      const uint32_t h = __builtin_amdgcn_raw_buffer_load_b16(r, vo, so, 0);
      const uint32_t l = __builtin_amdgcn_raw_buffer_load_b8(s, lv, ls, 0);
      x.v[0] = __uint_as_float((h << 16) | (l << 8));
    } else if constexpr (N == 2) {
      const uint32_t h = __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0);
      const uint32_t l = __builtin_amdgcn_raw_buffer_load_b16(s, lv, ls, 0);
      x.v[0] = __uint_as_float((h << 16) | ((l & 0xFFu) << 8));
      x.v[1] = __uint_as_float((h & 0xFFFF0000u) | (l & 0xFF00u));
    } else {
      const auto h = __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0);
      const uint32_t l = __builtin_amdgcn_raw_buffer_load_b32(s, lv, ls, 0);
      const uint32_t h0 = (uint32_t)h[0], h1 = (uint32_t)h[1];
      x.v[0] = __uint_as_float((h0 << 16) | ((l & 0xFFu) << 8));
      x.v[1] = __uint_as_float((h0 & 0xFFFF0000u) | (l & 0xFF00u));
      x.v[2] = __uint_as_float((h1 << 16) | ((l >> 8) & 0xFF00u));
      x.v[3] = __uint_as_float((h1 & 0xFFFF0000u) | ((l >> 16) & 0xFF00u));
    }
  } else if constexpr (K == ST_F32) {
    if constexpr (N == 1) {
      x.v[0] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
    } else if constexpr (N == 2) {
      // elements go through a scalar: clang's __builtin_bit_cast of an
      // ext-vector element reads element 0 whatever the index
      const auto w = __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0);
      x.v[0] = __uint_as_float((uint32_t)w[0]);
      x.v[1] = __uint_as_float((uint32_t)w[1]);
    } else {
      const auto w = __builtin_amdgcn_raw_buffer_load_b128(r, vo, so, 0);
#pragma unroll
      for (int c = 0; c < 4; ++c) x.v[c] = __uint_as_float((uint32_t)w[c]);
    }
  } else {
    if constexpr (N == 1) {
      x.v[0] = scl<K>(from_bits<K>(__builtin_amdgcn_raw_buffer_load_b16(r, vo, so, 0)), s);
    } else if constexpr (N == 2) {
      const uint32_t w = __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0);
      x.v[0] = scl<K>(from_bits<K>(w & 0xFFFFu), s);
      x.v[1] = scl<K>(from_bits<K>(w >> 16), s);
    } else {
      const auto w = __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0);
      x.v[0] = scl<K>(from_bits<K>(w[0] & 0xFFFFu), s);
      x.v[1] = scl<K>(from_bits<K>(w[0] >> 16), s);
      x.v[2] = scl<K>(from_bits<K>(w[1] & 0xFFFFu), s);
      x.v[3] = scl<K>(from_bits<K>(w[1] >> 16), s);
    }
  }
  return x;
}
template <int K, class S>
__device__ __forceinline__ void bstore(Rsrc r, uint32_t vo, uint32_t so, float x, S s) {
  if constexpr (K == ST_F32) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, x), r, vo, so, 0);
  } else if constexpr (K == ST_F24) {
    const uint32_t b = f24_bits(x);
    __builtin_amdgcn_raw_buffer_store_b16((unsigned short)(b >> 16), r, vo, so, 0);
    __builtin_amdgcn_raw_buffer_store_b8((unsigned char)(b >> 8), s, vo >> 1, so >> 1, 0);
  } else {
    __builtin_amdgcn_raw_buffer_store_b16((unsigned short)to_bits<K>(scl<K>(x, s)), r, vo, so, 0);
  }
}

// A lane's two adjacent elements at one offset: one 64-bit (FP32) or 32-bit
// (16-bit kinds) store, the same bits as two bstore calls.
template <int K, class S>
__device__ __forceinline__ void bstore2(Rsrc r, uint32_t vo, uint32_t so, float x0, float x1, S s) {
  if constexpr (K == ST_F32) {
    u2v w;
    w.x = __float_as_uint(x0);
    w.y = __float_as_uint(x1);
    __builtin_amdgcn_raw_buffer_store_b64(w, r, vo, so, 0);
  } else if constexpr (K == ST_F24) {
    const uint32_t b0 = f24_bits(x0), b1 = f24_bits(x1);
    __builtin_amdgcn_raw_buffer_store_b32((b0 >> 16) | (b1 & 0xFFFF0000u), r, vo, so, 0);
    __builtin_amdgcn_raw_buffer_store_b16((unsigned short)(((b0 >> 8) & 0xFFu) | (b1 & 0xFF00u)), s, vo >> 1,
                                          so >> 1, 0);
  } else {
    __builtin_amdgcn_raw_buffer_store_b32(to_bits2<K>(scl<K>(x0, s), scl<K>(x1, s)), r, vo, so, 0);
  }
}

// x as the storage holds it (stored with scale s, read back with is = 1/s).
template <int K, class S>
__device__ __forceinline__ float as_stored(float x, S s, S is) {
  if constexpr (K == ST_F32) return x;
  else if constexpr (K == ST_F24) return __uint_as_float(f24_bits(x));
  else return scl<K>(from_bits<K>(to_bits<K>(scl<K>(x, s))), is);
}

template <class T>
__device__ __forceinline__ const T* rowt(const T* base, int j, int pitch) {
  return base + (ptrdiff_t)j * pitch;
}
template <class T>
__device__ __forceinline__ T* rowt(T* base, int j, int pitch) {
  return base + (ptrdiff_t)j * pitch;
}

__device__ __forceinline__ const float* row(const float* base, int j, int pitch) {
  return base + (ptrdiff_t)j * pitch;
}
__device__ __forceinline__ float* row(float* base, int j, int pitch) {
  return base + (ptrdiff_t)j * pitch;
}

// HABC strip membership of copy_boundary_strip! (kernels/boundary.jl:20-38), 0-based.
__device__ __forceinline__ bool in_strip(int nx, int nz, int n, int fs, int i, int j) {
  const int jtop = fs ? n : 0;
  return j >= jtop && j < nz && i < nx &&
         (i <= n + 1 || i >= nx - n - 2 || j <= n + 1 || j >= nz - n - 2);
}
__device__ __forceinline__ bool in_strip(const ShotDev& sd, int i, int j) {
  return in_strip(sd.nx, sd.nz, sd.nbc, sd.fs, i, j);
}

// Step number (1-based) this launch works on for shot sd, or 0 if the shot is idle.
__device__ __forceinline__ int shot_step(const ShotDev& sd, int k_off) {
  int k = *sd.kp + k_off;
  return (k >= 1 && k <= sd.nt) ? k : 0;
}

// ---------------------------------------------------------------------------
// Velocity update (kernels/velocity.jl:21-46) with the strip backup of vx, vz
// (boundary.jl:45-56) and the source injection (source_receiver.jl:18-30;
// force sources backend/cpu.jl:171-187) fused in.
//
// A wavefront owns a tile of 64*CPT columns x tz rows and marches down z.
// Per row: txx and txz of the row go through LDS for the x-derivatives; txz
// and tzz columns are held in register windows of 2M rows for the
// z-derivatives; the next row's loads are issued before the current row is
// computed.  Cells outside [M, n-M) are visited too (strip backup) but keep
// their value.  A pressure source is added to the local copies of txx/tzz
// only; the stress kernel adds it to memory, so every reader sees
// txx + w exactly once.
// ---------------------------------------------------------------------------
template <int M, int CPT, int WPB, bool SAME>
__global__ __launch_bounds__(64 * WPB) void k_velocity(BatchArgs ba, int flags) {
  constexpr int TW = 64 * CPT;
  constexpr int LW = TW + 2 * M;
  __shared__ float lds[WPB][2][LW];
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  float* s_txx = lds[wid][0];
  float* s_txz = lds[wid][1];
  const int nsh = *ba.n_shots_p;
  const int total = ba.prefix[nsh];
  for (int item = uniform(blockIdx.x * WPB + wid); item < total;
       item = uniform(item + gridDim.x * WPB)) {
    const int s = uniform(find_shot(ba.prefix, nsh, item));
    const ShotDev& sd = ba.shots[s];
    const int k = shot_step(sd, ba.k_off);
    if (k == 0) continue;
    // Everything the row loop needs, read once from the shot table into
    // wave-uniform registers (the loop's stores could alias the table).
    const int t = item - ba.prefix[s];
    const int nx = uniform(sd.nx), nz = uniform(sd.nz), P = uniform(sd.pitch);
    const int nbc = uniform(sd.nbc), fs = uniform(sd.fs);
    const int tiles_x = uniform(sd.tiles_x), tz = uniform(sd.tz);
    const int i0 = (t % tiles_x) * TW;
    const int j0 = (t / tiles_x) * tz;
    const int j1 = min(j0 + tz, nz);
    const int ib = i0 + lane * CPT;
    int hx = -1;
    if (lane < M) hx = lane;
    else if (lane >= 64 - M) hx = TW + M + (lane - (64 - M));
    const bool has_halo = hx >= 0;
    const int hcol = i0 - M + hx;

    float* vx = uniform_p(sd.f[F_VX]);
    float* vz = uniform_p(sd.f[F_VZ]);
    const float* txx = uniform_p(sd.f[F_TXX]);
    const float* tzz = uniform_p(sd.f[F_TZZ]);
    const float* txz = uniform_p(sd.f[F_TXZ]);
    float* ovx = uniform_p(sd.fo[F_VX]);
    float* ovz = uniform_p(sd.fo[F_VZ]);
    const float* b1p = uniform_p(sd.bvx_x);
    const float* b2p = uniform_p(sd.bvx_z);
    const float* b3p = uniform_p(sd.bvz_x);
    const float* b4p = uniform_p(sd.bvz_z);
    const int nw = uniform(sd.nw), stype = uniform(sd.stype);
    const bool src_on = (flags & FLAG_INJECT) && k <= nw;
    const float wav = src_on ? uniform_f(uniform_p(sd.wavelet)[k - 1]) : 0.f;
    const bool src_p = src_on && stype == SRC_PRESSURE;
    const bool src_vz = src_on && stype == SRC_FORCE_Z;
    const bool src_vx = src_on && stype == SRC_FORCE_X;
    const int si = uniform(sd.si), sj = uniform(sd.sj);
    const float src_scale = uniform_f(sd.src_scale);
    float a[M];
#pragma unroll
    for (int l = 0; l < M; ++l) a[l] = uniform_f(sd.a[l]);
    // ---- end of shot-table reads

    float wtxz[CPT][2 * M];  // txz rows j-M .. j+M-1
    float wtzz[CPT][2 * M];  // tzz rows j-M+1 .. j+M
#pragma unroll
    for (int q = 0; q < 2 * M; ++q) {
      V<CPT> x = ldv<CPT>(row(txz, j0 - M + q, P) + ib);
      V<CPT> z = ldv<CPT>(row(tzz, j0 - M + 1 + q, P) + ib);
#pragma unroll
      for (int c = 0; c < CPT; ++c) {
        wtxz[c][q] = x.v[c];
        wtzz[c][q] = z.v[c] + ((src_p && j0 - M + 1 + q == sj && ib + c == si) ? wav : 0.f);
      }
    }
    V<CPT> c_txx = ldv<CPT>(row(txx, j0, P) + ib);
    float c_htxx = has_halo ? ldg(row(txx, j0, P) + hcol) : 0.f;
    float c_htxz = has_halo ? ldg(row(txz, j0, P) + hcol) : 0.f;
    V<CPT> c_vx = ldv<CPT>(row(vx, j0, P) + ib);
    V<CPT> c_vz = ldv<CPT>(row(vz, j0, P) + ib);
    V<CPT> c_b1 = ldv<CPT>(row(b1p, j0, P) + ib);
    V<CPT> c_b3 = ldv<CPT>(row(b3p, j0, P) + ib);
    V<CPT> c_b2 = c_b1, c_b4 = c_b3;
    if constexpr (!SAME) {
      c_b2 = ldv<CPT>(row(b2p, j0, P) + ib);
      c_b4 = ldv<CPT>(row(b4p, j0, P) + ib);
    }

    for (int j = j0; j < j1; ++j) {
      // Prefetch row j+1 (and the window tops) while row j is computed.
      const int jn = j + 1;
      V<CPT> n_txx = ldv<CPT>(row(txx, jn, P) + ib);
      float n_htxx = has_halo ? ldg(row(txx, jn, P) + hcol) : 0.f;
      float n_htxz = has_halo ? ldg(row(txz, jn, P) + hcol) : 0.f;
      V<CPT> n_txz = ldv<CPT>(row(txz, j + M, P) + ib);
      V<CPT> n_tzz = ldv<CPT>(row(tzz, j + M + 1, P) + ib);
      V<CPT> n_vx = ldv<CPT>(row(vx, jn, P) + ib);
      V<CPT> n_vz = ldv<CPT>(row(vz, jn, P) + ib);
      V<CPT> n_b1 = ldv<CPT>(row(b1p, jn, P) + ib);
      V<CPT> n_b3 = ldv<CPT>(row(b3p, jn, P) + ib);
      V<CPT> n_b2 = n_b1, n_b4 = n_b3;
      if constexpr (!SAME) {
        n_b2 = ldv<CPT>(row(b2p, jn, P) + ib);
        n_b4 = ldv<CPT>(row(b4p, jn, P) + ib);
      }

      const bool src_row = src_p && j == sj;
#pragma unroll
      for (int c = 0; c < CPT; ++c) {
        s_txx[M + lane * CPT + c] = c_txx.v[c] + ((src_row && ib + c == si) ? wav : 0.f);
        s_txz[M + lane * CPT + c] = wtxz[c][M];
      }
      if (has_halo) {
        s_txx[hx] = c_htxx + ((src_row && hcol == si) ? wav : 0.f);
        s_txz[hx] = c_htxz;
      }
      wave_sync();

      V<CPT> o_vx, o_vz;
#pragma unroll
      for (int c = 0; c < CPT; ++c) {
        const int x = M + lane * CPT + c;
        float d1 = 0.f, d2 = 0.f, d3 = 0.f, d4 = 0.f;
#pragma unroll
        for (int l = 1; l <= M; ++l) {
          d1 = fmaf_(a[l - 1], s_txx[x + l - 1] - s_txx[x - l], d1);
          d2 = fmaf_(a[l - 1], wtxz[c][M + l - 1] - wtxz[c][M - l], d2);
          d3 = fmaf_(a[l - 1], s_txz[x + l] - s_txz[x - l + 1], d3);
          d4 = fmaf_(a[l - 1], wtzz[c][M - 1 + l] - wtzz[c][M - l], d4);
        }
        const int i = ib + c;
        const bool inside = (i >= M) & (i < nx - M) & (j >= M) & (j < nz - M);
        float vxs = c_vx.v[c], vzs = c_vz.v[c];
        if (i == si && j == sj) {
          if (src_vz) vzs = vzs + wav * src_scale;
          if (src_vx) vxs = vxs + wav * src_scale;
        }
        const float ix = fmaf_(c_b2.v[c], d2, c_b1.v[c] * d1);
        const float iz = fmaf_(c_b4.v[c], d4, c_b3.v[c] * d3);
        o_vx.v[c] = inside ? vxs + ix : vxs;
        o_vz.v[c] = inside ? vzs + iz : vzs;
      }
      stv<CPT>(row(vx, j, P) + ib, o_vx);
      stv<CPT>(row(vz, j, P) + ib, o_vz);
      if (flags & FLAG_BACKUP) {
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
          if (in_strip(nx, nz, nbc, fs, ib + c, j)) {
            stg(row(ovx, j, P) + ib + c, c_vx.v[c]);
            stg(row(ovz, j, P) + ib + c, c_vz.v[c]);
          }
        }
      }
      wave_sync();

      const bool src_top = src_p && j + M + 1 == sj;
#pragma unroll
      for (int c = 0; c < CPT; ++c) {
#pragma unroll
        for (int q = 0; q < 2 * M - 1; ++q) {
          wtxz[c][q] = wtxz[c][q + 1];
          wtzz[c][q] = wtzz[c][q + 1];
        }
        wtxz[c][2 * M - 1] = n_txz.v[c];
        wtzz[c][2 * M - 1] = n_tzz.v[c] + ((src_top && ib + c == si) ? wav : 0.f);
      }
      c_txx = n_txx;
      c_htxx = n_htxx;
      c_htxz = n_htxz;
      c_vx = n_vx;
      c_vz = n_vz;
      c_b1 = n_b1;
      c_b2 = n_b2;
      c_b3 = n_b3;
      c_b4 = n_b4;
    }
  }
}

// ---------------------------------------------------------------------------
// Stress update (kernels/stress.jl:18-47) with the strip backup of txx, tzz,
// txz fused in, and the pressure source added to txx/tzz in memory.
// Same tiling as the velocity kernel: vx and vz rows through LDS for the
// x-derivatives, register windows for the z-derivatives.
// ---------------------------------------------------------------------------
template <int M, int CPT, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_stress(BatchArgs ba, int flags) {
  constexpr int TW = 64 * CPT;
  constexpr int LW = TW + 2 * M;
  __shared__ float lds[WPB][2][LW];
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  float* s_vx = lds[wid][0];
  float* s_vz = lds[wid][1];
  const int nsh = *ba.n_shots_p;
  const int total = ba.prefix[nsh];
  for (int item = uniform(blockIdx.x * WPB + wid); item < total;
       item = uniform(item + gridDim.x * WPB)) {
    const int s = uniform(find_shot(ba.prefix, nsh, item));
    const ShotDev& sd = ba.shots[s];
    const int k = shot_step(sd, ba.k_off);
    if (k == 0) continue;
    const int t = item - ba.prefix[s];
    const int nx = uniform(sd.nx), nz = uniform(sd.nz), P = uniform(sd.pitch);
    const int nbc = uniform(sd.nbc), fs = uniform(sd.fs);
    const int tiles_x = uniform(sd.tiles_x), tz = uniform(sd.tz);
    const int i0 = (t % tiles_x) * TW;
    const int j0 = (t / tiles_x) * tz;
    const int j1 = min(j0 + tz, nz);
    const int ib = i0 + lane * CPT;
    int hx = -1;
    if (lane < M) hx = lane;
    else if (lane >= 64 - M) hx = TW + M + (lane - (64 - M));
    const bool has_halo = hx >= 0;
    const int hcol = i0 - M + hx;

    const float* vx = uniform_p(sd.f[F_VX]);
    const float* vz = uniform_p(sd.f[F_VZ]);
    float* txx = uniform_p(sd.f[F_TXX]);
    float* tzz = uniform_p(sd.f[F_TZZ]);
    float* txz = uniform_p(sd.f[F_TXZ]);
    float* otxx = uniform_p(sd.fo[F_TXX]);
    float* otzz = uniform_p(sd.fo[F_TZZ]);
    float* otxz = uniform_p(sd.fo[F_TXZ]);
    const float* lamp = uniform_p(sd.lam);
    const float* l2mp = uniform_p(sd.l2m);
    const float* mup = uniform_p(sd.mu_txz);
    const int nw = uniform(sd.nw), stype = uniform(sd.stype);
    const bool src_on = (flags & FLAG_INJECT) && k <= nw && stype == SRC_PRESSURE;
    const float wav = src_on ? uniform_f(uniform_p(sd.wavelet)[k - 1]) : 0.f;
    const int si = uniform(sd.si), sj = uniform(sd.sj);
    const float dtx = uniform_f(sd.dtx), dtz = uniform_f(sd.dtz);
    float a[M];
#pragma unroll
    for (int l = 0; l < M; ++l) a[l] = uniform_f(sd.a[l]);
    // ---- end of shot-table reads

    float wvx[CPT][2 * M];  // vx rows j-M+1 .. j+M
    float wvz[CPT][2 * M];  // vz rows j-M .. j+M-1
#pragma unroll
    for (int q = 0; q < 2 * M; ++q) {
      V<CPT> x = ldv<CPT>(row(vx, j0 - M + 1 + q, P) + ib);
      V<CPT> z = ldv<CPT>(row(vz, j0 - M + q, P) + ib);
#pragma unroll
      for (int c = 0; c < CPT; ++c) {
        wvx[c][q] = x.v[c];
        wvz[c][q] = z.v[c];
      }
    }
    float c_hvx = has_halo ? ldg(row(vx, j0, P) + hcol) : 0.f;
    float c_hvz = has_halo ? ldg(row(vz, j0, P) + hcol) : 0.f;
    V<CPT> c_lam = ldv<CPT>(row(lamp, j0, P) + ib);
    V<CPT> c_l2m = ldv<CPT>(row(l2mp, j0, P) + ib);
    V<CPT> c_mu = ldv<CPT>(row(mup, j0, P) + ib);
    V<CPT> c_txx = ldv<CPT>(row(txx, j0, P) + ib);
    V<CPT> c_tzz = ldv<CPT>(row(tzz, j0, P) + ib);
    V<CPT> c_txz = ldv<CPT>(row(txz, j0, P) + ib);

    for (int j = j0; j < j1; ++j) {
      const int jn = j + 1;
      float n_hvx = has_halo ? ldg(row(vx, jn, P) + hcol) : 0.f;
      float n_hvz = has_halo ? ldg(row(vz, jn, P) + hcol) : 0.f;
      V<CPT> n_vx = ldv<CPT>(row(vx, j + M + 1, P) + ib);
      V<CPT> n_vz = ldv<CPT>(row(vz, j + M, P) + ib);
      V<CPT> n_lam = ldv<CPT>(row(lamp, jn, P) + ib);
      V<CPT> n_l2m = ldv<CPT>(row(l2mp, jn, P) + ib);
      V<CPT> n_mu = ldv<CPT>(row(mup, jn, P) + ib);
      V<CPT> n_txx = ldv<CPT>(row(txx, jn, P) + ib);
      V<CPT> n_tzz = ldv<CPT>(row(tzz, jn, P) + ib);
      V<CPT> n_txz = ldv<CPT>(row(txz, jn, P) + ib);

#pragma unroll
      for (int c = 0; c < CPT; ++c) {
        s_vx[M + lane * CPT + c] = wvx[c][M - 1];
        s_vz[M + lane * CPT + c] = wvz[c][M];
      }
      if (has_halo) {
        s_vx[hx] = c_hvx;
        s_vz[hx] = c_hvz;
      }
      wave_sync();

      V<CPT> o_xx, o_zz, o_xz;
#pragma unroll
      for (int c = 0; c < CPT; ++c) {
        const int x = M + lane * CPT + c;
        float d1 = 0.f, d2 = 0.f, d3 = 0.f, d4 = 0.f;
#pragma unroll
        for (int l = 1; l <= M; ++l) {
          d1 = fmaf_(a[l - 1], s_vx[x + l] - s_vx[x - l + 1], d1);          // dvxdx
          d2 = fmaf_(a[l - 1], wvz[c][M + l - 1] - wvz[c][M - l], d2);      // dvzdz
          d3 = fmaf_(a[l - 1], wvx[c][M - 1 + l] - wvx[c][M - l], d3);      // dvxdz
          d4 = fmaf_(a[l - 1], s_vz[x + l - 1] - s_vz[x - l], d4);          // dvzdx
        }
        const int i = ib + c;
        const bool inside = (i >= M) & (i < nx - M) & (j >= M) & (j < nz - M);
        float xx = c_txx.v[c], zz = c_tzz.v[c];
        if (src_on && i == si && j == sj) {
          xx = xx + wav;
          zz = zz + wav;
        }
        const float t1 = d1 * dtx, t2 = d2 * dtz;
        const float lv = c_lam.v[c], l2 = c_l2m.v[c];
        const float ixx = fmaf_(lv, t2, l2 * t1);
        const float izz = fmaf_(l2, t2, lv * t1);
        const float sxz = fmaf_(d4, dtx, d3 * dtz);
        const float ixz = c_mu.v[c] * sxz;
        o_xx.v[c] = inside ? xx + ixx : xx;
        o_zz.v[c] = inside ? zz + izz : zz;
        o_xz.v[c] = inside ? c_txz.v[c] + ixz : c_txz.v[c];
      }
      stv<CPT>(row(txx, j, P) + ib, o_xx);
      stv<CPT>(row(tzz, j, P) + ib, o_zz);
      stv<CPT>(row(txz, j, P) + ib, o_xz);
      if (flags & FLAG_BACKUP) {
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
          if (in_strip(nx, nz, nbc, fs, ib + c, j)) {
            stg(row(otxx, j, P) + ib + c, c_txx.v[c]);
            stg(row(otzz, j, P) + ib + c, c_tzz.v[c]);
            stg(row(otxz, j, P) + ib + c, c_txz.v[c]);
          }
        }
      }
      wave_sync();

#pragma unroll
      for (int c = 0; c < CPT; ++c) {
#pragma unroll
        for (int q = 0; q < 2 * M - 1; ++q) {
          wvx[c][q] = wvx[c][q + 1];
          wvz[c][q] = wvz[c][q + 1];
        }
        wvx[c][2 * M - 1] = n_vx.v[c];
        wvz[c][2 * M - 1] = n_vz.v[c];
      }
      c_hvx = n_hvx;
      c_hvz = n_hvz;
      c_lam = n_lam;
      c_l2m = n_l2m;
      c_mu = n_mu;
      c_txx = n_txx;
      c_tzz = n_tzz;
      c_txz = n_txz;
    }
  }
}

// ---------------------------------------------------------------------------
// Fused step on ping-pong buffers: velocity (velocity.jl:21-46) and stress
// (stress.jl:18-47) of one time step in a single z-march, reading the previous
// state A (fo[]) and writing the new state B (f[]).  A is never written, so it
// also serves as the HABC strip backup (the backup of boundary.jl:45-56 is the
// state at the start of the step).
//
// A wavefront covers 64*CPT lane columns and writes the inner ftw of them
// (H = M rounded up to CPT lanes on each side are computed but not written):
// new velocities are needed M columns and M rows beyond the stress cells.
// Iteration jv computes the new velocity row jv and the new stress row
// j = jv - M.  Register windows hold old txz (2M rows), old tzz (2M+1), old
// txx (M+1) and the new vx (2M) / vz (2M+1); LDS carries the rows needed
// for x-derivatives.  Stress cells in the bands next to the HABC strips are
// left to k_band (they need HABC-corrected velocities).
// Arithmetic is that of k_velocity / k_stress exactly.
// ---------------------------------------------------------------------------
// SAME: dtx == dtz for every shot, so the dt/rho arrays of x and z are one
// array each (a compile-time fact: a run-time choice between the two makes
// the compiler copy the prefetched values right after issuing their loads).
//
// CMP (FP32 with dx == dz only): the medium comes as three arrays, rho_vx,
// lam and mu_txx (12 B per cell instead of the 20 B of dt/rho_vx, dt/rho_vz,
// lam, lam+2mu, mu_txz), and the kernel forms the rest per cell with the
// setup's formulas and IEEE operations: dtx/rho_vx; rho_vz =
// 0.25*(((r(i,j) + r(i+1,j)) + r(i,j+1)) + r(i+1,j+1)) (init.jl:146-149);
// lam + 2mu; mu_txz = 2/(1/mu(i,j) + 1/mu(i,j+1)) (init.jl:151-155), with
// 1/mu of row j+1 carried to the next row.  The model uses it only when its
// rho_vz and mu_txz are exactly those averages, so results are unchanged.
// The column right of a lane's last column comes from the next lane (DPP);
// the tile's last column gets 0 there, and its vz is neither stored nor
// read by a stored stress.
//
// Register windows are rings of L = 2M+2 rows: row r lives in slot
// (r - jb) mod L, and the z-march is unrolled by L so that every slot index
// is a compile-time constant.  Rotating the windows is then register renaming
// (no copies), and each prefetched row is first waited for one row later, at
// its first use.  The march runs in whole blocks of L rows; rows past the
// tile's last needed row are computed from clamped (zero guard) rows and not
// stored.
// Rows the fused march loads ahead of their first use, and its ring length /
// unroll (host and device: tile heights are chosen in whole march blocks).
#ifndef FUSED_PD
#define FUSED_PD 1
#endif
constexpr int fused_block(int M) { return 2 * M + 2 * FUSED_PD; }

// Tile t of a fused launch: first written column o0, end of the written
// columns xend, rows [j0, j1).  Whole grid: ftiles_x x ftiles_z tiles of
// ftw x tz.  FLAG_INT: the interior rectangle in ftw x tzi tiles.  FLAG_EDGE:
// the top and bottom blocks across the grid, then the left and right blocks
// beside the interior, in ftw x tze tiles (see ShotDev).
__device__ __forceinline__ void fused_tile(const ShotDev& sd, int t, int flags, int& o0, int& xend,
                                           int& j0, int& j1) {
  const int ftw = uniform(sd.ftw), nx = uniform(sd.nx), nz = uniform(sd.nz);
  if (flags & FLAG_INT) {
    const int nix = uniform(sd.nix), tzi = uniform(sd.tzi);
    o0 = uniform(sd.xi0) + (t % nix) * ftw;
    xend = uniform(sd.xi1);
    j0 = uniform(sd.zi0) + (t / nix) * tzi;
    j1 = min(j0 + tzi, uniform(sd.zi1));
  } else if (flags & FLAG_EDGE) {
    const int tze = uniform(sd.tze), ftx = uniform(sd.ftiles_x);
    const int e_top = uniform(sd.e_top), e_bot = uniform(sd.e_bot);
    const int zi0 = uniform(sd.zi0), zi1 = uniform(sd.zi1);
    if (t < e_top) {
      o0 = (t % ftx) * ftw;
      xend = nx;
      j0 = (t / ftx) * tze;
      j1 = min(j0 + tze, zi0);
    } else if (t < e_top + e_bot) {
      const int u = t - e_top;
      o0 = (u % ftx) * ftw;
      xend = nx;
      j0 = zi1 + (u / ftx) * tze;
      j1 = min(j0 + tze, nz);
    } else {
      const int u = t - e_top - e_bot;
      const int ntl = uniform(sd.ntx_l), ntr = uniform(sd.ntx_r), e_mid = uniform(sd.e_mid);
      int row;
      if (u < ntl * e_mid) {
        o0 = (u % ntl) * ftw;
        xend = uniform(sd.xi0);
        row = u / ntl;
      } else {
        const int v = u - ntl * e_mid;
        o0 = uniform(sd.xi1) + (v % ntr) * ftw;
        xend = nx;
        row = v / ntr;
      }
      j0 = zi0 + row * tze;
      j1 = min(j0 + tze, zi1);
    }
  } else {
    const int ftiles_x = uniform(sd.ftiles_x), tz = uniform(sd.tz);
    o0 = (t % ftiles_x) * ftw;
    xend = nx;
    j0 = (t / ftiles_x) * tz;
    j1 = min(j0 + tz, nz);
  }
}

template <int M, int CPT, int WPB, int K, bool SAME, bool CMP>
__global__ __launch_bounds__(64 * WPB) void k_fused(BatchArgs ba, int flags) {
  static_assert(!CMP || (Mode<K>::C == ST_F32 && SAME), "derived coefficients: FP32, dtx == dtz");
  constexpr int KF = Mode<K>::F, KC = Mode<K>::C;
  typedef typename Store<KF>::T T;
  typedef typename Store<KC>::T TC;
  constexpr int HL = ((M + CPT - 1) / CPT) * CPT;  // computed-not-written lane columns per side
  constexpr int LC = 64 * CPT;
  constexpr int LW = LC + 2 * M;
  constexpr int PD = FUSED_PD;
  constexpr int L = fused_block(M);
  __shared__ float lds[WPB][4][LW];
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  float* s_txx = lds[wid][0];
  float* s_txz = lds[wid][1];
  float* s_vx = lds[wid][2];
  float* s_vz = lds[wid][3];
  const int nsh = *ba.n_shots_p;
  const int total = ba.prefix[nsh];
  for (int item = uniform(xcd_block() * WPB + wid); item < total;
       item = uniform(item + gridDim.x * WPB)) {
    const int s = uniform(find_shot(ba.prefix, nsh, item));
    const ShotDev& sd = ba.shots[s];
    const int k = shot_step(sd, ba.k_off);
    if (k == 0) continue;
    const int t = item - ba.prefix[s];
    const int nx = uniform(sd.nx), nz = uniform(sd.nz), P = uniform(sd.pitch);
    const int fs = uniform(sd.fs), band = uniform(sd.band);
    int o0, xend, j0, j1;  // first written column, end of the written columns, rows
    fused_tile(sd, t, flags, o0, xend, j0, j1);
    const int ib0 = o0 - HL;                  // first lane column of the tile
    const int lo = lane * CPT;                // this lane's column offset in the tile
    const int ib = ib0 + lo;                  // this lane's first column
    int hx = -1;
    if (lane < M) hx = lane;
    else if (lane >= 64 - M) hx = LC + M + (lane - (64 - M));
    const bool has_halo = hx >= 0;
    const int hcol = ib0 - M + hx;
    constexpr int FES = (int)sizeof(T), CES = (int)sizeof(TC);
    const Rsrc rA = make_rsrc(uniform_p(sd.bufA), (uint32_t)uniform((int)sd.recF));
    const Rsrc rB = make_rsrc(uniform_p(sd.bufB), (uint32_t)uniform((int)sd.recF));
    // F24: the byte-plane images of the read and written ranges
    const uint64_t loc = uniform_u64(sd.loc);
    const Rsrc rAl = make_rsrc((const void*)(((uint64_t)uniform_p(sd.bufA) >> 1) + loc),
                               (uint32_t)uniform((int)sd.recF) >> 1);
    const Rsrc rBl = make_rsrc((const void*)(((uint64_t)uniform_p(sd.bufB) >> 1) + loc),
                               (uint32_t)uniform((int)sd.recF) >> 1);
    const Rsrc rC = make_rsrc(uniform_p(sd.bufC), (uint32_t)uniform((int)sd.recC));
    const uint32_t fstr = (uint32_t)uniform((int)sd.fstr_b), forg = (uint32_t)uniform((int)sd.forg_b);
    const uint32_t cstr = (uint32_t)uniform((int)sd.cstr_b), corg = (uint32_t)uniform((int)sd.corg_b);
    // coefficient blocks (dtz arrays alias the dtx ones when SAME)
    constexpr int QB1 = 0, QB3 = 1, QLAM = 2, QL2M = 3, QMU = 4, QB2 = SAME ? 0 : 5, QB4 = SAME ? 1 : 6;
    constexpr int QRHO = 5, QMUX = 6;  // CMP: rho_vx and mu_txx blocks (after the five of SAME)
    const int nw = uniform(sd.nw), stype = uniform(sd.stype);
    const bool src_on = (flags & FLAG_INJECT) && k <= nw;
    const float wav = src_on ? uniform_f(uniform_p(sd.wavelet)[k - 1]) : 0.f;
    const bool src_p = src_on && stype == SRC_PRESSURE;
    const bool src_vz = src_on && stype == SRC_FORCE_Z;
    const bool src_vx = src_on && stype == SRC_FORCE_X;
    const int si = uniform(sd.si), sj = uniform(sd.sj);
    const float src_scale = uniform_f(sd.src_scale);
    const float dtx = uniform_f(sd.dtx), dtz = uniform_f(sd.dtz);
    // load (r*) and store (w*) scales of reduced-precision storage (unused for FP32)
    const auto rvx = bcodec<KF>(sd.fisc[F_VX], rAl), rvz = bcodec<KF>(sd.fisc[F_VZ], rAl);
    const auto rxx = bcodec<KF>(sd.fisc[F_TXX], rAl), rzz = bcodec<KF>(sd.fisc[F_TZZ], rAl);
    const auto rxz = bcodec<KF>(sd.fisc[F_TXZ], rAl);
    const auto wvxs = bcodec<KF>(sd.fsc[F_VX], rBl), wvzs = bcodec<KF>(sd.fsc[F_VZ], rBl);
    const auto wxx = bcodec<KF>(sd.fsc[F_TXX], rBl), wzz = bcodec<KF>(sd.fsc[F_TZZ], rBl);
    const auto wxz = bcodec<KF>(sd.fsc[F_TXZ], rBl);
    const float r1 = uniform_f(sd.csc[C_BVX_X]), r2 = uniform_f(sd.csc[C_BVX_Z]);
    const float r3 = uniform_f(sd.csc[C_BVZ_X]), r4 = uniform_f(sd.csc[C_BVZ_Z]);
    const float rl = uniform_f(sd.csc[C_LAM]), rl2 = uniform_f(sd.csc[C_L2M]);
    const float rmu = uniform_f(sd.csc[C_MU]);
    float a[M];
#pragma unroll
    for (int l = 0; l < M; ++l) a[l] = uniform_f(sd.a[l]);
    // ---- end of shot-table reads
    const int zlast = nz + GUARD_ROWS - 1;  // last row a load may touch (zero guard)
    // byte offsets: uniform part (block + row + tile column ib0), lane part
    const uint32_t lvo = (uint32_t)(lo * FES), lco = (uint32_t)(lo * CES);
    // Loads: the written columns [o0, min(o0 + ftw, xend)) need velocity
    // within M of them and that velocity the previous state within M more,
    // so lanes (and halo lanes) wholly right of the written columns + 2M load
    // nothing (the last tile of a side strip or of a row reads only what it
    // uses); their values are never written nor read by a written column.
    const int cend = min(o0 + uniform(sd.ftw), xend) + 2 * M;
    const uint32_t lvl = ib < cend ? lvo : OOB, lcl = ib < cend ? lco : OOB;
    const uint32_t hvo = has_halo && hcol < cend ? (uint32_t)(hx * FES) : OOB;  // column ib0-M+hx
    auto frow = [&](int c, int r, int col) {
      return (uint32_t)c * fstr + forg + (uint32_t)((min(r, zlast) * P + col) * FES);
    };
    auto fld = [&](int c, int r, auto sc) { return bload<KF, CPT>(rA, lvl, frow(c, r, ib0), sc); };
    auto cof = [&](int q, int r, float sc) {
      return bload<KC, CPT>(rC, lcl, (uint32_t)q * cstr + corg + (uint32_t)((min(r, zlast) * P + ib0) * CES),
                            sc);
    };
    auto hal = [&](int c, int r, auto sc) { return bload<KF, 1>(rA, hvo, frow(c, r, ib0 - M), sc).v[0]; };
    // pressure source as seen by every reader of old txx / tzz at row r; it is
    // added where a row is first used (one row after its load is issued), so
    // that no prefetched value is touched before it is needed
    auto add_src = [&](V<CPT>& x, int r) {
      if (src_p && r == sj) {
#pragma unroll
        for (int c = 0; c < CPT; ++c)
          if (ib + c == si) x.v[c] = x.v[c] + wav;
      }
    };

    V<CPT> rtxx[L], rtxz[L], rtzz[L];  // old stress rings
    V<CPT> qvx[L], qvz[L];             // new velocity rings
    // per-iteration operands, slot = iteration index mod L (the iteration of
    // velocity row jv uses slot (jv - jb) mod L; stress coefficients there are
    // those of row jv - M)
    V<CPT> cvx[L], cvz[L], cb1[L], cb2[L], cb3[L], cb4[L], clam[L], cl2m[L], cmu[L];
    V<CPT> crho[L], cmux[L], imu;  // CMP: rho rows (velocity), mu rows (stress), 1/mu of row j
    float htxx[L], htxz[L];
    const int jb = j0 - M;  // first velocity row; row r sits in ring slot (r - jb) mod L
    // rows the march reads before its first row (PD rows of lookahead):
    // txx jb-M..jb+PD-1, txz jb-M..jb+M+PD-2, tzz jb-M..jb+M+PD-1; the source is
    // added here to txx rows below jb and tzz rows below jb+M, the later rows
    // get it at their first use in the march
#pragma unroll
    for (int q = 0; q < 2 * M + PD; ++q) {
      const int r = jb - M + q, sl = (q - M + L) % L;
      if (q < 2 * M + PD - 1) rtxz[sl] = fld(F_TXZ, r, rxz);
      rtzz[sl] = fld(F_TZZ, r, rzz);
      if (q < 2 * M) add_src(rtzz[sl], r);
      if (q < M + PD) {
        rtxx[sl] = fld(F_TXX, r, rxx);
        if (q < M) add_src(rtxx[sl], r);
      }
    }
#pragma unroll
    for (int k = 0; k < PD; ++k) {
      htxx[k] = hal(F_TXX, jb + k, rxx);
      htxz[k] = hal(F_TXZ, jb + k, rxz);
      cvx[k] = fld(F_VX, jb + k, rvx);
      cvz[k] = fld(F_VZ, jb + k, rvz);
      if constexpr (!CMP) {
        cb1[k] = cof(QB1, jb + k, r1);
        cb3[k] = cof(QB3, jb + k, r3);
        if constexpr (!SAME) {
          cb2[k] = cof(QB2, jb + k, r2);
          cb4[k] = cof(QB4, jb + k, r4);
        }
        cl2m[k] = cof(QL2M, jb + k - M, rl2);
        cmu[k] = cof(QMU, jb + k - M, rmu);
      }
      clam[k] = cof(QLAM, jb + k - M, rl);
    }
    if constexpr (CMP) {
      // rho rows sit in ring slot (r - jb) mod L, mu rows likewise
#pragma unroll
      for (int k = 0; k <= PD; ++k) {
        crho[k] = cof(QRHO, jb + k, 1.f);
        cmux[(k - M + L) % L] = cof(QMUX, jb - M + k, 1.f);
      }
#pragma unroll
      for (int c = 0; c < CPT; ++c) imu.v[c] = 1.0f / cmux[(L - M) % L].v[c];
    }
    const bool own_lane = lo >= HL && lo < LC - HL;
    // Two columns per lane: a lane's first column ib is even (tile widths and
    // HL are even), so with nx even no column mask separates a lane's pair and
    // each field row goes out as one paired store.  The stress mask is then
    // widened to even bounds [band_lo, band_hi): the at most one extra column
    // per side lies in the side bands, which the band kernel rewrites later
    // in the step (from the HABC'd velocities) before anything reads it.
    const bool pairs = CPT == 2 && (nx & 1) == 0;
    const int band_lo = band & ~1, band_hi = (nx - band + 1) & ~1;

    const int jend = j1 + M;
    for (int jv0 = jb; jv0 < jend; jv0 += L) {
#pragma unroll
      for (int u = 0; u < L; ++u) {
        const int jv = jv0 + u;
        const int cu = u % L, nu = (u + PD) % L;  // this iteration's slot, the prefetched one
        // ---- prefetch: velocity inputs of row jv+PD, stress coefficients of row jv+PD-M
        const int jn = jv + PD;
        htxx[nu] = hal(F_TXX, jn, rxx);
        htxz[nu] = hal(F_TXZ, jn, rxz);
        rtxx[(u + PD) % L] = fld(F_TXX, jn, rxx);
        rtxz[(u + M + PD - 1) % L] = fld(F_TXZ, jv + M + PD - 1, rxz);
        rtzz[(u + M + PD) % L] = fld(F_TZZ, jv + M + PD, rzz);
        cvx[nu] = fld(F_VX, jn, rvx);
        cvz[nu] = fld(F_VZ, jn, rvz);
        if constexpr (CMP) {
          crho[(u + 1 + PD) % L] = cof(QRHO, jv + 1 + PD, 1.f);
          cmux[(u - M + 1 + PD + L) % L] = cof(QMUX, jv - M + 1 + PD, 1.f);
        } else {
          cb1[nu] = cof(QB1, jn, r1);
          cb3[nu] = cof(QB3, jn, r3);
          if constexpr (!SAME) {
            cb2[nu] = cof(QB2, jn, r2);
            cb4[nu] = cof(QB4, jn, r4);
          }
          cl2m[nu] = cof(QL2M, jn - M, rl2);
          cmu[nu] = cof(QMU, jn - M, rmu);
        }
        clam[nu] = cof(QLAM, jn - M, rl);

        // ---- first uses of rows loaded PD iterations ago: txx row jv, tzz row jv+M
        add_src(rtxx[u % L], jv);
        add_src(rtzz[(u + M) % L], jv + M);
        // ---- new velocity, row jv, all lane columns
        const V<CPT>& xr = rtxx[u % L];
        const V<CPT>& zr = rtxz[u % L];
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
          s_txx[M + lo + c] = xr.v[c];
          s_txz[M + lo + c] = zr.v[c];
        }
        if (has_halo) {
          float hxx = htxx[cu];
          if (src_p && jv == sj && hcol == si) hxx = hxx + wav;
          s_txx[hx] = hxx;
          s_txz[hx] = htxz[cu];
        }
        wave_sync();
        V<CPT> nvx, nvz;
        float nb0 = 0.f, nb1 = 0.f;  // CMP: rho right of this lane's last column, rows jv, jv+1
        if constexpr (CMP) {
          nb0 = from_upper_lane(crho[u % L].v[0]);
          nb1 = from_upper_lane(crho[(u + 1) % L].v[0]);
        }
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
          const int x = M + lo + c;
          float d1 = 0.f, d2 = 0.f, d3 = 0.f, d4 = 0.f;
#pragma unroll
          for (int l = 1; l <= M; ++l) {
            d1 = fmaf_(a[l - 1], s_txx[x + l - 1] - s_txx[x - l], d1);
            d2 = fmaf_(a[l - 1], rtxz[(u + l - 1) % L].v[c] - rtxz[(u - l + L) % L].v[c], d2);
            d3 = fmaf_(a[l - 1], s_txz[x + l] - s_txz[x - l + 1], d3);
            d4 = fmaf_(a[l - 1], rtzz[(u + l) % L].v[c] - rtzz[(u - l + 1 + L) % L].v[c], d4);
          }
          const int i = ib + c;
          const bool inside = (i >= M) & (i < nx - M) & (jv >= M) & (jv < nz - M);
          float vxs = cvx[cu].v[c], vzs = cvz[cu].v[c];
          if (i == si && jv == sj) {
            if (src_vz) vzs = vzs + wav * src_scale;
            if (src_vx) vxs = vxs + wav * src_scale;
          }
          float b1, b3;
          if constexpr (CMP) {
            const float q0 = crho[u % L].v[c], q1 = crho[(u + 1) % L].v[c];
            const float q0n = c + 1 < CPT ? crho[u % L].v[min(c + 1, CPT - 1)] : nb0;
            const float q1n = c + 1 < CPT ? crho[(u + 1) % L].v[min(c + 1, CPT - 1)] : nb1;
            b1 = dtx / q0;
            b3 = dtx / (0.25f * (((q0 + q0n) + q1) + q1n));
          } else {
            b1 = cb1[cu].v[c];
            b3 = cb3[cu].v[c];
          }
          const float b2 = SAME ? b1 : cb2[cu].v[c], b4 = SAME ? b3 : cb4[cu].v[c];
          const float ix = fmaf_(b2, d2, b1 * d1);
          const float iz = fmaf_(b4, d4, b3 * d3);
          // the stress half below consumes the velocity as stored
          nvx.v[c] = as_stored<KF>(inside ? vxs + ix : vxs, wvxs, rvx);
          nvz.v[c] = as_stored<KF>(inside ? vzs + iz : vzs, wvzs, rvz);
        }
        {
          const bool vrow = jv >= j0 && jv < j1;
          if (CPT == 2 && pairs) {
            const uint32_t vo = (vrow && own_lane && ib < xend) ? lvo : OOB;
            bstore2<KF>(rB, vo, frow(F_VX, jv, ib0), nvx.v[0], nvx.v[CPT - 1], wvxs);
            bstore2<KF>(rB, vo, frow(F_VZ, jv, ib0), nvz.v[0], nvz.v[CPT - 1], wvzs);
          } else {
#pragma unroll
            for (int c = 0; c < CPT; ++c) {
              const uint32_t vo = (vrow && own_lane && ib + c < xend) ? lvo + (uint32_t)(c * FES) : OOB;
              bstore<KF>(rB, vo, frow(F_VX, jv, ib0), nvx.v[c], wvxs);
              bstore<KF>(rB, vo, frow(F_VZ, jv, ib0), nvz.v[c], wvzs);
            }
          }
        }
        qvx[u % L] = nvx;
        qvz[u % L] = nvz;

        // ---- new stress, row j = jv - M (stores outside the branch, masked,
        // so that every store is issued on every row)
        const int j = jv - M;
        const bool srow = j >= j0 && j < j1;
        V<CPT> dl2, dmt;  // CMP: lam + 2mu and mu_txz of row j
        if constexpr (CMP) {
#pragma unroll
          for (int c = 0; c < CPT; ++c) {
            const float m0 = cmux[(u - M + L) % L].v[c], m1 = cmux[(u - M + 1 + L) % L].v[c];
            const float i1 = 1.0f / m1;
            dmt.v[c] = 2.0f / (imu.v[c] + i1);
            dl2.v[c] = clam[cu].v[c] + 2.0f * m0;
            imu.v[c] = i1;
          }
        }
        V<CPT> oxx, ozz, oxz;
#pragma unroll
        for (int c = 0; c < CPT; ++c) oxx.v[c] = ozz.v[c] = oxz.v[c] = 0.f;
        if (srow) {
          const int sm = (u - M + L) % L;  // slot of row j
#pragma unroll
          for (int c = 0; c < CPT; ++c) {
            s_vx[lo + c] = qvx[sm].v[c];
            s_vz[lo + c] = qvz[sm].v[c];
          }
          wave_sync();
#pragma unroll
          for (int c = 0; c < CPT; ++c) {
            const int x = lo + c;
            float d1 = 0.f, d2 = 0.f, d3 = 0.f, d4 = 0.f;
            if (own_lane) {
#pragma unroll
              for (int l = 1; l <= M; ++l) {
                d1 = fmaf_(a[l - 1], s_vx[x + l] - s_vx[x - l + 1], d1);  // dvxdx
                d2 = fmaf_(a[l - 1], qvz[(u - M + l - 1 + L) % L].v[c] -
                                         qvz[(u - M - l + L) % L].v[c], d2);  // dvzdz
                d3 = fmaf_(a[l - 1], qvx[(u - M + l + L) % L].v[c] -
                                         qvx[(u - M - l + 1 + L) % L].v[c], d3);  // dvxdz
                d4 = fmaf_(a[l - 1], s_vz[x + l - 1] - s_vz[x - l], d4);  // dvzdx
              }
            }
            const int i = ib + c;
            const bool inside = (i >= M) & (i < nx - M) & (j >= M) & (j < nz - M);
            const float xx = rtxx[sm].v[c], zz = rtzz[sm].v[c], xz = rtxz[sm].v[c];
            const float t1 = d1 * dtx, t2 = d2 * dtz;
            const float lv = clam[cu].v[c], l2 = CMP ? dl2.v[c] : cl2m[cu].v[c];
            const float ixx = fmaf_(lv, t2, l2 * t1);
            const float izz = fmaf_(l2, t2, lv * t1);
            const float sxz = fmaf_(d4, dtx, d3 * dtz);
            const float ixz = (CMP ? dmt.v[c] : cmu[cu].v[c]) * sxz;
            oxx.v[c] = inside ? xx + ixx : xx;
            ozz.v[c] = inside ? zz + izz : zz;
            oxz.v[c] = inside ? xz + ixz : xz;
          }
        }
        {
          const bool band_row = j >= nz - band || (!fs && j < band);
          const bool wrow = srow && !band_row;
          if (CPT == 2 && pairs) {
            const uint32_t vo =
                (wrow && own_lane && ib >= band_lo && ib < band_hi && ib < xend) ? lvo : OOB;
            bstore2<KF>(rB, vo, frow(F_TXX, j, ib0), oxx.v[0], oxx.v[CPT - 1], wxx);
            bstore2<KF>(rB, vo, frow(F_TZZ, j, ib0), ozz.v[0], ozz.v[CPT - 1], wzz);
            bstore2<KF>(rB, vo, frow(F_TXZ, j, ib0), oxz.v[0], oxz.v[CPT - 1], wxz);
          } else {
#pragma unroll
            for (int c = 0; c < CPT; ++c) {
              const int i = ib + c;
              const uint32_t vo = (wrow && own_lane && i >= band && i < nx - band && i < xend)
                                      ? lvo + (uint32_t)(c * FES) : OOB;
              bstore<KF>(rB, vo, frow(F_TXX, j, ib0), oxx.v[c], wxx);
              bstore<KF>(rB, vo, frow(F_TZZ, j, ib0), ozz.v[c], wzz);
              bstore<KF>(rB, vo, frow(F_TXZ, j, ib0), oxz.v[c], wxz);
            }
          }
        }
        wave_sync();
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Two time steps in one z-march (the interior tiles of the double step,
// step mode 3): velocity and stress of step k from the state in A, then
// velocity and stress of step k + 1 from those step-k values, which stay in
// registers and LDS; only the step-(k+1) state is written (to B).  FP32
// storage with derived coefficients and dtx == dtz (k_fused<M, 2, WPB,
// ST_F32, true, true>), two columns per lane.
//
// Iteration jv computes velocity(k) of row jv, stress(k) of row jv - M,
// velocity(k+1) of row jv - 2M and stress(k+1) of row jv - 3M.  In x each
// half-step narrows the exact columns by M per side: velocity(k) is exact on
// all 128 lane columns (halo lanes stage the old stress M beyond them),
// stress(k) on [M, 127 - M], velocity(k+1) on [2M, 127 - 2M], stress(k+1) on
// [3M, 127 - 3M]; the written columns are the inner 128 - 2 dbl_halo(M).
// Lanes outside a half-step's exact range compute values that only other
// inexact lanes read.
//
// The arithmetic of each step is k_fused's, operation for operation, so two
// steps here equal two k_fused steps bit for bit:
//   - the step-k pressure source is added to old txx / tzz at their first use
//     (as k_fused); the step-(k+1) one to the step-k txx / tzz values of the
//     source cell as they enter the rings, which every later reader sees;
//   - the coefficients k_fused forms per cell (dtx/rho_vx, dtx/rho_vz,
//     lam + 2mu, mu_txz) are formed once, at step k, and kept in rings for
//     step k + 1 (same rows, so the same values).
// All register rings have L = fused_block(M) slots, row r in slot
// (r - jb) mod L; the longest live span (lam of rows jv - 3M .. jv + PD - M)
// is 2M + 1 + PD <= L.
// ---------------------------------------------------------------------------
constexpr int dbl_halo(int M) { return (3 * M + 1) / 2 * 2; }  // per side, two columns per lane
// Rows k_fused2 loads ahead of their first use, and its ring length / unroll.
#ifndef DBL_PD
#define DBL_PD 1
#endif
constexpr int dbl_block(int M) { return 2 * M + 2 * DBL_PD; }

template <bool B>
struct BoolTag {
  static constexpr bool value = B;
};

// CMP: coefficients formed from rho_vx, lam, mu_txx (k_fused's derived form);
// otherwise read from the five stored arrays (dt/rho_vx, dt/rho_vz, lam,
// lam + 2mu, mu_txz), the same values (the model holds both only when they
// agree bit for bit), with no divisions.
template <int M, int WPB, bool CMP>
__global__ __launch_bounds__(64 * WPB) void k_fused2(BatchArgs ba, int flags) {
  constexpr int CPT = 2, LC = 64 * CPT, LW = LC + 2 * M, PD = DBL_PD, L = dbl_block(M);
  constexpr int HL = dbl_halo(M);
  static_assert(2 * M + 1 + PD <= L, "ring too short");
  __shared__ float lds[WPB][4][LW];
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  float* s_txx = lds[wid][0];
  float* s_txz = lds[wid][1];
  float* s_vx = lds[wid][2];
  float* s_vz = lds[wid][3];
  const int nsh = *ba.n_shots_p;
  const int total = ba.prefix[nsh];
  for (int item = uniform(xcd_block() * WPB + wid); item < total;
       item = uniform(item + gridDim.x * WPB)) {
    const int s = uniform(find_shot(ba.prefix, nsh, item));
    const ShotDev& sd = ba.shots[s];
    const int k = shot_step(sd, ba.k_off);
    // the host issues double steps only where both steps are live
    if (k == 0 || shot_step(sd, ba.k_off + 1) == 0) continue;
    const int t = item - ba.prefix[s];
    const int P = uniform(sd.pitch);
    int o0, xend, j0, j1;
    fused_tile(sd, t, FLAG_INT, o0, xend, j0, j1);
    const int ib0 = o0 - HL;    // first lane column of the tile
    const int lo = lane * CPT;  // this lane's column offset in the tile
    const int ib = ib0 + lo;    // this lane's first column
    int hx = -1;  // LDS slot of this lane's halo column (velocity(k) x-derivatives)
    if (lane < M) hx = lane;
    else if (lane >= 64 - M) hx = LC + M + (lane - (64 - M));
    const bool has_halo = hx >= 0;
    const int hcol = ib0 - M + hx;
    const Rsrc rA = make_rsrc(uniform_p(sd.bufA), (uint32_t)uniform((int)sd.recF));
    const Rsrc rB = make_rsrc(uniform_p(sd.bufB), (uint32_t)uniform((int)sd.recF));
    const Rsrc rC = make_rsrc(uniform_p(sd.bufC), (uint32_t)uniform((int)sd.recC));
    const uint32_t fstr = (uint32_t)uniform((int)sd.fstr_b), forg = (uint32_t)uniform((int)sd.forg_b);
    const uint32_t cstr = (uint32_t)uniform((int)sd.cstr_b), corg = (uint32_t)uniform((int)sd.corg_b);
    constexpr int QB1 = 0, QB3 = 1, QLAM = 2, QL2M = 3, QMU = 4, QRHO = 5, QMUX = 6;
    const int nw = uniform(sd.nw), stype = uniform(sd.stype);
    const bool inj = (flags & FLAG_INJECT) != 0;
    const bool on1 = inj && k <= nw, on2 = inj && k + 1 <= nw;
    const float w1 = on1 ? uniform_f(uniform_p(sd.wavelet)[k - 1]) : 0.f;
    const float w2 = on2 ? uniform_f(uniform_p(sd.wavelet)[k]) : 0.f;
    const bool p1 = on1 && stype == SRC_PRESSURE, p2 = on2 && stype == SRC_PRESSURE;
    const bool fz1 = on1 && stype == SRC_FORCE_Z, fz2 = on2 && stype == SRC_FORCE_Z;
    const bool f1 = fz1 || (on1 && stype == SRC_FORCE_X), f2 = fz2 || (on2 && stype == SRC_FORCE_X);
    const int si = uniform(sd.si), sj = uniform(sd.sj);
    const float src_scale = uniform_f(sd.src_scale);
    const float dtx = uniform_f(sd.dtx);  // == dtz (derived coefficients need dtx == dtz)
    float a[M];
#pragma unroll
    for (int l = 0; l < M; ++l) a[l] = uniform_f(sd.a[l]);
    // ---- end of shot-table reads
    const int jb = j0 - 3 * M;  // first velocity(k) row
    const int jend = j1 + 3 * M;
    // The source cell can change this tile's values only inside the columns
    // its lanes and halo lanes hold and the rows it loads; other tiles run
    // the march without any source code.
    const bool src_tile = (on1 || on2) && si >= ib0 - M && si < ib0 + LC + M && sj >= jb - M - PD &&
                          sj < jend + L + M + PD;
    // byte offsets: per-tile constants + row * pitch (rows stay inside the
    // guard rows: the interior is at least nbc + 3M + 3 rows from an edge)
    const uint32_t P4 = (uint32_t)P * 4u;
    const uint32_t lvo = (uint32_t)(lo * 4);
    // Loads: the written columns need the previous state within 4M of them
    // (two steps of velocity and stress, M each), so lanes and halo lanes
    // wholly right of the written columns + 4M (in the last tile of a row)
    // load nothing; their values are never written nor read by a written column.
    const int cend = min(o0 + uniform(sd.ftw), xend) + 4 * M;
    const uint32_t lvl = ib < cend ? lvo : OOB;
    const uint32_t hvo = has_halo && hcol < cend ? (uint32_t)(hx * 4) : OOB;
    const uint32_t fb = forg + (uint32_t)ib0 * 4u, cb = corg + (uint32_t)ib0 * 4u;
    const uint32_t hb = forg + (uint32_t)(ib0 - M) * 4u;
    auto foff = [&](int c, int r) { return fb + (uint32_t)c * fstr + (uint32_t)r * P4; };
    auto coff = [&](int q, int r) { return cb + (uint32_t)q * cstr + (uint32_t)r * P4; };
    auto fld = [&](int c, int r) { return bload<ST_F32, CPT>(rA, lvl, foff(c, r), 1.f); };
    auto cof = [&](int q, int r) { return bload<ST_F32, CPT>(rC, lvl, coff(q, r), 1.f); };
    auto hal = [&](int c, int r) {
      return bload<ST_F32, 1>(rA, hvo, hb + (uint32_t)c * fstr + (uint32_t)r * P4, 1.f).v[0];
    };
    // the march's loads and stores: offset of row jv + these constants
    const uint32_t o_hxx = hb + F_TXX * fstr + PD * P4, o_hxz = hb + F_TXZ * fstr + PD * P4;
    const uint32_t o_txx = fb + F_TXX * fstr + PD * P4;
    const uint32_t o_txz = fb + F_TXZ * fstr + (uint32_t)(M + PD - 1) * P4;
    const uint32_t o_tzz = fb + F_TZZ * fstr + (uint32_t)(M + PD) * P4;
    const uint32_t o_vx = fb + F_VX * fstr + PD * P4, o_vz = fb + F_VZ * fstr + PD * P4;
    const uint32_t o_rho = cb + QRHO * cstr + (uint32_t)(1 + PD) * P4;
    const uint32_t o_mu = cb + QMUX * cstr - (uint32_t)(M - 1 - PD) * P4;
    const uint32_t o_lam = cb + QLAM * cstr - (uint32_t)(M - PD) * P4;
    const uint32_t o_b1 = cb + QB1 * cstr + PD * P4, o_b3 = cb + QB3 * cstr + PD * P4;
    const uint32_t o_l2m = cb + QL2M * cstr - (uint32_t)(M - PD) * P4;
    const uint32_t o_mt = cb + QMU * cstr - (uint32_t)(M - PD) * P4;
    const uint32_t w_vx = fb + F_VX * fstr - (uint32_t)(2 * M) * P4;
    const uint32_t w_vz = fb + F_VZ * fstr - (uint32_t)(2 * M) * P4;
    const uint32_t w_xx = fb + F_TXX * fstr - (uint32_t)(3 * M) * P4;
    const uint32_t w_zz = fb + F_TZZ * fstr - (uint32_t)(3 * M) * P4;
    const uint32_t w_xz = fb + F_TXZ * fstr - (uint32_t)(3 * M) * P4;
    const bool own_lane = lo >= HL && lo < LC - HL;

    auto march = [&](auto src_tag) {
      constexpr bool SRC = decltype(src_tag)::value;
      auto add_src1 = [&](V<CPT>& x, int r) {
        if (SRC && p1 && r == sj) {
#pragma unroll
          for (int c = 0; c < CPT; ++c)
            if (ib + c == si) x.v[c] = x.v[c] + w1;
        }
      };
      V<CPT> rtxx[L], rtxz[L], rtzz[L];  // old (step k-1) stress
      V<CPT> cvx[L], cvz[L];             // old velocity (slot of row jv)
      float htxx[L], htxz[L];            // old stress of the halo columns
      V<CPT> crho[L], cmux[L], clam[L];  // rho_vx rows, mu_txx rows, lam (slot jv: row jv - M)
      V<CPT> imu;                        // 1 / mu_txx of the step-k stress row
      V<CPT> qvx[L], qvz[L];             // step-k velocity
      V<CPT> sxx[L], szz[L], sxz[L];     // step-k stress (+ the step-(k+1) pressure source)
      V<CPT> wvx[L], wvz[L];             // step-(k+1) velocity
      V<CPT> b1r[L], b3r[L];             // dtx / rho_vx, dtx / rho_vz of row jv
      V<CPT> l2r[L], mtr[L];             // lam + 2 mu, mu_txz of row jv - M (slot jv)
      // (formed at step k with CMP, loaded with the stored form)
      // rows read before the first iteration (as k_fused, from jb)
#pragma unroll
      for (int q = 0; q < 2 * M + PD; ++q) {
        const int r = jb - M + q, sl = (q - M + L) % L;
        if (q < 2 * M + PD - 1) rtxz[sl] = fld(F_TXZ, r);
        rtzz[sl] = fld(F_TZZ, r);
        if (q < 2 * M) add_src1(rtzz[sl], r);
        if (q < M + PD) {
          rtxx[sl] = fld(F_TXX, r);
          if (q < M) add_src1(rtxx[sl], r);
        }
      }
#pragma unroll
      for (int q = 0; q < PD; ++q) {
        htxx[q] = hal(F_TXX, jb + q);
        htxz[q] = hal(F_TXZ, jb + q);
        cvx[q] = fld(F_VX, jb + q);
        cvz[q] = fld(F_VZ, jb + q);
        clam[q] = cof(QLAM, jb + q - M);
        if constexpr (!CMP) {
          b1r[q] = cof(QB1, jb + q);
          b3r[q] = cof(QB3, jb + q);
          l2r[q] = cof(QL2M, jb + q - M);
          mtr[q] = cof(QMU, jb + q - M);
        }
      }
      if constexpr (CMP) {
#pragma unroll
        for (int q = 0; q <= PD; ++q) {
          crho[q] = cof(QRHO, jb + q);
          cmux[(q - M + L) % L] = cof(QMUX, jb - M + q);
        }
#pragma unroll
        for (int c = 0; c < CPT; ++c) imu.v[c] = 1.0f / cmux[(L - M) % L].v[c];
      }

      for (int jv0 = jb; jv0 < jend; jv0 += L) {
#pragma unroll
        for (int u = 0; u < L; ++u) {
          const int jv = jv0 + u;
          const uint32_t rv = (uint32_t)jv * P4;
          const int cu = u % L, nu = (u + PD) % L;
          // ---- prefetch: the inputs of velocity(k) row jv+PD and stress(k) row jv+PD-M
          htxx[nu] = bload<ST_F32, 1>(rA, hvo, o_hxx + rv, 1.f).v[0];
          htxz[nu] = bload<ST_F32, 1>(rA, hvo, o_hxz + rv, 1.f).v[0];
          rtxx[(u + PD) % L] = bload<ST_F32, CPT>(rA, lvl, o_txx + rv, 1.f);
          rtxz[(u + M + PD - 1) % L] = bload<ST_F32, CPT>(rA, lvl, o_txz + rv, 1.f);
          rtzz[(u + M + PD) % L] = bload<ST_F32, CPT>(rA, lvl, o_tzz + rv, 1.f);
          cvx[nu] = bload<ST_F32, CPT>(rA, lvl, o_vx + rv, 1.f);
          cvz[nu] = bload<ST_F32, CPT>(rA, lvl, o_vz + rv, 1.f);
          if constexpr (CMP) {
            crho[(u + 1 + PD) % L] = bload<ST_F32, CPT>(rC, lvl, o_rho + rv, 1.f);
            cmux[(u - M + 1 + PD + L) % L] = bload<ST_F32, CPT>(rC, lvl, o_mu + rv, 1.f);
          } else {
            b1r[nu] = bload<ST_F32, CPT>(rC, lvl, o_b1 + rv, 1.f);
            b3r[nu] = bload<ST_F32, CPT>(rC, lvl, o_b3 + rv, 1.f);
            l2r[nu] = bload<ST_F32, CPT>(rC, lvl, o_l2m + rv, 1.f);
            mtr[nu] = bload<ST_F32, CPT>(rC, lvl, o_mt + rv, 1.f);
          }
          clam[nu] = bload<ST_F32, CPT>(rC, lvl, o_lam + rv, 1.f);
          add_src1(rtxx[u % L], jv);
          add_src1(rtzz[(u + M) % L], jv + M);

          // ---- velocity(k), row jv, all lane columns
          {
            const V<CPT>& xr = rtxx[u % L];
            const V<CPT>& zr = rtxz[u % L];
#pragma unroll
            for (int c = 0; c < CPT; ++c) {
              s_txx[M + lo + c] = xr.v[c];
              s_txz[M + lo + c] = zr.v[c];
            }
            if (has_halo) {
              float hxx = htxx[cu];
              if (SRC && p1 && jv == sj && hcol == si) hxx = hxx + w1;
              s_txx[hx] = hxx;
              s_txz[hx] = htxz[cu];
            }
            wave_sync();
            float nb0 = 0.f, nb1 = 0.f;  // CMP: rho right of this lane's last column, rows jv, jv+1
            if constexpr (CMP) {
              nb0 = from_upper_lane(crho[u % L].v[0]);
              nb1 = from_upper_lane(crho[(u + 1) % L].v[0]);
            }
#pragma unroll
            for (int c = 0; c < CPT; ++c) {
              const int x = M + lo + c;
              float d1 = 0.f, d2 = 0.f, d3 = 0.f, d4 = 0.f;
#pragma unroll
              for (int l = 1; l <= M; ++l) {
                d1 = fmaf_(a[l - 1], s_txx[x + l - 1] - s_txx[x - l], d1);
                d2 = fmaf_(a[l - 1], rtxz[(u + l - 1) % L].v[c] - rtxz[(u - l + L) % L].v[c], d2);
                d3 = fmaf_(a[l - 1], s_txz[x + l] - s_txz[x - l + 1], d3);
                d4 = fmaf_(a[l - 1], rtzz[(u + l) % L].v[c] - rtzz[(u - l + 1 + L) % L].v[c], d4);
              }
              float vxs = cvx[cu].v[c], vzs = cvz[cu].v[c];
              if (SRC && f1 && jv == sj && ib + c == si) {
                if (fz1) vzs = vzs + w1 * src_scale;
                else vxs = vxs + w1 * src_scale;
              }
              float b1, b3;
              if constexpr (CMP) {
                const float q0 = crho[u % L].v[c], q1 = crho[(u + 1) % L].v[c];
                const float q0n = c + 1 < CPT ? crho[u % L].v[min(c + 1, CPT - 1)] : nb0;
                const float q1n = c + 1 < CPT ? crho[(u + 1) % L].v[min(c + 1, CPT - 1)] : nb1;
                b1 = dtx / q0;
                b3 = dtx / (0.25f * (((q0 + q0n) + q1) + q1n));
                b1r[cu].v[c] = b1;
                b3r[cu].v[c] = b3;
              } else {
                b1 = b1r[cu].v[c];
                b3 = b3r[cu].v[c];
              }
              const float ix = fmaf_(b1, d2, b1 * d1);
              const float iz = fmaf_(b3, d4, b3 * d3);
              qvx[cu].v[c] = vxs + ix;
              qvz[cu].v[c] = vzs + iz;
            }
          }

          // ---- stress(k), row j = jv - M (coefficients on every row: imu is carried)
          {
            const int j = jv - M;
            const int sm = (u - M + L) % L;  // slot of row j
            if constexpr (CMP) {
#pragma unroll
              for (int c = 0; c < CPT; ++c) {
                const float m0 = cmux[(u - M + L) % L].v[c], m1 = cmux[(u - M + 1 + L) % L].v[c];
                const float i1 = 1.0f / m1;
                mtr[cu].v[c] = 2.0f / (imu.v[c] + i1);
                l2r[cu].v[c] = clam[cu].v[c] + 2.0f * m0;
                imu.v[c] = i1;
              }
            }
            if (j >= j0 - 2 * M && j < j1 + 2 * M) {  // rows velocity(k+1) reads
#pragma unroll
              for (int c = 0; c < CPT; ++c) {
                s_vx[M + lo + c] = qvx[sm].v[c];
                s_vz[M + lo + c] = qvz[sm].v[c];
              }
              wave_sync();
#pragma unroll
              for (int c = 0; c < CPT; ++c) {
                const int x = M + lo + c;
                float d1 = 0.f, d2 = 0.f, d3 = 0.f, d4 = 0.f;
#pragma unroll
                for (int l = 1; l <= M; ++l) {
                  d1 = fmaf_(a[l - 1], s_vx[x + l] - s_vx[x - l + 1], d1);
                  d2 = fmaf_(a[l - 1], qvz[(u - M + l - 1 + L) % L].v[c] - qvz[(u - M - l + L) % L].v[c], d2);
                  d3 = fmaf_(a[l - 1], qvx[(u - M + l + L) % L].v[c] - qvx[(u - M - l + 1 + L) % L].v[c], d3);
                  d4 = fmaf_(a[l - 1], s_vz[x + l - 1] - s_vz[x - l], d4);
                }
                const float xx = rtxx[sm].v[c], zz = rtzz[sm].v[c], xz = rtxz[sm].v[c];
                const float t1 = d1 * dtx, t2 = d2 * dtx;
                const float lv = clam[cu].v[c], l2 = l2r[cu].v[c];
                const float ixx = fmaf_(lv, t2, l2 * t1);
                const float izz = fmaf_(l2, t2, lv * t1);
                const float sxzv = fmaf_(d4, dtx, d3 * dtx);
                const float ixz = mtr[cu].v[c] * sxzv;
                float oxx = xx + ixx;
                float ozz = zz + izz;
                if (SRC && p2 && j == sj && ib + c == si) {  // step k+1 reads txx, tzz + its source
                  oxx = oxx + w2;
                  ozz = ozz + w2;
                }
                sxx[sm].v[c] = oxx;
                szz[sm].v[c] = ozz;
                sxz[sm].v[c] = xz + ixz;
              }
            }
          }

          // ---- velocity(k+1), row jw = jv - 2M
          {
            const int jw = jv - 2 * M;
            const int cw = (u - 2 * M + 2 * L) % L;  // slot of row jw
            V<CPT> ovx, ovz;
#pragma unroll
            for (int c = 0; c < CPT; ++c) ovx.v[c] = ovz.v[c] = 0.f;
            if (jw >= j0 - M && jw < j1 + M) {  // rows stress(k+1) reads
#pragma unroll
              for (int c = 0; c < CPT; ++c) {
                s_txx[M + lo + c] = sxx[cw].v[c];
                s_txz[M + lo + c] = sxz[cw].v[c];
              }
              wave_sync();
#pragma unroll
              for (int c = 0; c < CPT; ++c) {
                const int x = M + lo + c;
                float d1 = 0.f, d2 = 0.f, d3 = 0.f, d4 = 0.f;
#pragma unroll
                for (int l = 1; l <= M; ++l) {
                  d1 = fmaf_(a[l - 1], s_txx[x + l - 1] - s_txx[x - l], d1);
                  d2 = fmaf_(a[l - 1], sxz[(u - 2 * M + l - 1 + 2 * L) % L].v[c] -
                                           sxz[(u - 2 * M - l + 2 * L) % L].v[c], d2);
                  d3 = fmaf_(a[l - 1], s_txz[x + l] - s_txz[x - l + 1], d3);
                  d4 = fmaf_(a[l - 1], szz[(u - 2 * M + l + 2 * L) % L].v[c] -
                                           szz[(u - 2 * M - l + 1 + 2 * L) % L].v[c], d4);
                }
                float vxs = qvx[cw].v[c], vzs = qvz[cw].v[c];
                if (SRC && f2 && jw == sj && ib + c == si) {
                  if (fz2) vzs = vzs + w2 * src_scale;
                  else vxs = vxs + w2 * src_scale;
                }
                const float b1 = b1r[cw].v[c], b3 = b3r[cw].v[c];
                const float ix = fmaf_(b1, d2, b1 * d1);
                const float iz = fmaf_(b3, d4, b3 * d3);
                ovx.v[c] = vxs + ix;
                ovz.v[c] = vzs + iz;
              }
              wvx[cw] = ovx;
              wvz[cw] = ovz;
            }
            const uint32_t vo = (jw >= j0 && jw < j1 && own_lane && ib < xend) ? lvo : OOB;
            bstore2<ST_F32>(rB, vo, w_vx + rv, ovx.v[0], ovx.v[1], 1.f);
            bstore2<ST_F32>(rB, vo, w_vz + rv, ovz.v[0], ovz.v[1], 1.f);
          }

          // ---- stress(k+1), row j2 = jv - 3M
          {
            const int j2 = jv - 3 * M;
            const int c2 = (u - 3 * M + 4 * L) % L;  // slot of row j2
            const int cw = (u - 2 * M + 2 * L) % L;  // slot holding lam, lam + 2mu, mu_txz of row j2
            V<CPT> oxx, ozz, oxz;
#pragma unroll
            for (int c = 0; c < CPT; ++c) oxx.v[c] = ozz.v[c] = oxz.v[c] = 0.f;
            if (j2 >= j0 && j2 < j1) {
#pragma unroll
              for (int c = 0; c < CPT; ++c) {
                s_vx[M + lo + c] = wvx[c2].v[c];
                s_vz[M + lo + c] = wvz[c2].v[c];
              }
              wave_sync();
#pragma unroll
              for (int c = 0; c < CPT; ++c) {
                const int x = M + lo + c;
                float d1 = 0.f, d2 = 0.f, d3 = 0.f, d4 = 0.f;
#pragma unroll
                for (int l = 1; l <= M; ++l) {
                  d1 = fmaf_(a[l - 1], s_vx[x + l] - s_vx[x - l + 1], d1);
                  d2 = fmaf_(a[l - 1], wvz[(u - 3 * M + l - 1 + 4 * L) % L].v[c] -
                                           wvz[(u - 3 * M - l + 4 * L) % L].v[c], d2);
                  d3 = fmaf_(a[l - 1], wvx[(u - 3 * M + l + 4 * L) % L].v[c] -
                                           wvx[(u - 3 * M - l + 1 + 4 * L) % L].v[c], d3);
                  d4 = fmaf_(a[l - 1], s_vz[x + l - 1] - s_vz[x - l], d4);
                }
                const float xx = sxx[c2].v[c], zz = szz[c2].v[c], xz = sxz[c2].v[c];
                const float t1 = d1 * dtx, t2 = d2 * dtx;
                const float lv = clam[cw].v[c], l2 = l2r[cw].v[c];
                const float ixx = fmaf_(lv, t2, l2 * t1);
                const float izz = fmaf_(l2, t2, lv * t1);
                const float sxzv = fmaf_(d4, dtx, d3 * dtx);
                const float ixz = mtr[cw].v[c] * sxzv;
                oxx.v[c] = xx + ixx;
                ozz.v[c] = zz + izz;
                oxz.v[c] = xz + ixz;
              }
            }
            const uint32_t vo = (j2 >= j0 && j2 < j1 && own_lane && ib < xend) ? lvo : OOB;
            bstore2<ST_F32>(rB, vo, w_xx + rv, oxx.v[0], oxx.v[1], 1.f);
            bstore2<ST_F32>(rB, vo, w_zz + rv, ozz.v[0], ozz.v[1], 1.f);
            bstore2<ST_F32>(rB, vo, w_xz + rv, oxz.v[0], oxz.v[1], 1.f);
          }
        }
      }
    };
    if (src_tile) march(BoolTag<true>{});
    else march(BoolTag<false>{});
  }
}

// Stress in the bands next to the HABC strips, after the velocity HABC:
// new velocities from B (f[]), old stress from A (fo[]), output to B.  One
// wavefront per (band, 64-column chunk, band_rows-row segment); the same
// arithmetic as k_stress.
template <int M, int WPB, int K>
__global__ __launch_bounds__(64 * WPB) void k_band(BatchArgs ba, int flags) {
  constexpr int KF = Mode<K>::F, KC = Mode<K>::C;
  typedef typename Store<KF>::T T;
  typedef typename Store<KC>::T TC;
  constexpr int LW = 64 + 2 * M;
  __shared__ float lds[WPB][2][LW];
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  float* s_vx = lds[wid][0];
  float* s_vz = lds[wid][1];
  const int nsh = *ba.n_shots_p;
  const int total = ba.prefix[nsh];
  for (int item = uniform(blockIdx.x * WPB + wid); item < total;
       item = uniform(item + gridDim.x * WPB)) {
    const int s = uniform(find_shot(ba.prefix, nsh, item));
    const ShotDev& sd = ba.shots[s];
    const int k = shot_step(sd, ba.k_off);
    if (k == 0) continue;
    int t = item - ba.prefix[s];
    const int nx = uniform(sd.nx), nz = uniform(sd.nz), P = uniform(sd.pitch);
    const int band = uniform(sd.band), nb_side = uniform(sd.nb_side), nb_mid = uniform(sd.nb_mid);
    const int brows = uniform(sd.band_rows);
    // band rectangle [c0, c1) x [r0, r1) of this item
    int c0, c1, r0, r1;
    {
      const int side_chunks = (band + 63) / 64;
      const int mid_chunks = (nx - 2 * band + 63) / 64;
      if (t < 2 * nb_side) {
        const bool right = t >= nb_side;
        const int u = right ? t - nb_side : t;
        const int cc = u % side_chunks, rs = u / side_chunks;
        c0 = (right ? nx - band : 0) + cc * 64;
        c1 = min(c0 + 64, right ? nx : band);
        r0 = rs * brows;
        r1 = min(r0 + brows, nz);
      } else {
        t -= 2 * nb_side;
        const bool top = t >= nb_mid;
        const int u = top ? t - nb_mid : t;
        const int cc = u % mid_chunks, rs = u / mid_chunks;
        c0 = band + cc * 64;
        c1 = min(c0 + 64, nx - band);
        r0 = (top ? 0 : nz - band) + rs * brows;
        r1 = min(r0 + brows, top ? band : nz);
      }
    }
    const int ib = c0 + lane;
    int hx = -1;
    if (lane < M) hx = lane;
    else if (lane >= 64 - M) hx = 64 + M + (lane - (64 - M));
    const bool has_halo = hx >= 0;
    // Loads: the written columns [c0, c1) need velocity up to column c1+M-1
    // and their own old stress and coefficients; lanes right of that read
    // the last needed column again (no further lines), and their values are
    // never written nor read by a written column.
    const int ibv = min(ib, c1 + M - 1), ibs = min(ib, c1 - 1);
    const int hcol = min(c0 - M + hx, c1 + M - 1);
    const T* vx = (const T*)uniform_p(sd.f[F_VX]);
    const T* vz = (const T*)uniform_p(sd.f[F_VZ]);
    const T* atxx = (const T*)uniform_p(sd.fo[F_TXX]);
    const T* atzz = (const T*)uniform_p(sd.fo[F_TZZ]);
    const T* atxz = (const T*)uniform_p(sd.fo[F_TXZ]);
    T* btxx = (T*)uniform_p(sd.f[F_TXX]);
    T* btzz = (T*)uniform_p(sd.f[F_TZZ]);
    T* btxz = (T*)uniform_p(sd.f[F_TXZ]);
    const TC* lamp = (const TC*)uniform_p(sd.lam);
    const TC* l2mp = (const TC*)uniform_p(sd.l2m);
    const TC* mup = (const TC*)uniform_p(sd.mu_txz);
    const int nw = uniform(sd.nw), stype = uniform(sd.stype);
    const bool src_on = (flags & FLAG_INJECT) && k <= nw && stype == SRC_PRESSURE;
    const float wav = src_on ? uniform_f(uniform_p(sd.wavelet)[k - 1]) : 0.f;
    const int si = uniform(sd.si), sj = uniform(sd.sj);
    const float dtx = uniform_f(sd.dtx), dtz = uniform_f(sd.dtz);
    const auto rvx = pcodec<KF>(sd.fisc[F_VX], sd.loc), rvz = pcodec<KF>(sd.fisc[F_VZ], sd.loc);
    const auto rxx = pcodec<KF>(sd.fisc[F_TXX], sd.loc), rzz = pcodec<KF>(sd.fisc[F_TZZ], sd.loc);
    const auto rxz = pcodec<KF>(sd.fisc[F_TXZ], sd.loc);
    const auto wxx = pcodec<KF>(sd.fsc[F_TXX], sd.loc), wzz = pcodec<KF>(sd.fsc[F_TZZ], sd.loc);
    const auto wxz = pcodec<KF>(sd.fsc[F_TXZ], sd.loc);
    const float rl = uniform_f(sd.csc[C_LAM]), rl2 = uniform_f(sd.csc[C_L2M]);
    const float rmu = uniform_f(sd.csc[C_MU]);
    float a[M];
#pragma unroll
    for (int l = 0; l < M; ++l) a[l] = uniform_f(sd.a[l]);
    // ---- end of shot-table reads

    float wvx[2 * M];  // vx rows j-M+1 .. j+M
    float wvz[2 * M];  // vz rows j-M .. j+M-1
#pragma unroll
    for (int q = 0; q < 2 * M; ++q) {
      wvx[q] = ld1<KF>(rowt(vx, r0 - M + 1 + q, P) + ibv, rvx);
      wvz[q] = ld1<KF>(rowt(vz, r0 - M + q, P) + ibv, rvz);
    }
    // row r0's pointwise operands; each iteration prefetches the next row's
    float hvx = has_halo ? ld1<KF>(rowt(vx, r0, P) + hcol, rvx) : 0.f;
    float hvz = has_halo ? ld1<KF>(rowt(vz, r0, P) + hcol, rvz) : 0.f;
    float lv = ld1<KC>(rowt(lamp, r0, P) + ibs, rl), l2 = ld1<KC>(rowt(l2mp, r0, P) + ibs, rl2);
    float mu = ld1<KC>(rowt(mup, r0, P) + ibs, rmu);
    float oxx = ld1<KF>(rowt(atxx, r0, P) + ibs, rxx), ozz = ld1<KF>(rowt(atzz, r0, P) + ibs, rzz);
    float oxz = ld1<KF>(rowt(atxz, r0, P) + ibs, rxz);
    for (int j = r0; j < r1; ++j) {
      const int jn = j + 1;
      const float n_hvx = has_halo ? ld1<KF>(rowt(vx, jn, P) + hcol, rvx) : 0.f;
      const float n_hvz = has_halo ? ld1<KF>(rowt(vz, jn, P) + hcol, rvz) : 0.f;
      const float n_lv = ld1<KC>(rowt(lamp, jn, P) + ibs, rl), n_l2 = ld1<KC>(rowt(l2mp, jn, P) + ibs, rl2);
      const float n_mu = ld1<KC>(rowt(mup, jn, P) + ibs, rmu);
      const float n_oxx = ld1<KF>(rowt(atxx, jn, P) + ibs, rxx), n_ozz = ld1<KF>(rowt(atzz, jn, P) + ibs, rzz);
      const float n_oxz = ld1<KF>(rowt(atxz, jn, P) + ibs, rxz);
      const float n_wvx = ld1<KF>(rowt(vx, j + M + 1, P) + ibv, rvx);
      const float n_wvz = ld1<KF>(rowt(vz, j + M, P) + ibv, rvz);
      s_vx[M + lane] = wvx[M - 1];
      s_vz[M + lane] = wvz[M];
      if (has_halo) {
        s_vx[hx] = hvx;
        s_vz[hx] = hvz;
      }
      wave_sync();
      float d1 = 0.f, d2 = 0.f, d3 = 0.f, d4 = 0.f;
      const int x = M + lane;
#pragma unroll
      for (int l = 1; l <= M; ++l) {
        d1 = fmaf_(a[l - 1], s_vx[x + l] - s_vx[x - l + 1], d1);
        d2 = fmaf_(a[l - 1], wvz[M + l - 1] - wvz[M - l], d2);
        d3 = fmaf_(a[l - 1], wvx[M - 1 + l] - wvx[M - l], d3);
        d4 = fmaf_(a[l - 1], s_vz[x + l - 1] - s_vz[x - l], d4);
      }
      const int i = ib;
      const bool inside = (i >= M) & (i < nx - M) & (j >= M) & (j < nz - M);
      float xx = oxx, zz = ozz;
      if (src_on && i == si && j == sj) {
        xx = xx + wav;
        zz = zz + wav;
      }
      const float t1 = d1 * dtx, t2 = d2 * dtz;
      const float ixx = fmaf_(lv, t2, l2 * t1);
      const float izz = fmaf_(l2, t2, lv * t1);
      const float sxz = fmaf_(d4, dtx, d3 * dtz);
      if (i < c1) {
        st1<KF>(rowt(btxx, j, P) + i, inside ? xx + ixx : xx, wxx);
        st1<KF>(rowt(btzz, j, P) + i, inside ? zz + izz : zz, wzz);
        st1<KF>(rowt(btxz, j, P) + i, inside ? oxz + mu * sxz : oxz, wxz);
      }
      wave_sync();
#pragma unroll
      for (int q = 0; q < 2 * M - 1; ++q) {
        wvx[q] = wvx[q + 1];
        wvz[q] = wvz[q + 1];
      }
      wvx[2 * M - 1] = n_wvx;
      wvz[2 * M - 1] = n_wvz;
      hvx = n_hvx;
      hvz = n_hvz;
      lv = n_lv;
      l2 = n_l2;
      mu = n_mu;
      oxx = n_oxx;
      ozz = n_ozz;
      oxz = n_oxz;
    }
  }
}

// ---------------------------------------------------------------------------
// HABC (kernels/boundary.jl:68-132) in the literal program order of the CPU
// loops: left, right, bottom, [top], left-bottom, right-bottom, [left-top,
// right-top]; j outer, i inner; every read sees the current array.  That
// order makes the left and top edges Jacobi sweeps over pre-HABC values, the
// right edge a recurrence along +x per row, the bottom edge a recurrence
// along +z per column, the bottom-right corner a 2D recurrence, and so on.
//
// One launch per pass covers every field and every strip piece.  A piece owns
// a rectangle of cells and reads only cells that no other piece writes, so
// the pieces are independent:
//   L, R   64-row chunks of the side edges, excluding the rows next to the
//          corners;
//   B, T   64-column chunks of the bottom / top edges, excluding the columns
//          next to the corners;
//   BL, BR the bottom corner plus the edge row above it and the edge column
//          beside it (the values the corner recurrence consumes);
//   TL, TR likewise at the top (no free surface only).
// Each piece stages its rectangle (field and backup) in LDS, runs the
// program-order recurrences there and writes the rectangle back.
// ---------------------------------------------------------------------------
struct Hq {
  float qx, qz, qt_x, qt_z, qxt;
  __device__ float sumx(float fn, float fo, float fon) const {
    float s = -qx * fn;
    s = fmaf_(-qt_x, fo, s);
    return fmaf_(-qxt, fon, s);
  }
  __device__ float sumz(float fn, float fo, float fon) const {
    float s = -qz * fn;
    s = fmaf_(-qt_z, fo, s);
    return fmaf_(-qxt, fon, s);
  }
  __device__ float edge(float w, float f, float s) const { return fmaf_(1.0f - w, s, w * f); }
  __device__ float corner(float w, float f, float sx, float sz) const {
    return fmaf_((1.0f - w) * 0.5f, sx + sz, w * f);
  }
};

__device__ __forceinline__ float habc_w(const ShotDev& sd, const float* ramp, int i, int j) {
  int d = min(min(i, sd.nx - 1 - i), min(j, sd.nz - 1 - j));
  return ramp[min(d, sd.nbc + 1)];
}
// Same, with the grid scalars in registers and the ramp in LDS.
struct HabcGeo {
  int nx, nz, nbc;
  const float* w;  // LDS copy of the field's ramp
  __device__ __forceinline__ float operator()(int i, int j) const {
    const int d = min(min(i, nx - 1 - i), min(j, nz - 1 - j));
    return w[min(d, nbc + 1)];
  }
};

struct HabcField {
  void* f;
  const void* fo;
  const float* ramp;
  float rs, ws;  // load / store scale of the field
};

__device__ __forceinline__ HabcField habc_field(const ShotDev& sd, int stress, int fi) {
  HabcField h;
  const int c = stress ? F_TXX + fi : fi;
  h.f = sd.f[c];
  h.fo = sd.fo[c];
  h.ramp = sd.ramp + (stress ? RAMP_TAU : fi) * sd.ramp_len;
  h.rs = sd.fisc[c];
  h.ws = sd.fsc[c];
  return h;
}

constexpr int HS = 65;  // LDS row stride: odd, so row and column walks are bank-conflict free
// The staging arrays hold `rows` = (largest nbc of the launch) + 1 LDS rows
// of HS floats each (dynamic LDS, so small strips leave room for more blocks
// per CU).  Column 64 of the last row is never read: recurrence steps of
// lanes with no cell store there instead of branching.  The bottom-right
// sweep reads one group beyond its last step: up to float 132*nbc + 155 of the
// interleaved block, past the 130*(nbc+2) floats of a launch with rows =
// nbc+1 when nbc > 52, so the launch adds habc_lds_slack(rows) floats.
__host__ __device__ constexpr int habc_lds_slack(int rows) { return 2 * (rows - 1) - 104 > 0 ? 2 * (rows - 1) - 104 : 0; }
constexpr int HABC_WT_LEN = 64 + 2 * MAX_NBC + 12;  // weight table (s_w / s_rw in k_habc)

// rows [r0, r0+nr) x cols [c0, c0+nc) of the field g and its backup go -> s, so
// (s[r*HS + c], so[r*HS + c]), nc <= 64.  Loads of both arrays go out a chunk
// of rows at a time before any LDS store, so a chunk costs one memory latency:
// 32 rows for FP32 fields, a whole piece (64 rows) for 16-bit ones (measured
// fastest on the SEAM shot and the 64-shot batch in each case), 32 for F24's
// two loads per element (64 rows of them need more than 256 registers).
template <int K>
constexpr int stage_chunk() { return K != ST_F24 && sizeof(typename Store<K>::T) == 2 ? 64 : 32; }
// Recurrence loops read their operands (pre-HABC values, which no earlier
// step of the loop writes) this many steps ahead, so LDS latency stays off the
// chain of dependent steps.
constexpr int HPF = 4;
template <int K, class S>
__device__ __forceinline__ void lds_in2(float* s, float* so, const typename Store<K>::T* g,
                                        const typename Store<K>::T* go, S rs, int P, int r0,
                                        int nr, int c0, int nc, int lane) {
  constexpr int STAGE_CHUNK = stage_chunk<K>();
  const bool on = lane < nc;
  const int cl = min(lane, nc - 1);  // lanes past the block load a valid cell and store nothing
  for (int rb = 0; rb < nr; rb += STAGE_CHUNK) {
    uint32_t v[STAGE_CHUNK], w[STAGE_CHUNK];
#pragma unroll
    for (int r = 0; r < STAGE_CHUNK; ++r) {
      const ptrdiff_t o = (ptrdiff_t)(r0 + min(rb + r, nr - 1)) * P + c0 + cl;
      v[r] = ld_bits<K>(g + o, rs);
      w[r] = ld_bits<K>(go + o, rs);
    }
#pragma unroll
    for (int r = 0; r < STAGE_CHUNK; ++r)
      if (on && rb + r < nr) {
        s[(rb + r) * HS + lane] = from_stored<K>(v[r], rs);
        so[(rb + r) * HS + lane] = from_stored<K>(w[r], rs);
      }
  }
}
// s[(sr+r)*HS + sc + c] -> rows [r0, r0+nr) x cols [c0, c0+nc) of g   (nc <= 64)
template <int K, class S>
__device__ __forceinline__ void lds_out(const float* s, int sr, int sc, typename Store<K>::T* g,
                                        S ws, int P, int r0, int nr, int c0, int nc, int lane) {
  constexpr int STAGE_CHUNK = stage_chunk<K>();
  const bool on = lane < nc;
  for (int rb = 0; rb < nr; rb += STAGE_CHUNK) {
    float v[STAGE_CHUNK];
#pragma unroll
    for (int r = 0; r < STAGE_CHUNK; ++r)
      v[r] = (on && rb + r < nr) ? s[(sr + rb + r) * HS + sc + lane] : 0.f;
#pragma unroll
    for (int r = 0; r < STAGE_CHUNK; ++r)
      if (on && rb + r < nr) st1<K>(g + (ptrdiff_t)(r0 + rb + r) * P + c0 + lane, v[r], ws);
  }
}

// Transposed staging for the side edges: rows [r0, r0+nr) x cols [c0, c0+nc)
// of g and its backup -> (s[c*HS + r], so[c*HS + r]), nr <= 64, nc <= 64, so a
// side piece takes nc (= nbc+1) LDS rows like every other piece.
template <int K, class S>
__device__ __forceinline__ void lds_in2_t(float* s, float* so, const typename Store<K>::T* g,
                                          const typename Store<K>::T* go, S rs, int P, int r0,
                                          int nr, int c0, int nc, int lane) {
  constexpr int STAGE_CHUNK = stage_chunk<K>();
  const bool on = lane < nc;
  const int cl = min(lane, nc - 1);
  for (int rb = 0; rb < nr; rb += STAGE_CHUNK) {
    uint32_t v[STAGE_CHUNK], w[STAGE_CHUNK];
#pragma unroll
    for (int r = 0; r < STAGE_CHUNK; ++r) {
      const ptrdiff_t o = (ptrdiff_t)(r0 + min(rb + r, nr - 1)) * P + c0 + cl;
      v[r] = ld_bits<K>(g + o, rs);
      w[r] = ld_bits<K>(go + o, rs);
    }
#pragma unroll
    for (int r = 0; r < STAGE_CHUNK; ++r)
      if (on && rb + r < nr) {
        s[lane * HS + rb + r] = from_stored<K>(v[r], rs);
        so[lane * HS + rb + r] = from_stored<K>(w[r], rs);
      }
  }
}
// s[(sc+c)*HS + r] -> rows [r0, r0+nr) x cols [c0, c0+nc) of g   (nr, nc <= 64)
template <int K, class S>
__device__ __forceinline__ void lds_out_t(const float* s, int sc, typename Store<K>::T* g, S ws,
                                          int P, int r0, int nr, int c0, int nc, int lane) {
  constexpr int STAGE_CHUNK = stage_chunk<K>();
  const bool on = lane < nc;
  for (int rb = 0; rb < nr; rb += STAGE_CHUNK) {
    float v[STAGE_CHUNK];
#pragma unroll
    for (int r = 0; r < STAGE_CHUNK; ++r)
      v[r] = (on && rb + r < nr) ? s[(sc + lane) * HS + rb + r] : 0.f;
#pragma unroll
    for (int r = 0; r < STAGE_CHUNK; ++r)
      if (on && rb + r < nr) st1<K>(g + (ptrdiff_t)(r0 + rb + r) * P + c0 + lane, v[r], ws);
  }
}

// Interleaved staging for the bottom-right corner: cell (r, c) of rows
// [r0, r0+nr) x cols [c0, c0+nc) -> s[2*((r+1)*HS + c + 1)] (field) and the
// float after it (backup).  Row -1 and column -1 are padding: only lanes whose
// result is discarded read them, and the first float is never read.  Every
// operand of a corner cell is then a fixed offset from one per-lane index.
template <int K, class S>
__device__ __forceinline__ void lds_in2_fo(float* s, const typename Store<K>::T* g,
                                           const typename Store<K>::T* go, S rs, int P, int r0,
                                           int nr, int c0, int nc, int lane) {
  constexpr int STAGE_CHUNK = stage_chunk<K>();
  const bool on = lane < nc;
  const int cl = min(lane, nc - 1);
  for (int rb = 0; rb < nr; rb += STAGE_CHUNK) {
    uint32_t v[STAGE_CHUNK], w[STAGE_CHUNK];
#pragma unroll
    for (int r = 0; r < STAGE_CHUNK; ++r) {
      const ptrdiff_t o = (ptrdiff_t)(r0 + min(rb + r, nr - 1)) * P + c0 + cl;
      v[r] = ld_bits<K>(g + o, rs);
      w[r] = ld_bits<K>(go + o, rs);
    }
#pragma unroll
    for (int r = 0; r < STAGE_CHUNK; ++r)
      if (on && rb + r < nr) {
        const int i = 2 * ((rb + r + 1) * HS + lane + 1);
        s[i] = from_stored<K>(v[r], rs);
        s[i + 1] = from_stored<K>(w[r], rs);
      }
  }
}
// fields of the interleaved block -> rows [r0, r0+nr) x cols [c0, c0+nc) of g
template <int K, class S>
__device__ __forceinline__ void lds_out_fo(const float* s, typename Store<K>::T* g, S ws, int P,
                                           int r0, int nr, int c0, int nc, int lane) {
  constexpr int STAGE_CHUNK = stage_chunk<K>();
  const bool on = lane < nc;
  for (int rb = 0; rb < nr; rb += STAGE_CHUNK) {
    float v[STAGE_CHUNK];
#pragma unroll
    for (int r = 0; r < STAGE_CHUNK; ++r)
      v[r] = (on && rb + r < nr) ? s[2 * ((rb + r + 1) * HS + lane + 1)] : 0.f;
#pragma unroll
    for (int r = 0; r < STAGE_CHUNK; ++r)
      if (on && rb + r < nr) st1<K>(g + (ptrdiff_t)(r0 + rb + r) * P + c0 + lane, v[r], ws);
  }
}

// Items per field of the HABC partition and the first side-edge row.
__device__ __forceinline__ int habc_side_r0(const ShotDev& sd) { return sd.fs ? sd.nbc : sd.nbc + 2; }

template <int K>
__global__ __launch_bounds__(64) void k_habc(BatchArgs ba, int stress, int rows) {
  constexpr int KF = Mode<K>::F;
  typedef typename Store<KF>::T T;
  extern __shared__ float s_dyn[];  // 2 * (rows + 1) * HS + habc_lds_slack(rows) floats
  float* const s_f = s_dyn;
  float* const s_o = s_dyn + rows * HS;
  const int hdummy = rows * HS - 1;
  // Weights of the current piece.  Most pieces: s_w[d] = ramp[d] by edge
  // distance d in [0, n+1].  The bottom-right corner instead: s_rw[64 + j] =
  // weight at corner distance j = max(row, column) of its block, j in [0,
  // n+1]; its sweep also reads [2, 64) and (65 + n, 64 + 2n + 12) and stores
  // into [0, 2*HPF), only on steps whose result is discarded.
  __shared__ float s_wt[HABC_WT_LEN];
  float* const s_w = s_wt;
  float* const s_rw = s_wt;
  const int lane = threadIdx.x;
  const int nsh = *ba.n_shots_p;
  const int total = ba.prefix[nsh];
  for (int item = uniform(blockIdx.x); item < total; item = uniform(item + gridDim.x)) {
    const int s = uniform(find_shot(ba.prefix, nsh, item));
    const ShotDev& sd = ba.shots[s];
    if (shot_step(sd, ba.k_off) == 0) continue;
    const int t = item - ba.prefix[s];
    const int ncorner = sd.fs ? 2 : 4;
    const int per = sd.n_l + sd.n_r + sd.n_b + sd.n_t + ncorner;
    const int fi = t / per;
    int p = t % per;
    // piece order: corners first (the longest recurrences start first), then
    // left/right edge chunks, then bottom/top edge chunks
    const bool is_corner = p < ncorner;
    const int pc = p;
    p -= ncorner;
    HabcField h = habc_field(sd, stress, fi);
    T* hf = (T*)uniform_p(h.f);
    const T* hfo = (const T*)uniform_p(h.fo);
    const auto hrs = pcodec<KF>(h.rs, sd.loc), hws = pcodec<KF>(h.ws, sd.loc);
    const Hq q{uniform_f(sd.qx), uniform_f(sd.qz), uniform_f(sd.qt_x), uniform_f(sd.qt_z),
               uniform_f(sd.qxt)};
    const int n = uniform(sd.nbc), nx = uniform(sd.nx), nz = uniform(sd.nz), P = uniform(sd.pitch);
    const int ra = habc_side_r0(sd), rb = nz - n - 2;  // side-edge rows [ra, rb)
    const int ca = n + 2, cb = nx - n - 2;             // bottom/top-edge columns [ca, cb)
    __syncthreads();  // LDS of the previous item is no longer read
    if (lane < n + 2) {
      if (is_corner && pc == 1) s_rw[64 + lane] = ldg(h.ramp + n + 1 - lane);
      else s_w[lane] = ldg(h.ramp + lane);
    }
    const HabcGeo W{nx, nz, n, s_w};

    if (!is_corner && p < sd.n_l + sd.n_r) {
      const bool right = p >= sd.n_l;
      const int j0 = ra + (right ? p - sd.n_l : p) * 64;
      const int nr = min(64, rb - j0);
      const int c0 = right ? nx - n - 2 : 1;  // staged columns [c0, c0+n+1), transposed
      lds_in2_t<KF>(s_f, s_o, hf, hfo, hrs, P, j0, nr, c0, n + 1, lane);
      __syncthreads();
      if (!right) {
        // left edge, i = 1..n: Jacobi on the row's pre-HABC values (rows independent)
#pragma unroll 4
        for (int r = 0; r < nr; ++r) {
          if (lane < n) {
            const float fv = s_f[lane * HS + r], fn = s_f[(lane + 1) * HS + r];
            const float ov = s_o[lane * HS + r], on = s_o[(lane + 1) * HS + r];
            const float out = q.edge(W(1 + lane, j0 + r), fv, q.sumx(fn, ov, on));
            s_f[lane * HS + r] = out;
          }
        }
      } else if (lane < nr) {
        // right edge, one row per lane, recurrence along +x from column nx-n-2
        // (column c of this lane's row at [c*HS + lane])
        struct Col {
          float* p;
          __device__ float& operator[](int c) const { return p[c * HS]; }
        };
        const Col fr{s_f + lane}, orow{s_o + lane};
        float prev = fr[0], op = orow[0];
        float pf[HPF], po[HPF], pw[HPF];
#pragma unroll
        for (int k = 0; k < HPF; ++k) {
          const int c = min(1 + k, n);
          pf[k] = fr[c];
          po[k] = orow[c];
          pw[k] = W(c0 + c, j0 + lane);
        }
        for (int cg = 1; cg <= n; cg += HPF) {
          float nf[HPF], no[HPF], nw[HPF];
#pragma unroll
          for (int k = 0; k < HPF; ++k) {  // the next group's pre-values
            const int c = min(cg + HPF + k, n);
            nf[k] = fr[c];
            no[k] = orow[c];
            nw[k] = W(c0 + c, j0 + lane);
          }
#pragma unroll
          for (int k = 0; k < HPF; ++k) {
            if (cg + k <= n) {
              const float out = q.edge(pw[k], pf[k], q.sumx(prev, po[k], op));
              fr[cg + k] = out;
              prev = out;
              op = po[k];
            }
          }
#pragma unroll
          for (int k = 0; k < HPF; ++k) {
            pf[k] = nf[k];
            po[k] = no[k];
            pw[k] = nw[k];
          }
        }
      }
      __syncthreads();
      // left: block columns 0..n-1 = columns 1..n; right: block columns 1..n
      lds_out_t<KF>(s_f, right ? 1 : 0, hf, hws, P, j0, nr, right ? c0 + 1 : 1, n, lane);
      continue;
    }
    p -= sd.n_l + sd.n_r;
    if (!is_corner) {
      const bool top = p >= sd.n_b;
      const int i0 = ca + (top ? p - sd.n_b : p) * 64;
      const int nc = min(64, cb - i0);
      const int r0 = top ? 1 : nz - n - 2;  // staged rows [r0, r0+n+1)
      lds_in2<KF>(s_f, s_o, hf, hfo, hrs, P, r0, n + 1, i0, nc, lane);
      __syncthreads();
      if (lane < nc) {
        const int i = i0 + lane;
        if (!top) {
          // bottom edge: recurrence along +z from row nz-n-2
          float prev = s_f[lane], op = s_o[lane];
          float pf[HPF], po[HPF], pw[HPF];
#pragma unroll
          for (int k = 0; k < HPF; ++k) {
            const int r = min(1 + k, n);
            pf[k] = s_f[r * HS + lane];
            po[k] = s_o[r * HS + lane];
            pw[k] = W(i, r0 + r);
          }
          for (int rg = 1; rg <= n; rg += HPF) {
            float nf[HPF], no[HPF], nw[HPF];
#pragma unroll
            for (int k = 0; k < HPF; ++k) {  // the next group's pre-values
              const int r = min(rg + HPF + k, n);
              nf[k] = s_f[r * HS + lane];
              no[k] = s_o[r * HS + lane];
              nw[k] = W(i, r0 + r);
            }
#pragma unroll
            for (int k = 0; k < HPF; ++k) {
              if (rg + k <= n) {
                const float out = q.edge(pw[k], pf[k], q.sumz(prev, po[k], op));
                s_f[(rg + k) * HS + lane] = out;
                prev = out;
                op = po[k];
              }
            }
#pragma unroll
            for (int k = 0; k < HPF; ++k) {
              pf[k] = nf[k];
              po[k] = no[k];
              pw[k] = nw[k];
            }
          }
        } else {
          // top edge (no free surface), rows 1..n: Jacobi in z, rows ascending
          for (int r = 0; r < n; ++r) {
            const float fv = s_f[r * HS + lane], fn = s_f[(r + 1) * HS + lane];
            const float ov = s_o[r * HS + lane], on = s_o[(r + 1) * HS + lane];
            s_f[r * HS + lane] = q.edge(W(i, r0 + r), fv, q.sumz(fn, ov, on));
          }
        }
      }
      __syncthreads();
      // bottom: block rows 1..n = rows nz-n-1..nz-2; top: block rows 0..n-1 = rows 1..n
      lds_out<KF>(s_f, top ? 0 : 1, 0, hf, hws, P, top ? r0 : r0 + 1, n, i0, nc, lane);
      continue;
    }
    // Corner pieces: an (n+1) x (n+1) block.
    const bool bottom = pc < 2;
    const bool leftc = (pc % 2) == 0;
    const int r0 = bottom ? nz - n - 2 : 1;
    const int c0 = leftc ? 1 : nx - n - 2;
    const bool br = bottom && !leftc;  // staged interleaved (lds_in2_fo)
    if (br)
      lds_in2_fo<KF>(s_dyn, hf, hfo, hrs, P, r0, n + 1, c0, n + 1, lane);
    else
      lds_in2<KF>(s_f, s_o, hf, hfo, hrs, P, r0, n + 1, c0, n + 1, lane);
    __syncthreads();
    const bool act = lane <= n;
    if (bottom && leftc) {
      // BL: block row 0 = left-edge row nz-n-2, block column n = bottom-edge
      // column n+1.  1) left edge on row 0 (Jacobi); 2) per row, the bottom
      // edge at column n+1 then the left-bottom corner (right neighbour read
      // before update; the column-(n+1) neighbour is the new bottom value).
      if (lane < n) {
        const float fv = s_f[lane], fn = s_f[lane + 1], ov = s_o[lane], on = s_o[lane + 1];
        s_f[lane] = q.edge(W(c0 + lane, r0), fv, q.sumx(fn, ov, on));
      }
      __syncthreads();
      float up = act ? s_f[lane] : 0.f, ou = act ? s_o[lane] : 0.f;
      const int la = act ? lane : 0;  // idle lanes read a valid cell and never store
      // edge distance of cell (1+la, r0+r) in this corner: min(1+la, n+1-r)
      const int wl = min(1 + la, n + 1);
      float pf[HPF], po[HPF], pw[HPF];
#pragma unroll
      for (int k = 0; k < HPF; ++k) {
        const int r = min(1 + k, n);
        pf[k] = s_f[r * HS + la];
        po[k] = s_o[r * HS + la];
        pw[k] = s_w[min(wl, n + 1 - r)];
      }
      for (int rg = 1; rg <= n; rg += HPF) {
        float nf[HPF], no[HPF], nw[HPF];
#pragma unroll
        for (int k = 0; k < HPF; ++k) {  // the next group's pre-values (rows not yet written)
          const int r = min(rg + HPF + k, n);
          nf[k] = s_f[r * HS + la];
          no[k] = s_o[r * HS + la];
          nw[k] = s_w[min(wl, n + 1 - r)];
        }
#pragma unroll
        for (int k = 0; k < HPF; ++k) {
          if (rg + k <= n) {
            const float fv = pf[k], ov = po[k], w = pw[k];
            float x = fv;
            if (lane == n) x = q.edge(w, fv, q.sumz(up, ov, ou));
            const float fr = from_upper_lane(x);
            const float orr = from_upper_lane(ov);
            const float out =
                (lane == n) ? x : q.corner(w, fv, q.sumx(fr, ov, orr), q.sumz(up, ov, ou));
            s_f[act ? (rg + k) * HS + lane : hdummy] = out;
            up = out;
            ou = ov;
          }
        }
#pragma unroll
        for (int k = 0; k < HPF; ++k) {
          pf[k] = nf[k];
          po[k] = no[k];
          pw[k] = nw[k];
        }
      }
    } else if (bottom) {
      // BR: one wavefront over the block; block row 0 is the right-edge row
      // nz-n-2 (recurrence along +x), block column 0 the bottom-edge column
      // nx-n-2 (recurrence along +z), the rest the right-bottom corner.
      // Lane r owns block row r and at step d updates cell (r, c = d - r):
      // one anti-diagonal per step.  Its left neighbour (r, c-1) is the
      // lane's own previous result and its upper neighbour (r-1, c) lane
      // r-1's (DPP).  Every operand of the lane's cells is at a fixed offset
      // from one per-lane LDS pointer that advances two floats per step, so a
      // step costs no address arithmetic.  Steps whose cell lies outside the
      // block (c < 0, c > n, idle lanes) compute a value nothing reads: their
      // own `last` is next read as a left neighbour only after an edge-column
      // step has replaced it by the x-term's exact 0, and by lane r+1 only on
      // its own out-of-block steps; their store goes to a dummy slot.
      const int rl = min(lane, n);              // idle lanes shadow row n
      const int off = act ? lane : (1 << 20);   // c = d - off: never in [0, n] when idle
      // cell (r, c) is at interleaved index 2*((r+1)*HS + c + 1) (its backup
      // at +1).  With pbi = 2*(rl*(HS-1) + 1), the step-d cell of this lane is
      // at pbi + 2d + 2*HS; its left neighbour's backup one float before, and
      // its upper neighbour's backup at pbi + 2d + 1.
      const int pbi = 2 * (rl * (HS - 1) + 1);
      const float* const pb = s_dyn + pbi;
      const float* const pw = s_rw + 64 - rl;   // pw[d]: weight at distance c, for c >= rl
      const float wl = s_rw[64 + rl];           // weight of the row's cells with c < rl
      const bool row0 = lane == 0;
      float last = s_dyn[2 * (HS + 1)];         // lane 0: cell (0,0), not updated, left of (0,1)
      struct Op {
        float fv, ov, ol, ou, w;
      };
      auto fetch = [&](int d) {
        const float* b = pb + 2 * d;
        Op o;
        o.fv = b[2 * HS];
        o.ov = b[2 * HS + 1];
        o.ol = b[2 * HS - 1];
        o.ou = b[1];
        o.w = d < 2 * rl ? wl : pw[d];
        return o;
      };
      // step d = dg + k of a group whose cells start at s_dyn[qs]: the store
      // goes to qs + 2k, or into s_rw[0..2*HPF) (read only by discarded steps)
      auto step = [&](int d, int k, const Op& o, float* qs) {
        const int c = d - off;
        const float up = from_lower_lane_z(last);
        float sx = q.sumx(last, o.ov, o.ol);
        float sz = q.sumz(up, o.ov, o.ou);
        // the bottom-edge column has no x-term, the right-edge row no z-term,
        // and their weight factor is (1-w) instead of (1-w)/2; adding an
        // exact 0 keeps each result identical to its own formula
        sx = (c == 0) ? 0.f : sx;
        sz = row0 ? 0.f : sz;
        const float om = 1.0f - o.w;
        const float cw = (c == 0 || row0) ? om : om * 0.5f;
        const float out = fmaf_(cw, sx + sz, o.w * o.fv);
        float* const dst = (unsigned)c <= (unsigned)n ? qs : s_rw;
        dst[2 * k] = out;
        last = out;
      };
      Op a[HPF], bq[HPF];
#pragma unroll
      for (int k = 0; k < HPF; ++k) a[k] = fetch(1 + k);
      // ping-pong over two groups of HPF steps, each group's operands read one
      // group ahead; steps past 2n fall outside the block for every lane
      for (int dg = 1; dg <= 2 * n; dg += 2 * HPF) {
        float* const qa = s_dyn + pbi + 2 * HS + 2 * dg;
#pragma unroll
        for (int k = 0; k < HPF; ++k) bq[k] = fetch(dg + HPF + k);
#pragma unroll
        for (int k = 0; k < HPF; ++k) step(dg + k, k, a[k], qa);
#pragma unroll
        for (int k = 0; k < HPF; ++k) a[k] = fetch(dg + 2 * HPF + k);
#pragma unroll
        for (int k = 0; k < HPF; ++k) step(dg + HPF + k, k, bq[k], qa + 2 * HPF);
      }
    } else if (leftc) {
      // TL (no free surface): block rows 0..n = rows 1..n+1, columns 0..n =
      // columns 1..n+1.  1) left edge on every block row (Jacobi); 2) top edge
      // on block column n (Jacobi in z); 3) left-top corner (Jacobi over 1-2).
      for (int r = 0; r <= n; ++r) {
        if (lane < n) {
          const float fv = s_f[r * HS + lane], fn = s_f[r * HS + lane + 1];
          const float ov = s_o[r * HS + lane], on = s_o[r * HS + lane + 1];
          s_f[r * HS + lane] = q.edge(W(c0 + lane, r0 + r), fv, q.sumx(fn, ov, on));
        }
      }
      __syncthreads();
      if (lane < n) {
        const float fv = s_f[lane * HS + n], fn = s_f[(lane + 1) * HS + n];
        const float ov = s_o[lane * HS + n], on = s_o[(lane + 1) * HS + n];
        s_f[lane * HS + n] = q.edge(W(c0 + n, r0 + lane), fv, q.sumz(fn, ov, on));
      }
      __syncthreads();
      for (int r = 0; r < n; ++r) {
        if (lane < n) {
          const float fv = s_f[r * HS + lane], fr = s_f[r * HS + lane + 1];
          const float fd = s_f[(r + 1) * HS + lane];
          const float ov = s_o[r * HS + lane], orr = s_o[r * HS + lane + 1];
          const float od = s_o[(r + 1) * HS + lane];
          s_f[r * HS + lane] = q.corner(W(c0 + lane, r0 + r), fv,
                                        q.sumx(fr, ov, orr), q.sumz(fd, ov, od));
        }
      }
    } else {
      // TR (no free surface): block rows 0..n = rows 1..n+1, column 0 =
      // column nx-n-2.  1) right edge on every block row (+x recurrence);
      // 2) top edge on block column 0 (Jacobi in z); 3) right-top corner, +x
      // recurrence per row reading the row below before it is updated.
      if (act) {
        float* fr = s_f + lane * HS;
        const float* orow = s_o + lane * HS;
        float prev = fr[0], op = orow[0];
        for (int c = 1; c <= n; ++c) {
          const float fv = fr[c], ov = orow[c];
          const float out = q.edge(W(c0 + c, r0 + lane), fv, q.sumx(prev, ov, op));
          fr[c] = out;
          prev = out;
          op = ov;
        }
      }
      __syncthreads();
      if (lane < n) {
        const float fv = s_f[lane * HS], fn = s_f[(lane + 1) * HS];
        const float ov = s_o[lane * HS], on = s_o[(lane + 1) * HS];
        s_f[lane * HS] = q.edge(W(c0, r0 + lane), fv, q.sumz(fn, ov, on));
      }
      __syncthreads();
      const int rr = min(lane, n - 1);  // lanes >= n shadow row n-1 and do not write
      float prev = s_f[rr * HS], op = s_o[rr * HS];
      for (int c = 1; c <= n; ++c) {
        const float fv = s_f[rr * HS + c], ov = s_o[rr * HS + c];
        const float fd = s_f[(rr + 1) * HS + c], od = s_o[(rr + 1) * HS + c];
        const float out = q.corner(W(c0 + c, r0 + rr), fv, q.sumx(prev, ov, op),
                                   q.sumz(fd, ov, od));
        wave_sync();
        if (lane < n) s_f[rr * HS + c] = out;
        prev = out;
        op = ov;
      }
    }
    __syncthreads();
    if (br)
      lds_out_fo<KF>(s_dyn, hf, hws, P, r0, n + 1, c0, n + 1, lane);
    else
      lds_out<KF>(s_f, 0, 0, hf, hws, P, r0, n + 1, c0, n + 1, lane);
  }
}

// ---------------------------------------------------------------------------
// Free surface (kernels/boundary.jl:168-181) and receiver sampling
// (kernels/source_receiver.jl:65-77).  Items [0, 6*nx) zero tzz/txz on rows
// pad-5..pad; the rest sample receivers into gather[k-1][r].  A pressure
// receiver on a free-surface row reads tzz as the 0 it is being set to.
// ---------------------------------------------------------------------------
template <int WPB, int K>
__global__ __launch_bounds__(64 * WPB) void k_fsrec(BatchArgs ba, int mode) {
  constexpr int KF = Mode<K>::F;
  typedef typename Store<KF>::T T;
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int nsh = *ba.n_shots_p;
  const int total = ba.prefix[nsh];
  for (int item = uniform(blockIdx.x * WPB + wid); item < total;
       item = uniform(item + gridDim.x * WPB)) {
    const int s = uniform(find_shot(ba.prefix, nsh, item));
    const ShotDev& sd = ba.shots[s];
    const int k = shot_step(sd, ba.k_off);
    if (k == 0) continue;
    const int idx = (item - ba.prefix[s]) * 64 + lane;
    const int nx = sd.nx, P = sd.pitch;
    const bool fs_on = (mode & MODE_FS) && sd.fs;
    const int nfs = 6 * nx;
    if (idx < nfs) {
      if (!fs_on) continue;
      const int j = sd.pad - 5 + idx / nx, i = idx % nx;
      if (j >= 0 && j < sd.nz) {
        const auto zc = pcodec<KF>(1.f, sd.loc);  // zero is stored exactly at any scale
        st1<KF>(rowt((T*)sd.f[F_TZZ], j, P) + i, 0.f, zc);
        st1<KF>(rowt((T*)sd.f[F_TXZ], j, P) + i, 0.f, zc);
      }
      continue;
    }
    const int r = idx - nfs;
    if (!(mode & MODE_RECORD) || r >= sd.n_rec) continue;
    const int ri = sd.rec[3 * r], rj = sd.rec[3 * r + 1], rt = sd.rec[3 * r + 2];
    float v;
    auto fld = [&](int c) {
      return ld1<KF>(rowt((const T*)sd.f[c], rj, P) + ri, pcodec<KF>(sd.fisc[c], sd.loc));
    };
    if (rt == REC_VZ) v = fld(F_VZ);
    else if (rt == REC_VX) v = fld(F_VX);
    else {
      const bool zero_tzz = fs_on && rj >= sd.pad - 5 && rj <= sd.pad;
      const float zz = zero_tzz ? 0.f : fld(F_TZZ);
      v = (fld(F_TXX) + zz) * 0.5f;
    }
    sd.gather[(size_t)(k - 1) * sd.n_rec + r] = v;
  }
}

// ---------------------------------------------------------------------------
// Wavefield frames (visualization/video_recorder.jl:207-283): on every step k
// with k % skip == 0 the decimated fields go to frame slot
// (k/skip - 1) mod capacity of a device frame store (slot >= 0 in the
// arguments: that slot, whatever the step).  frame[j_d][i_d] = field(min(i_d*ds, nx-1), min(j_d*ds, nz-1)),
// 0-based (_extract_frame!, :260-268); p = (txx + tzz) * 0.5 and
// vel = sqrt(vx*vx + vz*vz) (_get_field_data, :270-283).  One shot (row 0).
// ---------------------------------------------------------------------------
struct FrameArgs {
  float* out;     // [capacity][nfields][nz_d][nx_d]
  int field[4];   // FRAME_* per stored field
  int nfields, skip, ds, nx_d, nz_d, capacity, slot;
  int own_sc;     // 1: decode with isc (the state's scales) instead of the shot row's
  float isc[5];
};
enum FrameField { FRAME_P = 0, FRAME_VEL = 1, FRAME_VX = 2, FRAME_VZ = 3 };

template <int K>
__global__ __launch_bounds__(256) void k_frame(BatchArgs ba, FrameArgs fa) {
  constexpr int KF = Mode<K>::F;
  typedef typename Store<KF>::T T;
  const ShotDev& sd = ba.shots[0];
  int slot = fa.slot;
  if (slot < 0) {
    const int k = shot_step(sd, ba.k_off);
    if (k == 0 || k % fa.skip != 0) return;
    slot = (k / fa.skip - 1) % fa.capacity;  // a ring of frames
  }
  const size_t per = (size_t)fa.nx_d * fa.nz_d;
  float* base = fa.out + (size_t)slot * fa.nfields * per;
  const int P = sd.pitch;
  for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < per * fa.nfields;
       o += (size_t)gridDim.x * blockDim.x) {
    const int q = (int)(o / per);
    const size_t r = o - (size_t)q * per;
    const int id = (int)(r % fa.nx_d), jd = (int)(r / fa.nx_d);
    const int i = min(id * fa.ds, sd.nx - 1), j = min(jd * fa.ds, sd.nz - 1);
    auto fld = [&](int c) {
      return ld1<KF>(rowt((const T*)sd.f[c], j, P) + i, pcodec<KF>(fa.own_sc ? fa.isc[c] : sd.fisc[c], sd.loc));
    };
    const int which = fa.field[q];
    float v;
    if (which == FRAME_P) {
      v = (fld(F_TXX) + fld(F_TZZ)) * 0.5f;
    } else if (which == FRAME_VEL) {
      const float x = fld(F_VX), z = fld(F_VZ);
      v = __builtin_sqrtf(x * x + z * z);
    } else {
      v = fld(which == FRAME_VX ? F_VX : F_VZ);
    }
    base[o] = v;
  }
}

// Standalone strip backup (backup_boundary!, boundary.jl:45-56) for the per-stage API.
template <int WPB>
__global__ __launch_bounds__(64 * WPB) void k_backup(BatchArgs ba) {
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int nsh = *ba.n_shots_p;
  const int total = ba.prefix[nsh];
  for (int item = uniform(blockIdx.x * WPB + wid); item < total;
       item = uniform(item + gridDim.x * WPB)) {
    const int s = uniform(find_shot(ba.prefix, nsh, item));
    const ShotDev& sd = ba.shots[s];
    if (shot_step(sd, ba.k_off) == 0) continue;
    const int idx = (item - ba.prefix[s]) * 64 + lane;
    const int i = idx % sd.nx, j = idx / sd.nx;
    if (j >= sd.nz || !in_strip(sd, i, j)) continue;
    const size_t o = (size_t)j * sd.pitch + i;
#pragma unroll
    for (int c = 0; c < 5; ++c) sd.fo[c][o] = sd.f[c][o];
  }
}

// Standalone source injection (inject_source!, source_receiver.jl:18-30;
// force sources backend/cpu.jl:171-187) for the per-stage API.
__global__ void k_inject(BatchArgs ba) {
  const int nsh = *ba.n_shots_p;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nsh; s += gridDim.x * blockDim.x) {
    const ShotDev& sd = ba.shots[s];
    const int k = shot_step(sd, ba.k_off);
    if (k == 0 || k > sd.nw) continue;
    const float w = sd.wavelet[k - 1];
    const size_t o = (size_t)sd.sj * sd.pitch + sd.si;
    if (sd.stype == SRC_PRESSURE) {
      sd.f[F_TXX][o] = sd.f[F_TXX][o] + w;
      sd.f[F_TZZ][o] = sd.f[F_TZZ][o] + w;
    } else if (sd.stype == SRC_FORCE_Z) {
      sd.f[F_VZ][o] = sd.f[F_VZ][o] + w * sd.src_scale;
    } else {
      sd.f[F_VX][o] = sd.f[F_VX][o] + w * sd.src_scale;
    }
  }
}

// One step counter set on the stream (a pool member joining between runs).
// Gather layout change for the download: device [nt][n_rec] -> Julia's
// [nt x n_rec] column-major (dst[k + r*nt] = src[k*n_rec + r]), through 64 x 64
// LDS tiles so that reads and writes are both row-contiguous.
__global__ __launch_bounds__(256) void k_transpose(const float* src, float* dst, int nt, int nrec) {
  __shared__ float t[64][65];
  const int r0 = blockIdx.x * 64, k0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4 threads
  for (int i = ty; i < 64; i += 4) {
    const int k = k0 + i, r = r0 + tx;
    if (k < nt && r < nrec) t[i][tx] = src[(size_t)k * nrec + r];
  }
  __syncthreads();
  for (int i = ty; i < 64; i += 4) {
    const int r = r0 + i, k = k0 + tx;
    if (k < nt && r < nrec) dst[(size_t)r * nt + k] = t[tx][i];
  }
}

__global__ void k_set_int(int* p, int v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

// Step counters: set (before a run) and advance (last node of a captured chunk).
__global__ void k_set_step(const ShotDev* shots, const int* n_shots_p, int k) {
  const int n = *n_shots_p;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x)
    *shots[s].kp = k;
}
__global__ void k_advance(const ShotDev* shots, const int* n_shots_p, int by) {
  const int n = *n_shots_p;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x)
    *shots[s].kp += by;
}

}  // namespace fomo
