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

enum { FLAG_BACKUP = 1, FLAG_INJECT = 2 };
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

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

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

template <int N>
__device__ __forceinline__ V<N> ldv(const float* p);
template <>
__device__ __forceinline__ V<1> ldv<1>(const float* p) {
  V<1> r;
  r.v[0] = *p;
  return r;
}
template <>
__device__ __forceinline__ V<2> ldv<2>(const float* p) {
  f2v t = *(const f2v*)p;
  V<2> r;
  r.v[0] = t.x;
  r.v[1] = t.y;
  return r;
}
template <>
__device__ __forceinline__ V<4> ldv<4>(const float* p) {
  f4v t = *(const f4v*)p;
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
  *p = a.v[0];
}
template <>
__device__ __forceinline__ void stv<2>(float* p, const V<2>& a) {
  f2v t;
  t.x = a.v[0];
  t.y = a.v[1];
  *(f2v*)p = t;
}
template <>
__device__ __forceinline__ void stv<4>(float* p, const V<4>& a) {
  f4v t;
  t.x = a.v[0];
  t.y = a.v[1];
  t.z = a.v[2];
  t.w = a.v[3];
  *(f4v*)p = t;
}

__device__ __forceinline__ const float* row(const float* base, int j, int pitch) {
  return base + (ptrdiff_t)j * pitch;
}
__device__ __forceinline__ float* row(float* base, int j, int pitch) {
  return base + (ptrdiff_t)j * pitch;
}

// HABC strip membership of copy_boundary_strip! (kernels/boundary.jl:20-38), 0-based.
__device__ __forceinline__ bool in_strip(const ShotDev& sd, int i, int j) {
  const int n = sd.nbc;
  const int jtop = sd.fs ? n : 0;
  return j >= jtop && j < sd.nz && i < sd.nx &&
         (i <= n + 1 || i >= sd.nx - n - 2 || j <= n + 1 || j >= sd.nz - n - 2);
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
template <int M, int CPT, int WPB>
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
    const int t = item - ba.prefix[s];
    const int nx = sd.nx, nz = sd.nz, P = sd.pitch;
    const int i0 = (t % sd.tiles_x) * TW;
    const int j0 = (t / sd.tiles_x) * sd.tz;
    const int j1 = min(j0 + sd.tz, nz);
    const int ib = i0 + lane * CPT;
    int hx = -1;
    if (lane < M) hx = lane;
    else if (lane >= 64 - M) hx = TW + M + (lane - (64 - M));
    const bool has_halo = hx >= 0;
    const int hcol = i0 - M + hx;

    float* vx = sd.f[F_VX];
    float* vz = sd.f[F_VZ];
    const float* txx = sd.f[F_TXX];
    const float* tzz = sd.f[F_TZZ];
    const float* txz = sd.f[F_TXZ];
    const bool src_on = (flags & FLAG_INJECT) && k <= sd.nw;
    const float wav = src_on ? sd.wavelet[k - 1] : 0.f;
    const bool src_p = src_on && sd.stype == SRC_PRESSURE;
    const bool src_vz = src_on && sd.stype == SRC_FORCE_Z;
    const bool src_vx = src_on && sd.stype == SRC_FORCE_X;
    const int si = sd.si, sj = sd.sj;
    float a[M];
#pragma unroll
    for (int l = 0; l < M; ++l) a[l] = sd.a[l];

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
    float c_htxx = has_halo ? row(txx, j0, P)[hcol] : 0.f;
    float c_htxz = has_halo ? row(txz, j0, P)[hcol] : 0.f;
    V<CPT> c_vx = ldv<CPT>(row(vx, j0, P) + ib);
    V<CPT> c_vz = ldv<CPT>(row(vz, j0, P) + ib);
    V<CPT> c_b1 = ldv<CPT>(row(sd.bvx_x, j0, P) + ib);
    V<CPT> c_b2 = ldv<CPT>(row(sd.bvx_z, j0, P) + ib);
    V<CPT> c_b3 = ldv<CPT>(row(sd.bvz_x, j0, P) + ib);
    V<CPT> c_b4 = ldv<CPT>(row(sd.bvz_z, j0, P) + ib);

    for (int j = j0; j < j1; ++j) {
      // Prefetch row j+1 (and the window tops) while row j is computed.
      const int jn = j + 1;
      V<CPT> n_txx = ldv<CPT>(row(txx, jn, P) + ib);
      float n_htxx = has_halo ? row(txx, jn, P)[hcol] : 0.f;
      float n_htxz = has_halo ? row(txz, jn, P)[hcol] : 0.f;
      V<CPT> n_txz = ldv<CPT>(row(txz, j + M, P) + ib);
      V<CPT> n_tzz = ldv<CPT>(row(tzz, j + M + 1, P) + ib);
      V<CPT> n_vx = ldv<CPT>(row(vx, jn, P) + ib);
      V<CPT> n_vz = ldv<CPT>(row(vz, jn, P) + ib);
      V<CPT> n_b1 = ldv<CPT>(row(sd.bvx_x, jn, P) + ib);
      V<CPT> n_b2 = ldv<CPT>(row(sd.bvx_z, jn, P) + ib);
      V<CPT> n_b3 = ldv<CPT>(row(sd.bvz_x, jn, P) + ib);
      V<CPT> n_b4 = ldv<CPT>(row(sd.bvz_z, jn, P) + ib);

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
          if (src_vz) vzs = vzs + wav * sd.src_scale;
          if (src_vx) vxs = vxs + wav * sd.src_scale;
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
          if (in_strip(sd, ib + c, j)) {
            row(sd.fo[F_VX], j, P)[ib + c] = c_vx.v[c];
            row(sd.fo[F_VZ], j, P)[ib + c] = c_vz.v[c];
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
    const int nx = sd.nx, nz = sd.nz, P = sd.pitch;
    const int i0 = (t % sd.tiles_x) * TW;
    const int j0 = (t / sd.tiles_x) * sd.tz;
    const int j1 = min(j0 + sd.tz, nz);
    const int ib = i0 + lane * CPT;
    int hx = -1;
    if (lane < M) hx = lane;
    else if (lane >= 64 - M) hx = TW + M + (lane - (64 - M));
    const bool has_halo = hx >= 0;
    const int hcol = i0 - M + hx;

    const float* vx = sd.f[F_VX];
    const float* vz = sd.f[F_VZ];
    float* txx = sd.f[F_TXX];
    float* tzz = sd.f[F_TZZ];
    float* txz = sd.f[F_TXZ];
    const bool src_on = (flags & FLAG_INJECT) && k <= sd.nw && sd.stype == SRC_PRESSURE;
    const float wav = src_on ? sd.wavelet[k - 1] : 0.f;
    const int si = sd.si, sj = sd.sj;
    const float dtx = sd.dtx, dtz = sd.dtz;
    float a[M];
#pragma unroll
    for (int l = 0; l < M; ++l) a[l] = sd.a[l];

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
    float c_hvx = has_halo ? row(vx, j0, P)[hcol] : 0.f;
    float c_hvz = has_halo ? row(vz, j0, P)[hcol] : 0.f;
    V<CPT> c_lam = ldv<CPT>(row(sd.lam, j0, P) + ib);
    V<CPT> c_l2m = ldv<CPT>(row(sd.l2m, j0, P) + ib);
    V<CPT> c_mu = ldv<CPT>(row(sd.mu_txz, j0, P) + ib);
    V<CPT> c_txx = ldv<CPT>(row(txx, j0, P) + ib);
    V<CPT> c_tzz = ldv<CPT>(row(tzz, j0, P) + ib);
    V<CPT> c_txz = ldv<CPT>(row(txz, j0, P) + ib);

    for (int j = j0; j < j1; ++j) {
      const int jn = j + 1;
      float n_hvx = has_halo ? row(vx, jn, P)[hcol] : 0.f;
      float n_hvz = has_halo ? row(vz, jn, P)[hcol] : 0.f;
      V<CPT> n_vx = ldv<CPT>(row(vx, j + M + 1, P) + ib);
      V<CPT> n_vz = ldv<CPT>(row(vz, j + M, P) + ib);
      V<CPT> n_lam = ldv<CPT>(row(sd.lam, jn, P) + ib);
      V<CPT> n_l2m = ldv<CPT>(row(sd.l2m, jn, P) + ib);
      V<CPT> n_mu = ldv<CPT>(row(sd.mu_txz, jn, P) + ib);
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
          if (in_strip(sd, ib + c, j)) {
            row(sd.fo[F_TXX], j, P)[ib + c] = c_txx.v[c];
            row(sd.fo[F_TZZ], j, P)[ib + c] = c_tzz.v[c];
            row(sd.fo[F_TXZ], j, P)[ib + c] = c_txz.v[c];
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
// HABC (kernels/boundary.jl:68-132) in the literal program order of the CPU
// loops: left, right, bottom, [top], left-bottom, right-bottom, [left-top,
// right-top]; j outer, i inner; every read sees the current array.  That
// order makes the left and top edges Jacobi sweeps over pre-HABC values, the
// right edge a recurrence along +x per row, the bottom edge a recurrence
// along +z per column, the bottom-right corner a 2D recurrence (run as
// anti-diagonal wavefronts), and so on; the pieces below follow it exactly.
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

struct HabcField {
  float* f;
  const float* fo;
  const float* ramp;
};

__device__ __forceinline__ HabcField habc_field(const ShotDev& sd, int stress, int fi) {
  HabcField h;
  const int c = stress ? F_TXX + fi : fi;
  h.f = sd.f[c];
  h.fo = sd.fo[c];
  h.ramp = sd.ramp + (stress ? RAMP_TAU : fi) * sd.ramp_len;
  return h;
}

constexpr int LROWS = 8;  // rows per left-edge piece

template <int WPB>
__global__ __launch_bounds__(64 * WPB) void k_habc1(BatchArgs ba, int stress) {
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int nsh = *ba.n_shots_p;
  const int total = ba.prefix[nsh];
  for (int item = uniform(blockIdx.x * WPB + wid); item < total;
       item = uniform(item + gridDim.x * WPB)) {
    const int s = uniform(find_shot(ba.prefix, nsh, item));
    const ShotDev& sd = ba.shots[s];
    if (shot_step(sd, ba.k_off) == 0) continue;
    const int t = item - ba.prefix[s];
    const int per = sd.n_l + sd.n_r + sd.n_b;
    const int fi = t / per;
    int p = t % per;
    const HabcField h = habc_field(sd, stress, fi);
    const Hq q{sd.qx, sd.qz, sd.qt_x, sd.qt_z, sd.qxt};
    const int n = sd.nbc, nx = sd.nx, nz = sd.nz, P = sd.pitch;
    const int r0 = sd.edge_r0, r1 = sd.edge_r0 + sd.edge_rows;  // rows [r0, r1)
    if (p < sd.n_l) {
      // Left edge: i in [1, n], Jacobi on pre-HABC values of the same row.
      const bool act = lane < n;
      const int i = 1 + lane;
      const int ja = r0 + p * LROWS, jb = min(ja + LROWS, r1);
      for (int j = ja; j < jb; ++j) {
        float* fr = row(h.f, j, P);
        const float* gr = row(h.fo, j, P);
        const float fv = act ? fr[i] : 0.f;
        const float ov = act ? gr[i] : 0.f;
        float fn = __shfl_down(fv, 1);
        float on = __shfl_down(ov, 1);
        if (lane == n - 1) {
          fn = fr[n + 1];
          on = gr[n + 1];
        }
        if (act) fr[i] = q.edge(habc_w(sd, h.ramp, i, j), fv, q.sumx(fn, ov, on));
      }
      continue;
    }
    p -= sd.n_l;
    if (p < sd.n_r) {
      // Right edge: one row per lane, recurrence along +x.
      const int j = r0 + p * 64 + lane;
      if (j >= r1) continue;
      float* fr = row(h.f, j, P);
      const float* gr = row(h.fo, j, P);
      float prev = fr[nx - n - 2];
      float op = gr[nx - n - 2];
      for (int i = nx - n - 1; i <= nx - 2; ++i) {
        const float fv = fr[i], ov = gr[i];
        const float out = q.edge(habc_w(sd, h.ramp, i, j), fv, q.sumx(prev, ov, op));
        fr[i] = out;
        prev = out;
        op = ov;
      }
      continue;
    }
    p -= sd.n_r;
    {
      // Bottom edge: one column per lane, recurrence along +z.
      const int i = n + 1 + p * 64 + lane;
      if (i > nx - n - 2) continue;
      float prev = row(h.f, nz - n - 2, P)[i];
      float op = row(h.fo, nz - n - 2, P)[i];
      for (int j = nz - n - 1; j <= nz - 2; ++j) {
        float* fp = row(h.f, j, P) + i;
        const float fv = *fp, ov = row(h.fo, j, P)[i];
        const float out = q.edge(habc_w(sd, h.ramp, i, j), fv, q.sumz(prev, ov, op));
        *fp = out;
        prev = out;
        op = ov;
      }
    }
  }
}

// Phase 2: top edge (no free surface), left-bottom and right-bottom corners.
template <int WPB>
__global__ __launch_bounds__(64 * WPB) void k_habc2(BatchArgs ba, int stress) {
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int nsh = *ba.n_shots_p;
  const int total = ba.prefix[nsh];
  for (int item = uniform(blockIdx.x * WPB + wid); item < total;
       item = uniform(item + gridDim.x * WPB)) {
    const int s = uniform(find_shot(ba.prefix, nsh, item));
    const ShotDev& sd = ba.shots[s];
    if (shot_step(sd, ba.k_off) == 0) continue;
    const int t = item - ba.prefix[s];
    const int per = sd.n_t + 2;
    const int fi = t / per;
    const int p = t % per;
    const HabcField h = habc_field(sd, stress, fi);
    const Hq q{sd.qx, sd.qz, sd.qt_x, sd.qt_z, sd.qxt};
    const int n = sd.nbc, nx = sd.nx, nz = sd.nz, P = sd.pitch;
    if (p < sd.n_t) {
      // Top edge: one column per lane, Jacobi in z over rows 1..n.
      const int i = n + 1 + p * 64 + lane;
      if (i > nx - n - 2) continue;
      float fv = row(h.f, 1, P)[i], ov = row(h.fo, 1, P)[i];
      for (int j = 1; j <= n; ++j) {
        const float fn = row(h.f, j + 1, P)[i], on = row(h.fo, j + 1, P)[i];
        row(h.f, j, P)[i] = q.edge(habc_w(sd, h.ramp, i, j), fv, q.sumz(fn, ov, on));
        fv = fn;
        ov = on;
      }
    } else if (p == sd.n_t) {
      // Left-bottom corner: one column per lane, recurrence along +z; the
      // right neighbour is read before it is updated (same row, i ascending).
      const bool act = lane < n;
      const int i = 1 + lane;
      float up = act ? row(h.f, nz - n - 2, P)[i] : 0.f;
      float ou = act ? row(h.fo, nz - n - 2, P)[i] : 0.f;
      for (int j = nz - n - 1; j <= nz - 2; ++j) {
        float* fr = row(h.f, j, P);
        const float* gr = row(h.fo, j, P);
        const float fv = act ? fr[i] : 0.f;
        const float ov = act ? gr[i] : 0.f;
        float fnb = __shfl_down(fv, 1);
        float onb = __shfl_down(ov, 1);
        if (lane == n - 1) {
          fnb = fr[n + 1];
          onb = gr[n + 1];
        }
        if (act) {
          const float out =
              q.corner(habc_w(sd, h.ramp, i, j), fv, q.sumx(fnb, ov, onb), q.sumz(up, ov, ou));
          fr[i] = out;
          up = out;
          ou = ov;
        }
      }
    } else {
      // Right-bottom corner: 2D recurrence (+x and +z), anti-diagonal
      // wavefronts; lane c owns column nx-n-1+c and hands its newest value
      // to lane c+1 one step later.
      const int c = lane;
      const int i = nx - n - 1 + c;
      const bool col_ok = c < n;
      float up = col_ok ? row(h.f, nz - n - 2, P)[i] : 0.f;
      float ou = col_ok ? row(h.fo, nz - n - 2, P)[i] : 0.f;
      float last = 0.f;
      for (int st = 0; st <= 2 * n - 2; ++st) {
        const int r = st - c;
        const bool act = col_ok && r >= 0 && r < n;
        const int j = nz - n - 1 + r;
        float left = __shfl_up(last, 1);
        if (act) {
          float* fr = row(h.f, j, P);
          const float* gr = row(h.fo, j, P);
          if (c == 0) left = fr[i - 1];
          const float fv = fr[i], ov = gr[i], ol = gr[i - 1];
          const float out =
              q.corner(habc_w(sd, h.ramp, i, j), fv, q.sumx(left, ov, ol), q.sumz(up, ov, ou));
          fr[i] = out;
          last = out;
          up = out;
          ou = ov;
        }
      }
    }
  }
}

// Phase 3 (no free surface): left-top and right-top corners.
template <int WPB>
__global__ __launch_bounds__(64 * WPB) void k_habc3(BatchArgs ba, int stress) {
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int nsh = *ba.n_shots_p;
  const int total = ba.prefix[nsh];
  for (int item = uniform(blockIdx.x * WPB + wid); item < total;
       item = uniform(item + gridDim.x * WPB)) {
    const int s = uniform(find_shot(ba.prefix, nsh, item));
    const ShotDev& sd = ba.shots[s];
    if (shot_step(sd, ba.k_off) == 0) continue;
    const int t = item - ba.prefix[s];
    const int fi = t / 2;
    const int p = t % 2;
    const HabcField h = habc_field(sd, stress, fi);
    const Hq q{sd.qx, sd.qz, sd.qt_x, sd.qt_z, sd.qxt};
    const int n = sd.nbc, nx = sd.nx, P = sd.pitch;
    const bool act = lane < n;
    if (p == 0) {
      // Left-top: Jacobi over the post-left-edge values, rows 1..n ascending.
      const int i = 1 + lane;
      float fv = act ? row(h.f, 1, P)[i] : 0.f;
      float ov = act ? row(h.fo, 1, P)[i] : 0.f;
      for (int j = 1; j <= n; ++j) {
        float* fr = row(h.f, j, P);
        const float* gr = row(h.fo, j, P);
        const float fd = act ? row(h.f, j + 1, P)[i] : 0.f;
        const float od = act ? row(h.fo, j + 1, P)[i] : 0.f;
        float fnb = __shfl_down(fv, 1);
        float onb = __shfl_down(ov, 1);
        if (lane == n - 1) {
          fnb = fr[n + 1];
          onb = gr[n + 1];
        }
        if (act)
          fr[i] = q.corner(habc_w(sd, h.ramp, i, j), fv, q.sumx(fnb, ov, onb), q.sumz(fd, ov, od));
        fv = fd;
        ov = od;
      }
    } else {
      // Right-top: one row per lane, recurrence along +x; the row below is
      // read before it is updated (rows ascending).
      const int j = 1 + lane;
      float prev = act ? row(h.f, j, P)[nx - n - 2] : 0.f;
      float op = act ? row(h.fo, j, P)[nx - n - 2] : 0.f;
      for (int i = nx - n - 1; i <= nx - 2; ++i) {
        const float fv = act ? row(h.f, j, P)[i] : 0.f;
        const float ov = act ? row(h.fo, j, P)[i] : 0.f;
        float fd = __shfl_down(fv, 1);
        float od = __shfl_down(ov, 1);
        if (lane == n - 1) {
          fd = row(h.f, n + 1, P)[i];
          od = row(h.fo, n + 1, P)[i];
        }
        if (act) {
          const float out =
              q.corner(habc_w(sd, h.ramp, i, j), fv, q.sumx(prev, ov, op), q.sumz(fd, ov, od));
          row(h.f, j, P)[i] = out;
          prev = out;
          op = ov;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Free surface (kernels/boundary.jl:168-181) and receiver sampling
// (kernels/source_receiver.jl:65-77).  Items [0, 6*nx) zero tzz/txz on rows
// pad-5..pad; the rest sample receivers into gather[k-1][r].  A pressure
// receiver on a free-surface row reads tzz as the 0 it is being set to.
// ---------------------------------------------------------------------------
template <int WPB>
__global__ __launch_bounds__(64 * WPB) void k_fsrec(BatchArgs ba, int mode) {
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
        row(sd.f[F_TZZ], j, P)[i] = 0.f;
        row(sd.f[F_TXZ], j, P)[i] = 0.f;
      }
      continue;
    }
    const int r = idx - nfs;
    if (!(mode & MODE_RECORD) || r >= sd.n_rec) continue;
    const int ri = sd.rec[3 * r], rj = sd.rec[3 * r + 1], rt = sd.rec[3 * r + 2];
    float v;
    if (rt == REC_VZ) v = row(sd.f[F_VZ], rj, P)[ri];
    else if (rt == REC_VX) v = row(sd.f[F_VX], rj, P)[ri];
    else {
      const bool zero_tzz = fs_on && rj >= sd.pad - 5 && rj <= sd.pad;
      const float zz = zero_tzz ? 0.f : row(sd.f[F_TZZ], rj, P)[ri];
      v = (row(sd.f[F_TXX], rj, P)[ri] + zz) * 0.5f;
    }
    sd.gather[(size_t)(k - 1) * sd.n_rec + r] = v;
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
