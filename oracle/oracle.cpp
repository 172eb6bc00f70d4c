// ORACLE -- test infrastructure only.
//
// A CPU restatement of the reference's time-stepping path, used by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg, to check results against.  The
// product (fomo_julia_amd/) never links, loads or calls it.
//
// Arrays use the reference layout: (nx, nz) column-major, element (i, j)
// (0-based here) at i + j*nx; HABC weights are (nz, nx): (j + i*nz).
//
// Functions and the reference code they follow (files under the reference's src/):
//   stage_backup    kernels/boundary.jl:20-56     (copy_boundary_strip!, backup_boundary!)
//   stage_inject    kernels/source_receiver.jl:18-30 (pressure);
//                   backend/cpu.jl:171-187          (force_z / force_x, v1 semantics)
//   stage_velocity  kernels/velocity.jl:21-46
//   stage_habc      kernels/boundary.jl:68-132 (program order), :139-157
//                   snapshot variant: kernels/boundary.jl:251-297 with every
//                   neighbour read taken from the pre-HABC field
//   stage_stress    kernels/stress.jl:18-47
//   stage_fs        kernels/boundary.jl:168-181
//   stage_record    kernels/source_receiver.jl:65-77
//   time_step       simulation/time_stepper.jl:39-63
//
// Arithmetic modes (flag bits):
//   ORACLE_FMA       accumulate with fused multiply-add:
//                    acc += a*b -> fma(a, b, acc); e1*f1 + e2*f2 -> fma(e2, f2, e1*f1)
//                    (the HIP kernels use exactly this, so FP32 parity is bit-exact)
//                    otherwise every product and sum is rounded separately.
//   ORACLE_SNAPSHOT  HABC neighbour reads from the pre-HABC snapshot, one
//                    formula per cell (the CUDA kernel without its race);
//                    otherwise the literal program order of the CPU loops.
//
// Reduced-precision storage (desc.storage 1 = BF16, 2 = FP16, 3 = FP16
// coefficients with FP32 fields; FP32 model only; the reference has no such
// mode): the storage contract of the MI355X path, restated sequentially.
//  - The seven per-cell coefficients (k = 0..6: dtx/rho_vx, dtz/rho_vx,
//    dtx/rho_vz, dtz/rho_vz, lam, lam+2mu, mu_txz) are computed in FP32, scaled
//    (FP16: by the power of two that puts the array's largest magnitude in
//    [2^14, 2^15)) and rounded once with a dither: to the grid point below
//    |x|, plus one spacing when the remainder in spacings exceeds
//    u = (fmix32(i ^ fmix32(j ^ fmix32(k' + 0x9E3779B9))) >> 8) / 2^24, with
//    fmix32 MurmurHash3's finalizer and k' = k, except k' = k - 1 for the
//    dtz arrays when dtz == dtx (they share the dtx arrays).
//  - After the velocity, HABC and stress stages every cell of the fields the
//    stage writes is rounded to nearest, ties to even (FP16: at the field's
//    scale desc.fscale[c]); stages compute in FP32 in between, so the injected
//    source and the HABC recurrences see unrounded values within a stage.
// Both roundings are written from the definition of the format's value grid
// (spacing 2^(e-p+1) for a p-bit significand, the subnormal spacing below
// the smallest normal exponent).
//
// Storage 4 (block floating point) is a design study of a 16-bit field
// format: no kernel implements it.  Coefficients stay FP32.  After each
// stage, every row of each field it writes is cut into segments of 64 cells
// from i = 0.  A segment keeps one exponent e, with its largest magnitude in
// [2^(e-1), 2^e), and per cell a signed 16-bit mantissa on the spacing
// 2^(e-15), rounded to nearest (ties to even) and clamped to +-32767.
// Storage 5 is the same field format with storage 3's FP16 coefficients.
// Storage 6 (F24): FP32 coefficients; after each stage every cell of the
// fields it writes is rounded to nearest, ties to even, on the grid of a
// 16-bit significand with FP32's exponent range (spacing 2^(e-15) for
// |x| in [2^(e-1), 2^e), 2^-141 below 2^-126).
//
// Pinning: see tests/test_oracle.py (polynomial-exactness of the stencils,
// closed-form HABC cases, free surface, symmetry, 2^k linearity, energy decay,
// the reference's shot_1.bin gather).
#include <cmath>
#include <cstdint>
#include <algorithm>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { ORACLE_FMA = 1, ORACLE_SNAPSHOT = 2 };
enum { SRC_PRESSURE = 0, SRC_FORCE_Z = 1, SRC_FORCE_X = 2 };
enum { REC_VZ = 0, REC_VX = 1, REC_P = 2 };

extern "C" {
// Plain description of one run (all indices 1-based as in Julia).
typedef struct {
  int32_t nx, nz, pad, nbc, fd_order, free_surface;
  double dt, dtx, dtz;
  double qx, qz, qt_x, qt_z, qxt;
  int32_t flags;
  int32_t threads;
  int32_t storage;   // 0 = FP32 (no rounding), 1 = BF16, 2 = FP16, 3-6 (see above)
  float fscale[5];   // FP16 store scales of vx vz txx tzz txz
} oracle_desc;
}

namespace {

template <typename T>
struct Model {
  const oracle_desc* d;
  int nx, nz, M, nbc, pad;
  bool fs, fma_mode, snapshot;
  T dt, dtx, dtz, qx, qz, qt_x, qt_z, qxt;
  const T* a;
  const T *rho_vx, *rho_vz, *lam, *mu_txx, *mu_txz;
  const T *w_vx, *w_vz, *w_tau;
  T* f[10];  // vx vz txx tzz txz, then the *_old of each
  int storage = 0;
  bool cq_on = false;    // stages read the rounded coefficients cq
  std::vector<T> cq[7];  // rounded coefficients (reduced-precision storage)
};

// Nearest value, ties to even, of a binary format with a p-bit significand
// and smallest normal exponent emin (values beyond the largest finite one
// are not expected here).
static double round_format(double x, int p, int emin) {
  if (x == 0 || !std::isfinite(x)) return x;
  int e;
  std::frexp(x, &e);  // |x| in [2^(e-1), 2^e)
  const int ex = std::max(e - 1, emin);
  const double spacing = std::ldexp(1.0, ex - (p - 1));
  return std::nearbyint(x / spacing) * spacing;  // default mode: ties to even
}
// Field rounding of the storage mode (mode 3 keeps FP32 fields).
static float round_storage(int storage, float x, float scale) {
  if (storage == 1) return (float)round_format((double)x, 8, -126);          // BF16
  if (storage == 2) return (float)(round_format((double)x * scale, 11, -14) / scale);  // FP16
  if (storage == 6) return (float)round_format((double)x, 16, -126);        // F24
  return x;
}
static uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}
// Coefficient k' of cell (i, j) (0-based), dither-rounded (see the header).
static float dither_coef(int storage, float x, float scale, uint32_t i, uint32_t j, uint32_t kp) {
  const uint32_t h = fmix32(i ^ fmix32(j ^ fmix32(kp + 0x9E3779B9u)));
  const double u = (double)(h >> 8) / 16777216.0;
  const int p = storage == 1 ? 8 : 11, emin = storage == 1 ? -126 : -14;
  const double v = (double)x * scale;
  if (v == 0) return 0.f;
  int e;
  std::frexp(v, &e);
  const double spacing = std::ldexp(1.0, std::max(e - 1, emin) - (p - 1));
  const double a = std::fabs(v) / spacing;
  const double fl = std::floor(a);
  const double q = std::copysign((a - fl > u ? fl + 1 : fl) * spacing, v);
  return (float)(q / scale);
}
// One block floating-point segment of n cells, rounded in place (storage 4).
template <typename T>
static void round_block(T* x, int n) {
  double mx = 0;
  for (int t = 0; t < n; ++t) mx = std::max(mx, std::fabs((double)x[t]));
  if (!(mx > 0) || !std::isfinite(mx)) return;
  int e;
  std::frexp(mx, &e);  // mx in [2^(e-1), 2^e)
  const double spacing = std::ldexp(1.0, e - 15);
  for (int t = 0; t < n; ++t) {
    const double q = std::min(32767.0, std::max(-32767.0, std::nearbyint((double)x[t] / spacing)));
    x[t] = (T)(float)(q * spacing);
  }
}
template <typename T>
static void round_fields(Model<T>& m, int c0, int c1) {
  // every cell (segment) is rounded on its own, so rows go to threads
  if (m.storage == 4 || m.storage == 5) {
    for (int c = c0; c < c1; ++c) {
#pragma omp parallel for schedule(static) num_threads(m.d->threads)
      for (int j = 0; j < m.nz; ++j)
        for (int i0 = 0; i0 < m.nx; i0 += 64)
          round_block(m.f[c] + (size_t)j * m.nx + i0, std::min(64, m.nx - i0));
    }
    return;
  }
  if (m.storage != 1 && m.storage != 2 && m.storage != 6) return;
  for (int c = c0; c < c1; ++c) {
    const float sc = m.d->fscale[c];
#pragma omp parallel for schedule(static) num_threads(m.d->threads)
    for (int j = 0; j < m.nz; ++j) {
      T* row = m.f[c] + (size_t)j * m.nx;
      for (int i = 0; i < m.nx; ++i) row[i] = (T)round_storage(m.storage, (float)row[i], sc);
    }
  }
}
// Power of two putting max |x| in [2^14, 2^15) (FP16 coefficient scale).
static float coef_scale(const std::vector<float>& x) {
  double mx = 0;
  for (float v : x) mx = std::max(mx, (double)std::fabs(v));
  if (!(mx > 0)) return 1.f;
  return std::ldexp(1.0f, 14 - (int)std::floor(std::log2(mx)));
}

template <typename T>
static inline T madd(bool fm, T a, T b, T c) {  // a*b + c
  return fm ? std::fma(a, b, c) : (T)(a * b) + c;
}

// kernels/boundary.jl:20-38
template <typename T>
static void copy_strip(T* old, const T* cur, int nx, int nz, int nbc, bool fs) {
  int j_top = fs ? nbc + 1 : 1;
  auto cp = [&](int i, int j) {  // 1-based
    size_t o = (size_t)(i - 1) + (size_t)(j - 1) * nx;
    old[o] = cur[o];
  };
  for (int i = 1; i <= nbc + 2; ++i)
    for (int j = j_top; j <= nz; ++j) cp(i, j);
  for (int i = nx - nbc - 1; i <= nx; ++i)
    for (int j = j_top; j <= nz; ++j) cp(i, j);
  for (int j = j_top; j <= nbc + 2; ++j)
    for (int i = nbc + 3; i <= nx - nbc - 2; ++i) cp(i, j);
  for (int j = nz - nbc - 1; j <= nz; ++j)
    for (int i = nbc + 3; i <= nx - nbc - 2; ++i) cp(i, j);
}

template <typename T>
static void stage_backup(Model<T>& m) {
  for (int c = 0; c < 5; ++c) copy_strip(m.f[5 + c], m.f[c], m.nx, m.nz, m.nbc, m.fs);
}

template <typename T>
static void stage_inject(Model<T>& m, int src_i, int src_j, int src_type, const T* w, int nw, int k) {
  if (k > nw || k < 1) return;
  T wav = w[k - 1];
  size_t o = (size_t)(src_i - 1) + (size_t)(src_j - 1) * m.nx;
  if (src_type == SRC_PRESSURE) {
    m.f[2][o] += wav;
    m.f[3][o] += wav;
  } else if (src_type == SRC_FORCE_Z) {
    T s = (T)(m.dt / m.rho_vz[o]);
    m.f[1][o] += (T)(wav * s);
  } else if (src_type == SRC_FORCE_X) {
    T s = (T)(m.dt / m.rho_vx[o]);
    m.f[0][o] += (T)(wav * s);
  }
}

// kernels/velocity.jl:21-46
template <typename T>
static void stage_velocity(Model<T>& m) {
  const int nx = m.nx, nz = m.nz, M = m.M;
  T *vx = m.f[0], *vz = m.f[1];
  const T *txx = m.f[2], *tzz = m.f[3], *txz = m.f[4];
  const bool fm = m.fma_mode;
#pragma omp parallel for schedule(static) num_threads(m.d->threads)
  for (int j = M; j < nz - M; ++j) {
    for (int i = M; i < nx - M; ++i) {
      auto at = [&](const T* p, int ii, int jj) { return p[(size_t)ii + (size_t)jj * nx]; };
      T d1 = 0, d2 = 0, d3 = 0, d4 = 0;
      for (int l = 1; l <= M; ++l) {
        T al = m.a[l - 1];
        d1 = madd<T>(fm, al, at(txx, i + l - 1, j) - at(txx, i - l, j), d1);
        d2 = madd<T>(fm, al, at(txz, i, j + l - 1) - at(txz, i, j - l), d2);
        d3 = madd<T>(fm, al, at(txz, i + l, j) - at(txz, i - l + 1, j), d3);
        d4 = madd<T>(fm, al, at(tzz, i, j + l) - at(tzz, i, j - l + 1), d4);
      }
      size_t o = (size_t)i + (size_t)j * nx;
      T bx = m.dtx / m.rho_vx[o], bz = m.dtz / m.rho_vx[o];
      if (m.cq_on) {
        bx = m.cq[0][o];
        bz = m.cq[1][o];
      }
      T inc = madd<T>(fm, bz, d2, (T)(bx * d1));
      vx[o] = vx[o] + inc;
      bx = m.dtx / m.rho_vz[o];
      bz = m.dtz / m.rho_vz[o];
      if (m.cq_on) {
        bx = m.cq[2][o];
        bz = m.cq[3][o];
      }
      inc = madd<T>(fm, bz, d4, (T)(bx * d3));
      vz[o] = vz[o] + inc;
    }
  }
}

// kernels/stress.jl:18-47
template <typename T>
static void stage_stress(Model<T>& m) {
  const int nx = m.nx, nz = m.nz, M = m.M;
  const T *vx = m.f[0], *vz = m.f[1];
  T *txx = m.f[2], *tzz = m.f[3], *txz = m.f[4];
  const bool fm = m.fma_mode;
#pragma omp parallel for schedule(static) num_threads(m.d->threads)
  for (int j = M; j < nz - M; ++j) {
    for (int i = M; i < nx - M; ++i) {
      auto at = [&](const T* p, int ii, int jj) { return p[(size_t)ii + (size_t)jj * nx]; };
      T d1 = 0, d2 = 0, d3 = 0, d4 = 0;
      for (int l = 1; l <= M; ++l) {
        T al = m.a[l - 1];
        d1 = madd<T>(fm, al, at(vx, i + l, j) - at(vx, i - l + 1, j), d1);      // dvxdx
        d2 = madd<T>(fm, al, at(vz, i, j + l - 1) - at(vz, i, j - l), d2);      // dvzdz
        d3 = madd<T>(fm, al, at(vx, i, j + l) - at(vx, i, j - l + 1), d3);      // dvxdz
        d4 = madd<T>(fm, al, at(vz, i + l - 1, j) - at(vz, i - l, j), d4);      // dvzdx
      }
      size_t o = (size_t)i + (size_t)j * nx;
      T l_val = m.lam[o], m_val = m.mu_txx[o];
      T l2m = l_val + (T)2 * m_val;
      T mu_xz = m.mu_txz[o];
      if (m.cq_on) {
        l_val = m.cq[4][o];
        l2m = m.cq[5][o];
        mu_xz = m.cq[6][o];
      }
      T t1 = d1 * m.dtx, t2 = d2 * m.dtz;
      txx[o] = txx[o] + madd<T>(fm, l_val, t2, (T)(l2m * t1));
      tzz[o] = tzz[o] + madd<T>(fm, l2m, t2, (T)(l_val * t1));
      T s = madd<T>(fm, d4, m.dtx, (T)(d3 * m.dtz));
      txz[o] = txz[o] + (T)(mu_xz * s);
    }
  }
}

// One HABC blend of f(i,j) (1-based) given its neighbour values.
template <typename T>
struct Habc {
  T qx, qz, qt_x, qt_z, qxt;
  bool fm;
  T sumx(T fn, T fo, T fon) const {  // -qx*fn - qt_x*fo - qxt*fon
    T s = (T)(-qx * fn);
    s = madd<T>(fm, -qt_x, fo, s);
    return madd<T>(fm, -qxt, fon, s);
  }
  T sumz(T fn, T fo, T fon) const {
    T s = (T)(-qz * fn);
    s = madd<T>(fm, -qt_z, fo, s);
    return madd<T>(fm, -qxt, fon, s);
  }
  T edge(T w, T f, T s) const { return madd<T>(fm, (T)((T)1 - w), s, (T)(w * f)); }
  T corner(T w, T f, T sx, T sz) const {
    T c = (T)((T)((T)1 - w) * (T)0.5);
    return madd<T>(fm, c, (T)(sx + sz), (T)(w * f));
  }
};

// kernels/boundary.jl:68-132, literal program order (regions L, R, B, T, LB, RB, LT, RT;
// j outer, i inner, every read sees the current array).
template <typename T>
static void habc_program_order(Model<T>& m, T* f, const T* fo, const T* w) {
  const int nx = m.nx, nz = m.nz, nbc = m.nbc;
  Habc<T> h{m.qx, m.qz, m.qt_x, m.qt_z, m.qxt, m.fma_mode};
  auto F = [&](int i, int j) -> T& { return f[(size_t)(i - 1) + (size_t)(j - 1) * nx]; };
  auto O = [&](int i, int j) { return fo[(size_t)(i - 1) + (size_t)(j - 1) * nx]; };
  auto W = [&](int i, int j) { return w[(size_t)(j - 1) + (size_t)(i - 1) * nz]; };
  const int js = m.fs ? nbc + 1 : 2;
  for (int j = js; j <= nz - nbc - 1; ++j)  // left
    for (int i = 2; i <= nbc + 1; ++i)
      F(i, j) = h.edge(W(i, j), F(i, j), h.sumx(F(i + 1, j), O(i, j), O(i + 1, j)));
  for (int j = js; j <= nz - nbc - 1; ++j)  // right
    for (int i = nx - nbc; i <= nx - 1; ++i)
      F(i, j) = h.edge(W(i, j), F(i, j), h.sumx(F(i - 1, j), O(i, j), O(i - 1, j)));
  for (int j = nz - nbc; j <= nz - 1; ++j)  // bottom
    for (int i = nbc + 2; i <= nx - nbc - 1; ++i)
      F(i, j) = h.edge(W(i, j), F(i, j), h.sumz(F(i, j - 1), O(i, j), O(i, j - 1)));
  if (!m.fs)
    for (int j = 2; j <= nbc + 1; ++j)  // top
      for (int i = nbc + 2; i <= nx - nbc - 1; ++i)
        F(i, j) = h.edge(W(i, j), F(i, j), h.sumz(F(i, j + 1), O(i, j), O(i, j + 1)));
  for (int j = nz - nbc; j <= nz - 1; ++j)  // left-bottom
    for (int i = 2; i <= nbc + 1; ++i)
      F(i, j) = h.corner(W(i, j), F(i, j), h.sumx(F(i + 1, j), O(i, j), O(i + 1, j)),
                         h.sumz(F(i, j - 1), O(i, j), O(i, j - 1)));
  for (int j = nz - nbc; j <= nz - 1; ++j)  // right-bottom
    for (int i = nx - nbc; i <= nx - 1; ++i)
      F(i, j) = h.corner(W(i, j), F(i, j), h.sumx(F(i - 1, j), O(i, j), O(i - 1, j)),
                         h.sumz(F(i, j - 1), O(i, j), O(i, j - 1)));
  if (!m.fs) {
    for (int j = 2; j <= nbc + 1; ++j)  // left-top
      for (int i = 2; i <= nbc + 1; ++i)
        F(i, j) = h.corner(W(i, j), F(i, j), h.sumx(F(i + 1, j), O(i, j), O(i + 1, j)),
                           h.sumz(F(i, j + 1), O(i, j), O(i, j + 1)));
    for (int j = 2; j <= nbc + 1; ++j)  // right-top
      for (int i = nx - nbc; i <= nx - 1; ++i)
        F(i, j) = h.corner(W(i, j), F(i, j), h.sumx(F(i - 1, j), O(i, j), O(i - 1, j)),
                           h.sumz(F(i, j + 1), O(i, j), O(i, j + 1)));
  }
}

// kernels/boundary.jl:251-297 with all reads from the pre-HABC snapshot.
template <typename T>
static void habc_snapshot(Model<T>& m, T* f, const T* fo, const T* w) {
  const int nx = m.nx, nz = m.nz, nbc = m.nbc;
  Habc<T> h{m.qx, m.qz, m.qt_x, m.qt_z, m.qxt, m.fma_mode};
  std::vector<T> s(f, f + (size_t)nx * nz);
  auto S = [&](int i, int j) { return s[(size_t)(i - 1) + (size_t)(j - 1) * nx]; };
  auto O = [&](int i, int j) { return fo[(size_t)(i - 1) + (size_t)(j - 1) * nx]; };
  auto W = [&](int i, int j) { return w[(size_t)(j - 1) + (size_t)(i - 1) * nz]; };
  const int js = m.fs ? nbc + 1 : 2;
  for (int j = js; j <= nz - 1; ++j)
    for (int i = 2; i <= nx - 1; ++i) {
      bool L = i <= nbc + 1, R = i >= nx - nbc, B = j >= nz - nbc, Tp = !m.fs && j <= nbc + 1;
      T wt = W(i, j), v = S(i, j), out;
      if (L && !B && !Tp) out = h.edge(wt, v, h.sumx(S(i + 1, j), O(i, j), O(i + 1, j)));
      else if (R && !B && !Tp) out = h.edge(wt, v, h.sumx(S(i - 1, j), O(i, j), O(i - 1, j)));
      else if (B && !L && !R) out = h.edge(wt, v, h.sumz(S(i, j - 1), O(i, j), O(i, j - 1)));
      else if (Tp && !L && !R) out = h.edge(wt, v, h.sumz(S(i, j + 1), O(i, j), O(i, j + 1)));
      else if (L && B) out = h.corner(wt, v, h.sumx(S(i + 1, j), O(i, j), O(i + 1, j)), h.sumz(S(i, j - 1), O(i, j), O(i, j - 1)));
      else if (R && B) out = h.corner(wt, v, h.sumx(S(i - 1, j), O(i, j), O(i - 1, j)), h.sumz(S(i, j - 1), O(i, j), O(i, j - 1)));
      else if (L && Tp) out = h.corner(wt, v, h.sumx(S(i + 1, j), O(i, j), O(i + 1, j)), h.sumz(S(i, j + 1), O(i, j), O(i, j + 1)));
      else if (R && Tp) out = h.corner(wt, v, h.sumx(S(i - 1, j), O(i, j), O(i - 1, j)), h.sumz(S(i, j + 1), O(i, j), O(i, j + 1)));
      else continue;
      f[(size_t)(i - 1) + (size_t)(j - 1) * nx] = out;
    }
}

template <typename T>
static void stage_habc(Model<T>& m, int which) {  // 0 = velocity, 1 = stress
  auto run = [&](int c, const T* w) {
    if (m.snapshot) habc_snapshot(m, m.f[c], m.f[5 + c], w);
    else habc_program_order(m, m.f[c], m.f[5 + c], w);
  };
  if (which == 0) {
    run(0, m.w_vx);
    run(1, m.w_vz);
  } else {
    run(2, m.w_tau);
    run(3, m.w_tau);
    run(4, m.w_tau);
  }
}

// kernels/boundary.jl:168-181
template <typename T>
static void stage_fs(Model<T>& m) {
  if (!m.fs) return;
  int j_fs = m.pad + 1;
  for (int j = j_fs - 5; j <= j_fs; ++j) {
    if (j < 1 || j > m.nz) continue;
    for (int i = 1; i <= m.nx; ++i) {
      size_t o = (size_t)(i - 1) + (size_t)(j - 1) * m.nx;
      m.f[3][o] = 0;
      m.f[4][o] = 0;
    }
  }
}

// kernels/source_receiver.jl:65-77 (gather is [nt x n_rec] column-major)
template <typename T>
static void stage_record(Model<T>& m, const int* ri, const int* rj, const int* rt, int n_rec,
                         int nt, int k, T* gather) {
  if (k < 1 || k > nt) return;
  for (int r = 0; r < n_rec; ++r) {
    size_t o = (size_t)(ri[r] - 1) + (size_t)(rj[r] - 1) * m.nx;
    T v;
    if (rt[r] == REC_VZ) v = m.f[1][o];
    else if (rt[r] == REC_VX) v = m.f[0][o];
    else v = (T)(m.f[2][o] + m.f[3][o]) * (T)0.5;
    gather[(size_t)(k - 1) + (size_t)r * nt] = v;
  }
}

template <typename T>
static Model<T> make_model(const oracle_desc* d, const T* a, const T* const* med, const T* const* w,
                           T* fields) {
  Model<T> m;
  m.d = d;
  m.nx = d->nx;
  m.nz = d->nz;
  m.M = d->fd_order / 2;
  m.nbc = d->nbc;
  m.pad = d->pad;
  m.fs = d->free_surface != 0;
  m.fma_mode = (d->flags & ORACLE_FMA) != 0;
  m.snapshot = (d->flags & ORACLE_SNAPSHOT) != 0;
  m.dt = (T)d->dt;
  m.dtx = (T)d->dtx;
  m.dtz = (T)d->dtz;
  m.qx = (T)d->qx;
  m.qz = (T)d->qz;
  m.qt_x = (T)d->qt_x;
  m.qt_z = (T)d->qt_z;
  m.qxt = (T)d->qxt;
  m.a = a;
  m.rho_vx = med[0];
  m.rho_vz = med[1];
  m.lam = med[2];
  m.mu_txx = med[3];
  m.mu_txz = med[4];
  m.w_vx = w[0];
  m.w_vz = w[1];
  m.w_tau = w[2];
  size_t n = (size_t)d->nx * d->nz;
  for (int c = 0; c < 10; ++c) m.f[c] = fields + c * n;
  m.storage = d->storage;
  m.cq_on = (m.storage >= 1 && m.storage <= 3) || m.storage == 5;
  if (m.cq_on) {
    std::vector<float> c(n);
    for (int k = 0; k < 7; ++k) {
      for (size_t o = 0; o < n; ++o) {
        float v;
        switch (k) {
          case 0: v = (float)m.dtx / (float)m.rho_vx[o]; break;
          case 1: v = (float)m.dtz / (float)m.rho_vx[o]; break;
          case 2: v = (float)m.dtx / (float)m.rho_vz[o]; break;
          case 3: v = (float)m.dtz / (float)m.rho_vz[o]; break;
          case 4: v = (float)m.lam[o]; break;
          case 5: v = (float)m.lam[o] + 2.0f * (float)m.mu_txx[o]; break;
          default: v = (float)m.mu_txz[o]; break;
        }
        c[o] = v;
      }
      const float sc = m.storage >= 2 ? coef_scale(c) : 1.f;
      const bool alias = (k == 1 || k == 3) && (float)m.dtx == (float)m.dtz;
      const uint32_t kp = alias ? k - 1 : k;
      m.cq[k].resize(n);
      for (size_t o = 0; o < n; ++o)
        m.cq[k][o] = (T)dither_coef(m.storage, c[o], sc, (uint32_t)(o % m.nx),
                                    (uint32_t)(o / m.nx), kp);
    }
  }
  return m;
}

template <typename T>
static int run_steps(const oracle_desc* d, const T* a, const T* const* med, const T* const* w,
                     T* fields, int src_i, int src_j, int src_type, const T* wavelet, int nw,
                     const int* ri, const int* rj, const int* rt, int n_rec, int nt, int k_first,
                     int n_steps, T* gather, int stage_mask) {
  if (d->nx < 2 * (d->nbc + 3) || d->nz < 2 * (d->nbc + 3)) return -1;
  if (d->fd_order < 2 || d->fd_order % 2) return -2;
  if (d->storage && sizeof(T) != sizeof(float)) return -3;
  Model<T> m = make_model<T>(d, a, med, w, fields);
  for (int k = k_first; k < k_first + n_steps; ++k) {
    if (stage_mask & 1) stage_backup(m);
    if (stage_mask & 2) stage_inject(m, src_i, src_j, src_type, wavelet, nw, k);
    if (stage_mask & 4) {
      stage_velocity(m);
      round_fields(m, 0, 2);
    }
    if (stage_mask & 8) {
      stage_habc(m, 0);
      round_fields(m, 0, 2);
    }
    if (stage_mask & 16) {
      stage_stress(m);
      round_fields(m, 2, 5);
    }
    if (stage_mask & 32) {
      stage_habc(m, 1);
      round_fields(m, 2, 5);
    }
    if (stage_mask & 64) stage_fs(m);
    if (stage_mask & 128) stage_record(m, ri, rj, rt, n_rec, nt, k, gather);
  }
  return 0;
}

}  // namespace

extern "C" {

// Runs time steps k_first .. k_first+n_steps-1 (1-based, time_stepper.jl:108-109)
// over `fields` (10 arrays of nx*nz, order vx vz txx tzz txz + *_old), writing
// gather[(k-1) + r*nt].  stage_mask selects stages (bit order of time_step!:
// backup, inject, velocity, habc_v, stress, habc_s, free surface, record);
// 255 = the whole step.  med = {rho_vx, rho_vz, lam, mu_txx, mu_txz};
// w = {w_vx, w_vz, w_tau} in the reference's (nz, nx) layout.
int oracle_run_f32(const oracle_desc* d, const float* a, const float* const* med,
                   const float* const* w, float* fields, int src_i, int src_j, int src_type,
                   const float* wavelet, int nw, const int* ri, const int* rj, const int* rt,
                   int n_rec, int nt, int k_first, int n_steps, float* gather, int stage_mask) {
  return run_steps<float>(d, a, med, w, fields, src_i, src_j, src_type, wavelet, nw, ri, rj, rt,
                          n_rec, nt, k_first, n_steps, gather, stage_mask);
}

int oracle_run_f64(const oracle_desc* d, const double* a, const double* const* med,
                   const double* const* w, double* fields, int src_i, int src_j, int src_type,
                   const double* wavelet, int nw, const int* ri, const int* rj, const int* rt,
                   int n_rec, int nt, int k_first, int n_steps, double* gather, int stage_mask) {
  return run_steps<double>(d, a, med, w, fields, src_i, src_j, src_type, wavelet, nw, ri, rj, rt,
                           n_rec, nt, k_first, n_steps, gather, stage_mask);
}

// One block floating-point segment rounded in place (storage 4's own tests).
void oracle_round_block(float* x, int n) { round_block(x, n); }
// One value rounded as the storage holds it (for the rounding's own tests).
float oracle_round_storage(int storage, float x, float scale) {
  return round_storage(storage, x, scale);
}

int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
