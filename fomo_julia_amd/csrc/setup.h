// Device-side model setup: the medium coefficients of one shot computed on
// the GPU from the unpadded vp, vs, rho arrays (x fastest, (nx_in, nz_in)).
//
// Per padded cell (i, j), with the reference's formulas in Float32 and their
// evaluation order (so the arrays are the ones the host path builds):
//   padding      edge replication: source cell (clamp(i-pad), clamp(j-pad))
//                (_pad_array, src/utils/init.jl:119-134);
//   mu, lam      rho*(vs*vs), rho*(vp*vp) - 2*mu (init.jl:141-142);
//   rho_vz       0.25*(((r(i,j) + r(i+1,j)) + r(i,j+1)) + r(i+1,j+1)), the last
//                row / column repeated (init.jl:146-149);
//   mu_txz       2 / (1/mu(i,j) + 1/mu(i,j+1)), the last column repeated
//                (init.jl:151-155);
//   coefficients dtx/rho_vx, dtz/rho_vx, dtx/rho_vz, dtz/rho_vz, lam,
//                lam + 2mu, mu_txz (IEEE divisions).
// Reduced storage quantizes each coefficient with the dither of the host
// path (same hash, same grid arithmetic in double precision, all exact).
#pragma once
#include "kernels.h"

namespace fomo {

struct SetupArgs {
  const float* vp;
  const float* vs;
  const float* rho;
  int nxi, nzi, pad, nx, nz, pitch;
  float dtx, dtz;
  int same;         // dtx == dtz: arrays 1 and 3 alias 0 and 2 (not written)
  void* out[7];     // coefficient block origins (element type per kind)
  float* out_rho;   // FP32 models of the derived-coefficient kernel: rho_vx and
  float* out_mu;    // mu_txx blocks (null otherwise)
  float sc[7];      // FP16 store scales (1 for the other kinds)
  float* partial;   // [gridDim.x][9]: max |c_k| (k < 7), min and max impedance
};

__device__ __forceinline__ void medium_at(const SetupArgs& A, int i, int j, float& rho, float& mu,
                                          float& lam) {
  const int ci = min(max(i - A.pad, 0), A.nxi - 1), cj = min(max(j - A.pad, 0), A.nzi - 1);
  const size_t o = (size_t)ci + (size_t)cj * A.nxi;
  const float r = A.rho[o], s = A.vs[o], p = A.vp[o];
  mu = r * (s * s);
  lam = r * (p * p) - 2.0f * mu;
  rho = r;
}

__device__ __forceinline__ void coefficients_at(const SetupArgs& A, int i, int j, float c[7],
                                                float& rho_vx, float& mu_txx) {
  float r, mu, lam, r2, mu2, lam2;
  medium_at(A, i, j, r, mu, lam);
  rho_vx = r;
  mu_txx = mu;
  const int i2 = min(i, A.nx - 2), j2 = min(j, A.nz - 2);
  float a, b, e, f, tmp;
  medium_at(A, i2, j2, a, tmp, tmp);
  medium_at(A, i2 + 1, j2, b, tmp, tmp);
  medium_at(A, i2, j2 + 1, e, tmp, tmp);
  medium_at(A, i2 + 1, j2 + 1, f, tmp, tmp);
  const float rho_vz = 0.25f * (((a + b) + e) + f);
  medium_at(A, i, j2, r2, mu2, lam2);
  float r3, mu3, lam3;
  medium_at(A, i, j2 + 1, r3, mu3, lam3);
  const float mu_txz = 2.0f / (1.0f / mu2 + 1.0f / mu3);
  c[0] = A.dtx / r;
  c[1] = A.dtz / r;
  c[2] = A.dtx / rho_vz;
  c[3] = A.dtz / rho_vz;
  c[4] = lam;
  c[5] = lam + 2.0f * mu;
  c[6] = mu_txz;
}

__device__ __forceinline__ uint32_t fmix32_d(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}

// Dithered rounding to a p-bit significand grid (see the host encoder).
__device__ __forceinline__ double dither_round_d(double x, int p, int emin, double u) {
  if (x == 0.0 || !isfinite(x)) return x;
  int e;
  (void)frexp(x, &e);
  const double spacing = ldexp(1.0, max(e - 1, emin) - (p - 1));
  const double a = fabs(x) / spacing;
  const double fl = floor(a);
  return copysign((a - fl > u ? fl + 1.0 : fl) * spacing, x);
}

// pass 0: per-block maxima of |c_k| and the impedance range; pass 1: write.
template <int KC>
__global__ __launch_bounds__(256) void k_setup(SetupArgs A, int pass) {
  __shared__ float red[9][256];
  float mx[7] = {0, 0, 0, 0, 0, 0, 0};
  float zmin = INFINITY, zmax = 0.f;
  const size_t n = (size_t)A.nx * A.nz;
  for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < n;
       o += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(o % A.nx), j = (int)(o / A.nx);
    float c[7], rvx, mux;
    coefficients_at(A, i, j, c, rvx, mux);
    if (pass == 0) {
#pragma unroll
      for (int k = 0; k < 7; ++k) mx[k] = fmaxf(mx[k], fabsf(c[k]));
      const double z = sqrt(fmax(0.0, (double)c[5] * (double)rvx));
      if (z > 0) zmin = fminf(zmin, (float)z);
      zmax = fmaxf(zmax, (float)z);
      continue;
    }
    const size_t off = (size_t)j * A.pitch + i;
    if constexpr (KC == ST_F32) {
      if (A.out_rho) {
        A.out_rho[off] = rvx;
        A.out_mu[off] = mux;
      }
    }
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      if (A.same && (k == 1 || k == 3)) continue;
      if constexpr (KC == ST_F32) {
        ((float*)A.out[k])[off] = c[k];
      } else {
        const uint32_t h = fmix32_d((uint32_t)i ^ fmix32_d((uint32_t)j ^ fmix32_d((uint32_t)k + 0x9E3779B9u)));
        const double u = (double)(h >> 8) * (1.0 / 16777216.0);
        const double v = (double)c[k] * (double)A.sc[k];
        const float q = (float)(KC == ST_BF16 ? dither_round_d(v, 8, -126, u) : dither_round_d(v, 11, -14, u));
        ((uint16_t*)A.out[k])[off] = (uint16_t)to_bits<KC>(q);  // exact: q is on the grid
      }
    }
  }
  if (pass != 0) return;
#pragma unroll
  for (int k = 0; k < 7; ++k) red[k][threadIdx.x] = mx[k];
  red[7][threadIdx.x] = zmin;
  red[8][threadIdx.x] = zmax;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int k = 0; k < 7; ++k) red[k][threadIdx.x] = fmaxf(red[k][threadIdx.x], red[k][threadIdx.x + s]);
      red[7][threadIdx.x] = fminf(red[7][threadIdx.x], red[7][threadIdx.x + s]);
      red[8][threadIdx.x] = fmaxf(red[8][threadIdx.x], red[8][threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 9) A.partial[(size_t)blockIdx.x * 9 + threadIdx.x] = red[threadIdx.x][0];
}

}  // namespace fomo
