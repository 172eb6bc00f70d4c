// Device-side data layout of the MI355X time-stepping path.
//
// Every field (wavefield component or medium coefficient) is a pitched 2D array,
// x fastest: element (i, j) (0-based; Julia's (i+1, j+1)) lives at
// base[j * pitch + i].  pitch is a multiple of 256 floats (1 KiB), so every
// 64-lane x 4-column tile row starts on a 1 KiB boundary, and it exceeds nx by
// at least 8, so x-halo reads of up to 8 cells past either end of a row stay
// inside the allocation (they land in zero padding).  GUARD_ROWS zero rows sit
// above row 0 and below row nz-1 for the z-halo, plus SLACK floats before and
// after the whole block.  Padding and guards are zeroed at allocation and
// never written, so stencil loads need no bounds checks.
#pragma once
#include <stdint.h>

namespace fomo {

constexpr int WAVE = 64;
constexpr int GUARD_ROWS = 18;  // >= 2 * max half-order (8) + 2: the fused z-march reads 2M rows out
constexpr int SLACK = 256;      // floats before/after a field block
constexpr int MAX_M = 8;        // fd_order <= 16
constexpr int MAX_NBC = 63;     // one wavefront spans an HABC strip's width

enum SrcType { SRC_PRESSURE = 0, SRC_FORCE_Z = 1, SRC_FORCE_X = 2 };
// Storage modes of wavefield and medium arrays (arithmetic is always FP32):
// element kinds ST_F32 / ST_BF16 / ST_F16 for both, ST_F16C = FP16
// coefficients with FP32 fields, or ST_F24 = fields rounded to a 16-bit
// significand (a 16-bit plane and a byte plane per field) with FP32
// coefficients.
enum StorageKind { ST_F32 = 0, ST_BF16 = 1, ST_F16 = 2, ST_F16C = 3, ST_F24 = 4 };
enum Coef { C_BVX_X = 0, C_BVX_Z, C_BVZ_X, C_BVZ_Z, C_LAM, C_L2M, C_MU };
enum RecType { REC_VZ = 0, REC_VX = 1, REC_P = 2 };
enum Field { F_VX = 0, F_VZ = 1, F_TXX = 2, F_TZZ = 3, F_TXZ = 4 };
enum Ramp { RAMP_VX = 0, RAMP_VZ = 1, RAMP_TAU = 2 };

// Work-partition kinds; each has a prefix-sum table over the shots of a batch.
enum Part {
  PART_STENCIL = 0,   // velocity / stress tiles (in-place two-pass kernels)
  PART_HABC_V,        // HABC strip pieces x 2 velocity fields
  PART_HABC_S,        // HABC strip pieces x 3 stress fields
  PART_FSREC,         // free-surface rows + receivers
  PART_BACKUP,        // standalone strip backup (per-stage API)
  PART_SHOT,          // one item per shot
  PART_FUSED,         // fused velocity+stress tiles (ping-pong step)
  PART_BAND,          // stress band tiles next to the HABC strips (ping-pong step)
  PART_FUSED_EDGE,    // split fused step: tiles of the frame (see ShotDev::xi0)
  PART_FUSED_INT,     // split fused step: tiles of the interior rectangle
  PART_DBL_EDGE_A,    // double step, step k: edge tiles around the interior shrunk by 2M
  PART_DBL_EDGE_B,    // double step, step k + 1: edge tiles around the interior
  PART_DBL_INT,       // double step: two-step tiles of the interior rectangle
  PART_COUNT
};

// Stress cells whose stencil reads velocity cells the HABC changes: the
// fused kernel leaves them to the band kernel, which runs after the velocity
// HABC.  Bands (0-based, half-open): left columns [0, b), right columns
// [nx-b, nx), bottom rows [nz-b, nz), top rows [0, b) without a free surface,
// with b = nbc + M + 1.
// rows per band-kernel item: few for a single shot (more, shorter items
// spread its thin bands over the GPU), more for batches (less halo overhead)
constexpr int BAND_ROWS_SINGLE = 4, BAND_ROWS_BATCH = 8;

// Everything a kernel needs to know about one shot.  Lives in device memory
// (a table indexed by shot), so a captured graph stays valid when the host
// rewrites the table between replays.
struct ShotDev {
  int nx, nz, pitch, nbc, pad, fs;
  int tz;               // rows per stencil tile
  int tiles_x, tiles_z; // stencil tiling
  int ftw;              // fused kernel: output columns per tile (64*CPT - 2*halo)
  int ftiles_x, ftiles_z;
  int band;             // band width b = nbc + M + 1
  int nb_side, nb_mid;  // band items: per side band, per bottom/top band
  int band_rows;        // rows per band item
  // Split fused step.  Interior: written columns [xi0, xi1), rows [zi0, zi1),
  // in nix x niz tiles of ftw x tzi; the cells any of them reads are at least
  // 2M + 2 cells inside the band frame that the strip kernels (HABC, band
  // stress, free surface) write, so interior tiles of a step may run while
  // the edge tiles and then those kernels run.  Edge: the rest, in ftw x tze tiles: top block rows
  // [0, zi0) and bottom block rows [zi1, nz) across the grid (e_top, e_bot
  // tiles), then the left block columns [0, xi0) and right block columns
  // [xi1, nx) of rows [zi0, zi1) (ntx_l, ntx_r tile columns of e_mid rows).
  int xi0, xi1, zi0, zi1, nix, niz, tzi, tze;
  int e_top, e_bot, e_mid, ntx_l, ntx_r;
  int n_l, n_r, n_b, n_t;  // HABC side/bottom/top edge pieces per field
  float dtx, dtz;
  // Reduced-precision storage: a stored value is x * 2^e (exact), read back as
  // stored * 2^-e.  fsc / fisc are 2^e / 2^-e of the five fields; csc the
  // 2^-e of the coefficient arrays (C_*).  All 1 for FP32 storage.
  float fsc[5], fisc[5], csc[7];
  float* f[5];          // vx vz txx tzz txz   (pointer to cell (0,0); element type per storage)
  float* fo[5];         // strip backups (*_old)
  const float* bvx_x;   // dtx / rho_vx
  const float* bvx_z;   // dtz / rho_vx
  const float* bvz_x;   // dtx / rho_vz
  const float* bvz_z;   // dtz / rho_vz
  const float* lam;
  const float* l2m;     // lam + 2 mu
  const float* mu_txz;
  const float* ramp;    // [3][ramp_len]: weight as a function of edge distance
  int ramp_len;         // nbc + 2
  float a[MAX_M];
  float qx, qz, qt_x, qt_z, qxt;
  int si, sj, stype;    // source cell (0-based) and type
  const float* wavelet;
  int nw;
  float src_scale;      // dt / rho at the source (force sources)
  const int* rec;       // [n_rec][3]: i, j (0-based), type
  int n_rec;
  float* gather;        // [nt][n_rec]
  int nt;
  int* kp;              // device step counter: step of k_off == 0 (1-based)
  // Fused step, buffer addressing: byte ranges holding the five read fields
  // (bufA), the five written fields (bufB) and the coefficient blocks (bufC),
  // each from the start of its first block.  Field c starts at
  // c * fstr_b + forg_b of its buffer; coefficient block q at q * cstr_b +
  // corg_b (blocks: dt/rho_vx, dt/rho_vz, lam, lam+2mu, mu_txz, then the
  // dtz arrays when dtz != dtx).  recF / recC: the ranges' lengths (< 2^31).
  const void* bufA;
  void* bufB;
  const void* bufC;
  uint32_t fstr_b, forg_b, cstr_b, corg_b, recF, recC;
  // F24 storage: the byte-plane element of the 16-bit element at address p
  // is at (p >> 1) + loc (one constant per wavefield allocation; 0 otherwise)
  uint64_t loc;
};

struct BatchArgs {
  const ShotDev* shots;
  const int* prefix;    // prefix-sum table of this kernel's partition kind
  const int* n_shots_p; // number of live table rows
  int k_off;            // step offset inside the captured chunk
};

}  // namespace fomo
