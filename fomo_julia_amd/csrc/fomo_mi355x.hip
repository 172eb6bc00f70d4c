// Host side of libfomo_mi355x.so: the C ABI declared in include/fomo_mi355x.h.
//
// Objects
//   fomo_ctx        device + one HIP stream + last error text
//   fomo_model      Medium + HABCConfig + SimParams on the device: the
//                   coefficient arrays the kernels read (dt/rho per staggered
//                   point and direction, lambda, lambda+2mu, mu_txz), the HABC
//                   weight ramps, FD coefficients and scalars
//   fomo_wavefield  two state buffers of the 5 fields (the fused step reads one
//                   and writes the other) + the 5 strip backups of Wavefield
//   fomo_shot       source cell + wavelet, receivers, gather [nt][n_rec], step counter
//   fomo_batch      a device table of ShotDev rows + per-kernel work lists;
//                   single-shot calls use a one-row batch cached on the wavefield
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fomo_mi355x.h"
#include "kernels.h"
#include "setup.h"

using namespace fomo;

struct fomo_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  float* scratch = nullptr;  // this device's copy of the `scratch` array (fomo_shot_gather_noalloc)
  std::string err;
  int cpt = 1;         // columns per lane of the stencil tiles
  int tz = 0;          // rows per stencil tile; 0 = sized to fill the GPU in whole rounds
  int graph_steps = 16;
  int cus = 0;         // compute units of the device
  // 1 = fused ping-pong step, 0 = two-pass in-place step, 2 = split fused step,
  // 3 = double step (two steps per pass over the interior, split step otherwise)
  int step_mode = 1;
  hipStream_t aux = nullptr;  // the split step's strip kernels (default priority, like `stream`)
  int tze = 16;        // split step: rows per edge tile
  int int_bpc = 0;     // split step: interior blocks per CU at most (0: as many as fit)
};

struct fomo_model {
  fomo_ctx* ctx = nullptr;
  fomo_model_desc d{};
  int M = 0;
  int pitch = 0;
  size_t fstride = 0;  // floats per field block (guards + slack included)
  int storage = FOMO_STORAGE_F32;  // storage mode (FOMO_STORAGE_*)
  int ckind = ST_F32, fkind = ST_F32;  // element kinds of coefficients / fields
  int es = 4, fes = 4;  // bytes per stored coefficient / field element
  void* dcoef = nullptr;
  const float* coef[7] = {};  // bvx_x bvx_z bvz_x bvz_z lam l2m mu_txz (element type per storage)
  float csc[7] = {1, 1, 1, 1, 1, 1, 1};  // load scales 2^-e of the coefficient arrays
  int ncoef_blocks = 5;
  // FP32 with dtx == dtz and rho_vz / mu_txz the setup's averages of rho_vx /
  // mu_txx: blocks 5 and 6 hold rho_vx and mu_txx, and the fused kernel forms
  // the other coefficients itself (12 B per cell read instead of 20)
  bool compact = false;
  float zmin = 0.f, zmax = 0.f;          // P impedance sqrt((lam+2mu) rho) extremes
  float* dramp = nullptr;
  std::vector<float> rho_vx, rho_vz;  // host copies, for force-source scaling
  std::vector<float> rho_in;          // or: the unpadded rho of a device-side setup
  int nxi = 0, nzi = 0;
  float a[MAX_M] = {};
  int64_t bytes = 0;
  int pool_refs = 0;  // pool members on this model (destroy is refused while > 0)
};

struct fomo_shot;
struct fomo_batch;

struct fomo_wavefield {
  fomo_model* m = nullptr;
  void* dbuf = nullptr;
  float* buf[2][5] = {};  // state buffers (element type per storage)
  float* old[5] = {};     // strip backups (*_old) of the in-place kernels; FP32 only
  // scales of the stored state (reduced-precision storage): a zeroed state
  // takes the scales of the first shot run on it
  float fsc[5] = {1, 1, 1, 1, 1};
  bool zero = true;
  int cur = 0;            // buffer holding the current state
  // after a fused step the backup is the previous buffer's strip cells; old[]
  // is refreshed from it only when the in-place path or a download needs it
  bool backup_is_prev = false;
  // the double step's third state buffer (allocated when a double-step batch
  // is built): step k of a double step goes there outside the interior, so
  // after one the previous state's strips are in it (prev_buf2)
  void* dbuf2 = nullptr;
  float* buf2[5] = {};
  int64_t bytes2 = 0;
  bool prev_buf2 = false;
  int64_t bytes = 0;
  fomo_shot* placeholder = nullptr;  // idle shot for stage calls that take none
  // one-shot batches cached per wavefield: [0] placeholder, [1] last real shot
  fomo_batch* single[2] = {nullptr, nullptr};
  uint64_t single_id[2] = {0, 0};
  int pool_refs = 0;  // 1 while in a pool (destroy is refused)
  // F24 storage: the byte planes are a half-size image of the 16-bit planes
  // after them in dbuf; the byte of the element at p is at (p >> 1) + loc
  uint64_t loc = 0;
};

struct fomo_shot {
  fomo_model* m = nullptr;
  int si = 0, sj = 0, stype = 0;
  float* dwav = nullptr;
  int nw = 0, nt = 0, n_rec = 0;
  int* drec = nullptr;
  float* dgather = nullptr;
  int* dk = nullptr;
  float src_scale = 0.f;
  float fsc[5] = {1, 1, 1, 1, 1};  // field store scales 2^e (FP16 storage)
  std::vector<int> hrec;  // receivers (i, j, type), 0-based, as on the device
  uint64_t id = 0;  // unique per shot object, keys the cached batches
  int pool_refs = 0;  // 1 while in a pool (destroy is refused)
};

static uint64_t g_next_shot_id = 1;

// Shot tables of a batch: in-place kernels on buffer 0 / 1 (f = buffer,
// fo = strip backups) and the fused step from buffer 0 / 1 (fo = read
// buffer, f = written buffer).
// Double step from buffer p (A = buffer p, B = buffer 1 - p, C = the third
// buffer): TBL_DA* step k outside the interior shrunk by 2M, A -> C;
// TBL_DB* step k + 1 outside the interior, C -> B; TBL_DI* the two-step
// interior tiles, A -> B.  TBL_IN2: the third buffer with the strip backups
// (refreshing old[] after a double step).
enum Table {
  TBL_IN0 = 0, TBL_IN1, TBL_FUSED0, TBL_FUSED1,
  TBL_IN2, TBL_DA0, TBL_DA1, TBL_DB0, TBL_DB1, TBL_DI0, TBL_DI1,
  TBL_COUNT
};

struct fomo_batch {
  fomo_ctx* ctx = nullptr;
  int n = 0, M = 0, cpt = 2, tz = 32, graph_steps = 16, step_mode = 1;
  int step_mode_req = 1;  // the context's step mode when built (step_mode may fall back to 0)
  int tz_req = 0;  // the context's rows_per_tile when built (0 = auto)
  int storage = FOMO_STORAGE_F32;
  bool same = true;  // dtx == dtz for every member
  // every member's model holds rho_vx / mu_txx for the derived-coefficient
  // fused kernel (FP32, dtx == dtz)
  bool compact = false;
  int habc_rows = 1;  // largest member nbc + 1: LDS rows of the HABC staging arrays
  std::vector<fomo_wavefield*> ws;
  std::vector<fomo_shot*> ss;
  std::vector<ShotDev> hs;  // [TBL_COUNT][n]
  ShotDev* ds = nullptr;
  int* dn = nullptr;
  int* dprefix = nullptr;  // [PART_COUNT][n+1]
  std::vector<int> hprefix;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};  // by the parity the chunk starts from
  hipGraph_t graph[2] = {nullptr, nullptr};
  // Double step (step mode 3): two steps per launch over the interior
  // (k_fused2), two split-step chains beside it; its own captured chunks.
  bool dbl = false;
  int tzd = 0;  // rows per two-step interior tile
  // columns per lane of the double step's edge tiles: one, whose 88 VGPRs fit
  // beside k_fused2's one wave per SIMD (a two-column edge wave does not, and
  // then waits for the whole interior launch)
  int ecpt = 1;
  hipGraphExec_t gexec2[2] = {nullptr, nullptr};
  hipGraph_t graph2[2] = {nullptr, nullptr};
  int grid_fixed[PART_COUNT] = {};  // > 0: fixed grid (pools: persistent over any membership)
  hipStream_t ls = nullptr;         // stream of all its launches (null: the context's)
  // Split fused step: the edge tiles run first; then the interior tiles on
  // the batch's stream beside the strip kernels on the context's aux stream,
  // joined before the next step.  ev_edge / ev_chain order the two streams;
  // rec_in_frame: every receiver lies outside the interior rectangles, so
  // recording joins the strip kernels instead of waiting for the interior.
  bool split = false, rec_in_frame = false;
  int tzi = 0, tze = 16, int_bpc = 0;
  hipEvent_t ev_edge = nullptr, ev_chain = nullptr;
  hipEvent_t ev_dedge = nullptr;  // double step: the step-k edge tiles are done
  mutable hipStream_t cur = nullptr;  // set while the strip kernels are queued on the aux stream
  // Two halves (whole-grid step, two or more members): members [0, nh) step
  // on the batch's stream and [nh, n) on the context's aux stream.
  bool alt = false;
  int nh = 0;
  mutable int half = -1;  // launches cover this half's members (-1: all)
  int* dprefix_h[2] = {nullptr, nullptr};  // [PART_COUNT][members of the half + 1]
  int* dn_h[2] = {nullptr, nullptr};
  std::vector<int> hprefix_h[2];
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // frame store written after every step (one-shot batches); 0: none.  The
  // captured chunks hold the store's arguments, so they are keyed by its id.
  uint64_t frames_id = 0;
  FrameArgs fargs{};
};

struct fomo_frames {
  fomo_ctx* ctx = nullptr;
  FrameArgs a{};          // out, fields, sizes; slot -1
  int nx = 0, nz = 0;     // the grid size the frames were laid out for
  int64_t bytes = 0;
  uint64_t id = 0;
};
static uint64_t g_next_frames_id = 1;

static hipStream_t lstream(const fomo_batch* b) {
  return b->cur ? b->cur : b->ls ? b->ls : b->ctx->stream;
}

static thread_local std::string g_err;

// Device array of fomo_shot_gather_noalloc: one copy per GPU the library runs
// on, held until the process ends, so that entry point allocates nothing.
// Host code must not pass this name to a kernel or copy: it stands for the
// array only on the device.  fomo_ctx_create looks up the device address
// (hipGetSymbolAddress, with the context's device current) and keeps it in
// fomo_ctx::scratch; that pointer is what host code uses.  Calls on one GPU
// share the array, so calls that overlap on one GPU race.
constexpr size_t NOALLOC_SCRATCH_FLOATS = (size_t)4 << 20;  // 16 MiB
constexpr int NOALLOC_MIN_STEPS = 64;  // time steps per chunk: at least one transpose tile
constexpr int NOALLOC_MAX_RECEIVERS = (int)(NOALLOC_SCRATCH_FLOATS / NOALLOC_MIN_STEPS);
static __device__ __align__(256) float scratch[NOALLOC_SCRATCH_FLOATS];

// Held while a graph is captured and around every call that allocates, frees
// or copies device memory outside the context's stream: such a call from
// another thread during a capture (hipMalloc, hipFree, synchronous copies)
// invalidates the capture.  Recursive: destroy paths nest.
static std::recursive_mutex g_api_mu;
#define API_LOCK std::lock_guard<std::recursive_mutex> api_lock_(g_api_mu)

// A context's message is written by whichever thread fails on it (a serving
// loop creates objects in worker threads while its main thread runs pools),
// so the assignment is serialized.
static std::mutex g_err_mu;

static int fail(fomo_ctx* c, int code, const std::string& msg) {
  if (c) {
    std::lock_guard<std::mutex> lock(g_err_mu);
    c->err = msg;
  }
  g_err = msg;
  return code;
}

#define HIP_TRY(ctx, call)                                                         \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail((ctx), FOMO_ERR_HIP,                                             \
                  std::string(#call) + ": " + hipGetErrorString(e_));              \
  } while (0)

static int round_up(int v, int m) { return (v + m - 1) / m * m; }

// ------------------------------------------------------------------ layout
// The fused kernel's last tile reaches up to ftw + halo + M < 272 columns past nx.
static int pitch_for(int nx) { return round_up(nx + 272, 256); }
static size_t field_stride(int nz, int pitch) {
  return (size_t)(nz + 2 * GUARD_ROWS) * pitch + 2 * SLACK;
}
// Pointer to cell (0,0) of field block b inside a buffer of blocks of
// `stride` elements of es bytes.
static float* block_origin(void* buf, size_t stride, int b, int pitch, int es) {
  return (float*)((char*)buf + ((size_t)b * stride + SLACK + (size_t)GUARD_ROWS * pitch) * es);
}

// ------------------------------------------------------------------ storage
// Host conversions, the same rounding as the kernels (nearest, ties to even).
static uint16_t to_bf16(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
static float from_bf16(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float x;
  std::memcpy(&x, &u, 4);
  return x;
}
static uint16_t to_f16(float x) {
  const _Float16 h = (_Float16)x;
  uint16_t b;
  std::memcpy(&b, &h, 2);
  return b;
}
static float from_f16(uint16_t b) {
  _Float16 h;
  std::memcpy(&h, &b, 2);
  return (float)h;
}
// x * scale stored as element kind `kind` (ST_*) into dst[0..n)
static void encode(int kind, const float* x, size_t n, float scale, void* dst) {
  if (kind == ST_F32) {
    std::memcpy(dst, x, n * 4);
    return;
  }
  uint16_t* d = (uint16_t*)dst;
  for (size_t i = 0; i < n; ++i) d[i] = kind == ST_BF16 ? to_bf16(x[i] * scale) : to_f16(x[i] * scale);
}
static void decode(int kind, const void* src, size_t n, float scale, float* x) {
  if (kind == ST_F32) {
    std::memcpy(x, src, n * 4);
    return;
  }
  const uint16_t* d = (const uint16_t*)src;
  for (size_t i = 0; i < n; ++i) x[i] = (kind == ST_BF16 ? from_bf16(d[i]) : from_f16(d[i])) * scale;
}
// Power-of-two scale placing `bound` in [2^target, 2^(target+1)); 1 for a zero
// or non-finite bound.
static float pow2_scale(double bound, int target) {
  if (!(bound > 0) || !std::isfinite(bound)) return 1.f;
  const int e = target - (int)std::floor(std::log2(bound));
  return std::ldexp(1.0f, std::max(-126, std::min(127, e)));
}
// Reduced-precision storage supports these half-orders (fd_order 4, 8, 12).
static bool reduced_m_ok(int M) { return M == 2 || M == 4 || M == 6; }

// MurmurHash3's 32-bit finalizer (fmix32, public domain, Austin Appleby).
static uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}
// Deterministic dither in [0, 1) for cell (i, j) of coefficient array a.
static double dither_u(uint32_t i, uint32_t j, uint32_t a) {
  const uint32_t h = fmix32(i ^ fmix32(j ^ fmix32(a + 0x9E3779B9u)));
  return (double)(h >> 8) * (1.0 / 16777216.0);
}
// Dithered rounding of x to the grid of a p-bit significand format with
// smallest normal exponent emin: the grid point below |x| plus one spacing
// when the remainder (in spacings) exceeds u.  The error has mean zero over
// cells, so the medium a wave averages over keeps its coefficients.
static double dither_round(double x, int p, int emin, double u) {
  if (x == 0 || !std::isfinite(x)) return x;
  int e;
  std::frexp(x, &e);
  const double spacing = std::ldexp(1.0, std::max(e - 1, emin) - (p - 1));
  const double a = std::fabs(x) / spacing;
  const double fl = std::floor(a);
  return std::copysign((a - fl > u ? fl + 1 : fl) * spacing, x);
}
// Coefficient array k (nx x nz) times scale, dither-rounded to `kind`.
static void encode_coef(int kind, const float* x, int nx, int nz, float scale, int k, uint16_t* d) {
  for (int j = 0; j < nz; ++j)
    for (int i = 0; i < nx; ++i) {
      const size_t o = (size_t)i + (size_t)j * nx;
      const double u = dither_u((uint32_t)i, (uint32_t)j, (uint32_t)k);
      const double v = (double)x[o] * scale;
      const double q = kind == ST_BF16 ? dither_round(v, 8, -126, u) : dither_round(v, 11, -14, u);
      d[o] = kind == ST_BF16 ? to_bf16((float)q) : to_f16((float)q);  // exact: q is on the grid
    }
}
static int coef_kind(int storage) {
  return storage == FOMO_STORAGE_BF16                                    ? ST_BF16
         : storage == FOMO_STORAGE_F32 || storage == FOMO_STORAGE_F24 ? ST_F32
                                                                       : ST_F16;
}
static int field_kind(int storage) {
  return storage == FOMO_STORAGE_BF16  ? ST_BF16
         : storage == FOMO_STORAGE_F16 ? ST_F16
         : storage == FOMO_STORAGE_F24 ? ST_F24
                                       : ST_F32;
}
// F24: x rounded to a 16-bit significand (nearest, ties to even) as FP32
// bits, the same integer rounding as the kernels' f24_bits
static uint32_t f24_round_bits(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  return (u + 0x7Fu + ((u >> 8) & 1u)) & 0xFFFFFF00u;
}

// ------------------------------------------------------------------ context
extern "C" int32_t fomo_ctx_create(int32_t device, fomo_ctx** out) {
  if (!out) return fail(nullptr, FOMO_ERR_ARG, "out is NULL");
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0)
    return fail(nullptr, FOMO_ERR_HIP, std::string("no HIP device: ") + hipGetErrorString(e));
  if (device < 0 || device >= ndev) return fail(nullptr, FOMO_ERR_ARG, "device out of range");
  auto c = std::make_unique<fomo_ctx>();
  c->device = device;
  HIP_TRY(nullptr, hipSetDevice(device));
  HIP_TRY(nullptr, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIP_TRY(nullptr, hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
  HIP_TRY(nullptr, hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device));
  void* sp = nullptr;  // this device's copy of `scratch` (the device is current)
  HIP_TRY(nullptr, hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
  c->scratch = (float*)sp;
  *out = c.release();
  return FOMO_OK;
}

extern "C" int32_t fomo_ctx_destroy(fomo_ctx* c) {
  if (!c) return FOMO_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  (void)hipStreamSynchronize(c->aux);
  (void)hipStreamDestroy(c->stream);
  (void)hipStreamDestroy(c->aux);
  delete c;
  return FOMO_OK;
}

extern "C" const char* fomo_ctx_last_error(const fomo_ctx* c) {
  return c ? c->err.c_str() : g_err.c_str();
}

extern "C" int32_t fomo_ctx_synchronize(fomo_ctx* c) {
  if (!c) return fail(nullptr, FOMO_ERR_ARG, "ctx is NULL");
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return FOMO_OK;
}

extern "C" void* fomo_ctx_stream(fomo_ctx* c) { return c ? (void*)c->stream : nullptr; }

extern "C" int32_t fomo_ctx_set_step_mode(fomo_ctx* c, int32_t mode) {
  if (!c) return fail(nullptr, FOMO_ERR_ARG, "ctx is NULL");
  if (mode < 0 || mode > 3) return fail(c, FOMO_ERR_ARG, "step mode must be 0, 1, 2 or 3");
  c->step_mode = mode;
  return FOMO_OK;
}

extern "C" int32_t fomo_ctx_set_tiling(fomo_ctx* c, int32_t cpt, int32_t tz, int32_t gsteps) {
  if (!c) return fail(nullptr, FOMO_ERR_ARG, "ctx is NULL");
  if (cpt != 1 && cpt != 2 && cpt != 4) return fail(c, FOMO_ERR_ARG, "cols_per_lane must be 1, 2 or 4");
  if (tz < 0 || tz > 4096) return fail(c, FOMO_ERR_ARG, "rows_per_tile out of range (0 = auto)");
  if (gsteps != 0 && (gsteps < 2 || gsteps > 256 || gsteps % 2))
    return fail(c, FOMO_ERR_ARG, "steps_per_graph must be even, 2..256");
  c->cpt = cpt;
  c->tz = tz;
  c->graph_steps = gsteps;
  return FOMO_OK;
}

extern "C" int32_t fomo_ctx_set_split(fomo_ctx* c, int32_t edge_rows, int32_t interior_blocks_per_cu) {
  if (!c) return fail(nullptr, FOMO_ERR_ARG, "ctx is NULL");
  if (edge_rows < 1 || edge_rows > 256) return fail(c, FOMO_ERR_ARG, "edge_rows must be 1..256");
  if (interior_blocks_per_cu < 0 || interior_blocks_per_cu > 64)
    return fail(c, FOMO_ERR_ARG, "interior_blocks_per_cu must be 0..64 (0: as many as fit)");
  c->tze = edge_rows;
  c->int_bpc = interior_blocks_per_cu;
  return FOMO_OK;
}

// ------------------------------------------------------------------ model
// Extracts the weight-vs-distance ramp from an (nz, nx) matrix and checks that
// the whole matrix is that ramp of dist = min(i-1, nx-i, j-1, nz-j).
static bool ramp_from_matrix(const float* w, int nx, int nz, int nbc, float* ramp) {
  const int jm = nz / 2;  // 0-based middle row: its distance to top/bottom exceeds nbc+1
  for (int d = 0; d <= nbc + 1; ++d) ramp[d] = w[(size_t)jm + (size_t)d * nz];
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < nz; ++j) {
      int d = std::min(std::min(i, nx - 1 - i), std::min(j, nz - 1 - j));
      if (w[(size_t)j + (size_t)i * nz] != ramp[std::min(d, nbc + 1)]) return false;
    }
  return true;
}

static void ramp_closed_form(int nbc, double off, float* ramp) {
  for (int d = 0; d <= nbc + 1; ++d) {
    double v = ((double)d - off) / (double)nbc;
    ramp[d] = (float)std::min(1.0, std::max(0.0, v));
  }
}

// rho_vz and mu_txz are bit for bit the staggered averages init_medium forms
// from rho_vx and mu_txx (init.jl:146-155, last row / column repeated).
static bool averages_match(const float* rho_vx, const float* rho_vz, const float* mu_txx,
                           const float* mu_txz, int nx, int nz) {
  auto same_bits = [](float a, float b) { return std::memcmp(&a, &b, sizeof(float)) == 0; };
  auto at = [nx](const float* x, int i, int j) { return x[(size_t)i + (size_t)j * nx]; };
  for (int j = 0; j < nz; ++j) {
    const int j2 = std::min(j, nz - 2);
    for (int i = 0; i < nx; ++i) {
      const int i2 = std::min(i, nx - 2);
      const float rvz = 0.25f * (((at(rho_vx, i2, j2) + at(rho_vx, i2 + 1, j2)) + at(rho_vx, i2, j2 + 1)) +
                                 at(rho_vx, i2 + 1, j2 + 1));
      const float mxz = 2.0f / (1.0f / at(mu_txx, i, j2) + 1.0f / at(mu_txx, i, j2 + 1));
      if (!same_bits(rvz, at(rho_vz, i, j)) || !same_bits(mxz, at(mu_txz, i, j))) return false;
    }
  }
  return true;
}

extern "C" int32_t fomo_model_create(fomo_ctx* c, const fomo_model_desc* d, const float* a,
                                     const float* rho_vx, const float* rho_vz, const float* lam,
                                     const float* mu_txx, const float* mu_txz, const float* w_vx,
                                     const float* w_vz, const float* w_tau, fomo_model** out) {
  API_LOCK;
  if (!c || !d || !a || !out) return fail(c, FOMO_ERR_ARG, "NULL argument");
  if (!rho_vx || !rho_vz || !lam || !mu_txx || !mu_txz)
    return fail(c, FOMO_ERR_ARG, "medium array is NULL");
  if (d->fd_order < 2 || d->fd_order > 2 * MAX_M || d->fd_order % 2)
    return fail(c, FOMO_ERR_UNSUPPORTED, "fd_order must be even, 2..16");
  if (d->nbc < 1 || d->nbc > MAX_NBC)
    return fail(c, FOMO_ERR_UNSUPPORTED, "nbc must be in 1..63");
  const int M = d->fd_order / 2;
  const int nmin = 2 * (d->nbc + 3) + 2 * M;
  if (d->nx < nmin || d->nz < nmin)
    return fail(c, FOMO_ERR_ARG, "grid too small for nbc and fd_order");
  if (d->nx > (1 << 20) || d->nz > (1 << 20) || (int64_t)d->nx * d->nz > (int64_t)1 << 31)
    return fail(c, FOMO_ERR_ARG, "grid too large");
  if (d->storage < FOMO_STORAGE_F32 || d->storage > FOMO_STORAGE_F24)
    return fail(c, FOMO_ERR_ARG, "unknown storage type");
  if (d->storage != FOMO_STORAGE_F32 && !reduced_m_ok(M))
    return fail(c, FOMO_ERR_UNSUPPORTED, "reduced-precision storage supports fd_order 4, 8 and 12");
  const bool any_w = w_vx || w_vz || w_tau;
  if (any_w && !(w_vx && w_vz && w_tau))
    return fail(c, FOMO_ERR_ARG, "pass all three HABC weight matrices or none");

  auto m = std::make_unique<fomo_model>();
  m->ctx = c;
  m->d = *d;
  m->M = M;
  m->storage = d->storage;
  m->ckind = coef_kind(d->storage);
  m->fkind = field_kind(d->storage);
  m->es = m->ckind == ST_F32 ? 4 : 2;
  m->fes = m->fkind == ST_F32 ? 4 : 2;
  for (int l = 0; l < M; ++l) m->a[l] = a[l];
  const int nx = d->nx, nz = d->nz, rl = d->nbc + 2;
  std::vector<float> ramps(3 * rl);
  if (any_w) {
    const float* ws[3] = {w_vx, w_vz, w_tau};
    for (int r = 0; r < 3; ++r)
      if (!ramp_from_matrix(ws[r], nx, nz, d->nbc, &ramps[r * rl]))
        return fail(c, FOMO_ERR_UNSUPPORTED,
                    "HABC weights are not a function of the edge distance (init_habc ramps)");
  } else {
    ramp_closed_form(d->nbc, 0.0, &ramps[0]);
    ramp_closed_form(d->nbc, 0.5, &ramps[rl]);
    ramp_closed_form(d->nbc, 0.75, &ramps[2 * rl]);
  }

  // Derived coefficients, IEEE Float32 on the host: identical bits to the
  // per-cell expressions of velocity.jl:41-42 (dtx / rho) and stress.jl:124-125
  // (lam + 2mu; 2mu is exact).
  const size_t n = (size_t)nx * nz;
  const bool same = d->dtx == d->dtz;
  std::vector<float> h[7];
  const float* src_direct[7] = {nullptr, nullptr, nullptr, nullptr, lam, nullptr, mu_txz};
  for (int k = 0; k < 7; ++k)
    if (!src_direct[k]) h[k].resize(n);
  for (size_t o = 0; o < n; ++o) {
    h[0][o] = d->dtx / rho_vx[o];
    h[1][o] = d->dtz / rho_vx[o];
    h[2][o] = d->dtx / rho_vz[o];
    h[3][o] = d->dtz / rho_vz[o];
    h[5][o] = lam[o] + 2.0f * mu_txx[o];
  }
  const int P = pitch_for(nx);
  const size_t fs = field_stride(nz, P);
  // FP32, dtx == dtz and init_medium's averages (init.jl:146-155): the fused
  // kernel can form the coefficients from rho_vx, lam and mu_txx
  m->compact = same && m->ckind == ST_F32 && averages_match(rho_vx, rho_vz, mu_txx, mu_txz, nx, nz);
  // dtx == dtz: the z-scaled arrays alias the x-scaled ones; compact: + rho_vx, mu_txx
  const int nblocks = (same ? 5 : 7) + (m->compact ? 2 : 0);
  const int es = m->es;
  m->ncoef_blocks = nblocks;
  m->pitch = P;
  m->fstride = fs;
  // P impedance range, for the field scales of reduced-precision storage
  double zmin = INFINITY, zmax = 0;
  for (size_t o = 0; o < n; ++o) {
    const double z = std::sqrt(std::max(0.0, (double)h[5][o] * (double)rho_vx[o]));
    if (z > 0) zmin = std::min(zmin, z);
    zmax = std::max(zmax, z);
  }
  m->zmin = std::isfinite(zmin) ? (float)zmin : 1.f;
  m->zmax = zmax > 0 ? (float)zmax : 1.f;
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipMalloc(&m->dcoef, nblocks * fs * es));
  HIP_TRY(c, hipMemset(m->dcoef, 0, nblocks * fs * es));
  // block order: bvx_x, bvz_x, lam, l2m, mu_txz, [bvx_z, bvz_z]
  const int blk_of[7] = {0, same ? 0 : 5, 1, same ? 1 : 6, 2, 3, 4};
  std::vector<uint16_t> enc(m->ckind == ST_F32 ? 0 : n);
  for (int k = 0; k < 7; ++k) {
    float* dst = block_origin(m->dcoef, fs, blk_of[k], P, es);
    m->coef[k] = dst;
    const float* src = src_direct[k] ? src_direct[k] : h[k].data();
    if (m->ckind != ST_F32) {
      // FP16: largest magnitude into [2^14, 2^15), below the FP16 maximum
      // 65504 after dithering; BF16 keeps FP32's range and needs no scale
      float sc = 1.f;
      if (m->ckind == ST_F16) {
        double mx = 0;
        for (size_t o = 0; o < n; ++o) mx = std::max(mx, (double)std::fabs(src[o]));
        sc = pow2_scale(mx, 14);
      }
      m->csc[k] = 1.0f / sc;
      if (same && (k == 1 || k == 3)) {  // aliases array 0 / 2
        m->csc[k] = m->csc[k - 1];
        continue;
      }
      encode_coef(m->ckind, src, nx, nz, sc, k, enc.data());
      src = (const float*)enc.data();
    } else if (same && (k == 1 || k == 3)) {
      continue;
    }
    HIP_TRY(c, hipMemcpy2D(dst, (size_t)P * es, src, (size_t)nx * es, (size_t)nx * es, nz,
                           hipMemcpyHostToDevice));
  }
  if (m->compact) {  // blocks 5, 6: rho_vx, mu_txx
    const float* extra[2] = {rho_vx, mu_txx};
    for (int q = 0; q < 2; ++q)
      HIP_TRY(c, hipMemcpy2D(block_origin(m->dcoef, fs, 5 + q, P, es), (size_t)P * es, extra[q],
                             (size_t)nx * es, (size_t)nx * es, nz, hipMemcpyHostToDevice));
  }
  HIP_TRY(c, hipMalloc(&m->dramp, ramps.size() * sizeof(float)));
  HIP_TRY(c, hipMemcpy(m->dramp, ramps.data(), ramps.size() * sizeof(float), hipMemcpyHostToDevice));
  m->rho_vx.assign(rho_vx, rho_vx + n);
  m->rho_vz.assign(rho_vz, rho_vz + n);
  m->bytes = (int64_t)nblocks * fs * es + (int64_t)ramps.size() * sizeof(float);
  *out = m.release();
  return FOMO_OK;
}

// Device-side setup (see setup.h): the same model as fomo_model_create on the
// arrays init_medium builds from (vp, vs, rho), without building them on the host.
extern "C" int32_t fomo_model_create_from_velocity(fomo_ctx* c, const fomo_model_desc* d,
                                                   const float* a, const float* vp, const float* vs,
                                                   const float* rho, int32_t nx_in, int32_t nz_in,
                                                   fomo_model** out) {
  API_LOCK;
  if (!c || !d || !a || !out || !vp || !vs || !rho) return fail(c, FOMO_ERR_ARG, "NULL argument");
  if (d->fd_order < 2 || d->fd_order > 2 * MAX_M || d->fd_order % 2)
    return fail(c, FOMO_ERR_UNSUPPORTED, "fd_order must be even, 2..16");
  if (d->nbc < 1 || d->nbc > MAX_NBC) return fail(c, FOMO_ERR_UNSUPPORTED, "nbc must be in 1..63");
  const int M = d->fd_order / 2;
  if (nx_in < 2 || nz_in < 2 || d->pad < 0 || d->nx != nx_in + 2 * d->pad || d->nz != nz_in + 2 * d->pad)
    return fail(c, FOMO_ERR_ARG, "nx/nz must be nx_in/nz_in + 2*pad");
  const int nmin = 2 * (d->nbc + 3) + 2 * M;
  if (d->nx < nmin || d->nz < nmin) return fail(c, FOMO_ERR_ARG, "grid too small for nbc and fd_order");
  if (d->nx > (1 << 20) || d->nz > (1 << 20) || (int64_t)d->nx * d->nz > (int64_t)1 << 31)
    return fail(c, FOMO_ERR_ARG, "grid too large");
  if (d->storage < FOMO_STORAGE_F32 || d->storage > FOMO_STORAGE_F24)
    return fail(c, FOMO_ERR_ARG, "unknown storage type");
  if (d->storage != FOMO_STORAGE_F32 && !reduced_m_ok(M))
    return fail(c, FOMO_ERR_UNSUPPORTED, "reduced-precision storage supports fd_order 4, 8 and 12");
  auto m = std::make_unique<fomo_model>();
  m->ctx = c;
  m->d = *d;
  m->M = M;
  m->storage = d->storage;
  m->ckind = coef_kind(d->storage);
  m->fkind = field_kind(d->storage);
  m->es = m->ckind == ST_F32 ? 4 : 2;
  m->fes = m->fkind == ST_F32 ? 4 : 2;
  for (int l = 0; l < M; ++l) m->a[l] = a[l];
  const int nx = d->nx, nz = d->nz, rl = d->nbc + 2;
  std::vector<float> ramps(3 * rl);
  ramp_closed_form(d->nbc, 0.0, &ramps[0]);
  ramp_closed_form(d->nbc, 0.5, &ramps[rl]);
  ramp_closed_form(d->nbc, 0.75, &ramps[2 * rl]);
  const bool same = d->dtx == d->dtz;
  const int P = pitch_for(nx);
  const size_t fs = field_stride(nz, P);
  m->compact = same && m->ckind == ST_F32;  // + rho_vx, mu_txx blocks for the fused kernel
  const int nblocks = (same ? 5 : 7) + (m->compact ? 2 : 0), es = m->es;
  m->ncoef_blocks = nblocks;
  m->pitch = P;
  m->fstride = fs;
  m->nxi = nx_in;
  m->nzi = nz_in;
  const size_t nin = (size_t)nx_in * nz_in;
  m->rho_in.assign(rho, rho + nin);
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipMalloc(&m->dcoef, nblocks * fs * es));
  HIP_TRY(c, hipMemset(m->dcoef, 0, nblocks * fs * es));
  const int blk_of[7] = {0, same ? 0 : 5, 1, same ? 1 : 6, 2, 3, 4};
  for (int k = 0; k < 7; ++k) m->coef[k] = block_origin(m->dcoef, fs, blk_of[k], P, es);
  // inputs and the per-block maxima of the scan pass in one staging buffer
  const int blocks = std::max(1, std::min(4 * std::max(c->cus, 1), (int)((size_t)nx * nz / 256 + 1)));
  float* staging = nullptr;
  HIP_TRY(c, hipMalloc(&staging, (3 * nin + (size_t)blocks * 9) * sizeof(float)));
  std::unique_ptr<float, hipError_t (*)(void*)> hold(staging, hipFree);
  HIP_TRY(c, hipMemcpy(staging, vp, nin * sizeof(float), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(staging + nin, vs, nin * sizeof(float), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(staging + 2 * nin, rho, nin * sizeof(float), hipMemcpyHostToDevice));
  SetupArgs A;
  std::memset(&A, 0, sizeof(A));
  A.vp = staging;
  A.vs = staging + nin;
  A.rho = staging + 2 * nin;
  A.nxi = nx_in;
  A.nzi = nz_in;
  A.pad = d->pad;
  A.nx = nx;
  A.nz = nz;
  A.pitch = P;
  A.dtx = d->dtx;
  A.dtz = d->dtz;
  A.same = same ? 1 : 0;
  for (int k = 0; k < 7; ++k) {
    A.out[k] = const_cast<float*>(m->coef[k]);
    A.sc[k] = 1.f;
  }
  A.partial = staging + 3 * nin;
  A.out_rho = m->compact ? block_origin(m->dcoef, fs, 5, P, es) : nullptr;
  A.out_mu = m->compact ? block_origin(m->dcoef, fs, 6, P, es) : nullptr;
  // pass 0: coefficient maxima (FP16 scales) and the impedance range
  auto launch = [&](int pass) {
    if (m->ckind == ST_BF16)
      hipLaunchKernelGGL(k_setup<ST_BF16>, dim3(blocks), dim3(256), 0, c->stream, A, pass);
    else if (m->ckind == ST_F16)
      hipLaunchKernelGGL(k_setup<ST_F16>, dim3(blocks), dim3(256), 0, c->stream, A, pass);
    else
      hipLaunchKernelGGL(k_setup<ST_F32>, dim3(blocks), dim3(256), 0, c->stream, A, pass);
  };
  launch(0);
  std::vector<float> part((size_t)blocks * 9);
  HIP_TRY(c, hipMemcpyAsync(part.data(), A.partial, part.size() * sizeof(float), hipMemcpyDeviceToHost,
                            c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  float mx[7] = {0, 0, 0, 0, 0, 0, 0}, zmin = INFINITY, zmax = 0.f;
  for (int q = 0; q < blocks; ++q) {
    for (int k = 0; k < 7; ++k) mx[k] = std::max(mx[k], part[(size_t)q * 9 + k]);
    zmin = std::min(zmin, part[(size_t)q * 9 + 7]);
    zmax = std::max(zmax, part[(size_t)q * 9 + 8]);
  }
  m->zmin = std::isfinite(zmin) ? zmin : 1.f;
  m->zmax = zmax > 0 ? zmax : 1.f;
  for (int k = 0; k < 7; ++k) {
    const float sc = m->ckind == ST_F16 ? pow2_scale((double)mx[k], 14) : 1.f;
    A.sc[k] = sc;
    m->csc[k] = 1.0f / sc;
  }
  if (same) {
    m->csc[1] = m->csc[0];
    m->csc[3] = m->csc[2];
  }
  launch(1);
  HIP_TRY(c, hipGetLastError());
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, hipMalloc(&m->dramp, ramps.size() * sizeof(float)));
  HIP_TRY(c, hipMemcpy(m->dramp, ramps.data(), ramps.size() * sizeof(float), hipMemcpyHostToDevice));
  m->bytes = (int64_t)nblocks * fs * es + (int64_t)ramps.size() * sizeof(float);
  *out = m.release();
  return FOMO_OK;
}

extern "C" int32_t fomo_model_destroy(fomo_model* m) {
  API_LOCK;
  if (!m) return FOMO_OK;
  if (m->pool_refs > 0)
    return fail(m->ctx, FOMO_ERR_STATE, "model still has pool members: remove them first");
  (void)hipSetDevice(m->ctx->device);
  (void)hipStreamSynchronize(m->ctx->stream);
  (void)hipFree(m->dcoef);
  (void)hipFree(m->dramp);
  delete m;
  return FOMO_OK;
}

extern "C" int64_t fomo_model_device_bytes(const fomo_model* m) { return m ? m->bytes : 0; }

// ------------------------------------------------------------------ wavefield
static void drop_single(fomo_wavefield* w);
static int refresh_backup(fomo_wavefield* w);
// The field scales of a zeroed wavefield (FP16 storage): a unit pressure source.
static void default_field_scales(const fomo_model* m, float* fsc);

extern "C" int32_t fomo_wavefield_create(fomo_model* m, fomo_wavefield** out) {
  API_LOCK;
  if (!m || !out) return fail(m ? m->ctx : nullptr, FOMO_ERR_ARG, "NULL argument");
  fomo_ctx* c = m->ctx;
  auto w = std::make_unique<fomo_wavefield>();
  w->m = m;
  // two state buffers + (FP32 only: the two-pass path's strip backups);
  // F24: + the byte-plane image of the two buffers' 16-bit planes
  const bool f32 = m->storage == FOMO_STORAGE_F32, f24 = m->fkind == ST_F24;
  const size_t hi_bytes = (f32 ? 15 : 10) * m->fstride * m->fes;
  const size_t bytes = hi_bytes + (f24 ? hi_bytes / 2 : 0);
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipMalloc(&w->dbuf, bytes));
  // zeroed on the context's stream: single-shot runs and batches use that
  // stream, and a pool orders its stream after it when the member joins
  HIP_TRY(c, hipMemsetAsync(w->dbuf, 0, bytes, c->stream));
  for (int k = 0; k < 5; ++k) {
    w->buf[0][k] = block_origin(w->dbuf, m->fstride, k, m->pitch, m->fes);
    w->buf[1][k] = block_origin(w->dbuf, m->fstride, 5 + k, m->pitch, m->fes);
    w->old[k] = f32 ? block_origin(w->dbuf, m->fstride, 10 + k, m->pitch, m->fes) : nullptr;
  }
  if (f24) w->loc = ((uint64_t)w->dbuf + hi_bytes) - ((uint64_t)w->dbuf >> 1);
  w->bytes = (int64_t)bytes;
  default_field_scales(m, w->fsc);
  *out = w.release();
  return FOMO_OK;
}

extern "C" int32_t fomo_wavefield_destroy(fomo_wavefield* w) {
  API_LOCK;
  if (!w) return FOMO_OK;
  if (w->pool_refs > 0)
    return fail(w->m->ctx, FOMO_ERR_STATE, "wavefield is in a pool: remove its shot first");
  (void)hipSetDevice(w->m->ctx->device);
  (void)hipStreamSynchronize(w->m->ctx->stream);
  drop_single(w);
  (void)hipFree(w->dbuf);
  if (w->dbuf2) (void)hipFree(w->dbuf2);
  delete w;
  return FOMO_OK;
}

extern "C" int32_t fomo_wavefield_reset(fomo_wavefield* w) {
  if (!w) return fail(nullptr, FOMO_ERR_ARG, "NULL wavefield");
  fomo_ctx* c = w->m->ctx;
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipMemsetAsync(w->dbuf, 0, (size_t)w->bytes, c->stream));
  w->cur = 0;
  w->backup_is_prev = false;
  w->prev_buf2 = false;
  w->zero = true;
  return FOMO_OK;
}

extern "C" int64_t fomo_wavefield_device_bytes(const fomo_wavefield* w) {
  return w ? w->bytes + w->bytes2 : 0;
}

static float* field_ptr(fomo_wavefield* w, int field) {
  return field < 5 ? w->buf[w->cur][field] : w->old[field - 5];
}
// F24: the byte-plane element of the 16-bit element at p
static void* f24_plane(const fomo_wavefield* w, const void* p) {
  return (void*)(((uint64_t)p >> 1) + w->loc);
}

extern "C" int32_t fomo_wavefield_download(fomo_wavefield* w, int32_t field, float* host) {
  if (!w || !host) return fail(w ? w->m->ctx : nullptr, FOMO_ERR_ARG, "NULL argument");
  fomo_ctx* c = w->m->ctx;
  if (field < 0 || field > 9) return fail(c, FOMO_ERR_ARG, "field id out of range");
  const int st = w->m->storage;
  if (field >= 5 && st != FOMO_STORAGE_F32)
    return fail(c, FOMO_ERR_UNSUPPORTED, "*_old fields exist only with FP32 storage");
  if (field >= 5) {
    int rc = refresh_backup(w);
    if (rc) return rc;
  }
  const int nx = w->m->d.nx, nz = w->m->d.nz, P = w->m->pitch, es = w->m->fes;
  const int fk = w->m->fkind;
  HIP_TRY(c, hipSetDevice(c->device));
  if (fk == ST_F32) {
    HIP_TRY(c, hipMemcpy2DAsync(host, nx * sizeof(float), field_ptr(w, field), P * sizeof(float),
                                nx * sizeof(float), nz, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return FOMO_OK;
  }
  std::vector<uint16_t> tmp((size_t)nx * nz);
  HIP_TRY(c, hipMemcpy2DAsync(tmp.data(), (size_t)nx * es, field_ptr(w, field), (size_t)P * es,
                              (size_t)nx * es, nz, hipMemcpyDeviceToHost, c->stream));
  if (fk == ST_F24) {  // + the byte plane: value bits = high << 16 | byte << 8
    std::vector<uint8_t> lo((size_t)nx * nz);
    HIP_TRY(c, hipMemcpy2DAsync(lo.data(), (size_t)nx, f24_plane(w, field_ptr(w, field)), (size_t)P,
                                (size_t)nx, nz, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < tmp.size(); ++i) {
      const uint32_t b = ((uint32_t)tmp[i] << 16) | ((uint32_t)lo[i] << 8);
      std::memcpy(&host[i], &b, 4);
    }
    return FOMO_OK;
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  decode(fk, tmp.data(), tmp.size(), 1.0f / w->fsc[field], host);
  return FOMO_OK;
}

extern "C" int32_t fomo_wavefield_upload(fomo_wavefield* w, int32_t field, const float* host) {
  if (!w || !host) return fail(w ? w->m->ctx : nullptr, FOMO_ERR_ARG, "NULL argument");
  fomo_ctx* c = w->m->ctx;
  if (field < 0 || field > 9) return fail(c, FOMO_ERR_ARG, "field id out of range");
  const int st = w->m->storage;
  if (field >= 5 && st != FOMO_STORAGE_F32)
    return fail(c, FOMO_ERR_UNSUPPORTED, "*_old fields exist only with FP32 storage");
  if (field >= 5) {
    int rc = refresh_backup(w);
    if (rc) return rc;
  }
  const int nx = w->m->d.nx, nz = w->m->d.nz, P = w->m->pitch, es = w->m->fes;
  const int fk = w->m->fkind;
  HIP_TRY(c, hipSetDevice(c->device));
  w->zero = false;
  if (fk == ST_F32) {
    HIP_TRY(c, hipMemcpy2DAsync(field_ptr(w, field), P * sizeof(float), host, nx * sizeof(float),
                                nx * sizeof(float), nz, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return FOMO_OK;
  }
  std::vector<uint16_t> tmp((size_t)nx * nz);
  if (fk == ST_F24) {  // each value rounded to F24, split into its two planes
    std::vector<uint8_t> lo(tmp.size());
    for (size_t i = 0; i < tmp.size(); ++i) {
      const uint32_t b = f24_round_bits(host[i]);
      tmp[i] = (uint16_t)(b >> 16);
      lo[i] = (uint8_t)(b >> 8);
    }
    HIP_TRY(c, hipMemcpy2DAsync(field_ptr(w, field), (size_t)P * es, tmp.data(), (size_t)nx * es,
                                (size_t)nx * es, nz, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpy2DAsync(f24_plane(w, field_ptr(w, field)), (size_t)P, lo.data(), (size_t)nx,
                                (size_t)nx, nz, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return FOMO_OK;
  }
  encode(fk, host, tmp.size(), w->fsc[field], tmp.data());
  HIP_TRY(c, hipMemcpy2DAsync(field_ptr(w, field), (size_t)P * es, tmp.data(), (size_t)nx * es,
                              (size_t)nx * es, nz, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return FOMO_OK;
}

static void default_field_scales(const fomo_model* m, float* fsc) {
  for (int k = 0; k < 5; ++k) fsc[k] = 1.f;
  if (m->fkind != ST_F16) return;
  const float ss = pow2_scale(1.0, 13), sv = pow2_scale(1.0 / m->zmin, 13);
  fsc[F_VX] = fsc[F_VZ] = sv;
  fsc[F_TXX] = fsc[F_TZZ] = fsc[F_TXZ] = ss;
}

// A wavefield holding a nonzero state can only continue at that state's
// scales; a zeroed one takes the shot's.
static int adopt_scales(fomo_wavefield* w, const fomo_shot* s) {
  if (w->m->fkind != ST_F16) return FOMO_OK;
  if (!w->zero && std::memcmp(w->fsc, s->fsc, sizeof(w->fsc)) != 0)
    return fail(w->m->ctx, FOMO_ERR_STATE,
                "wavefield holds a state stored at another shot's FP16 scales: reset it first");
  std::memcpy(w->fsc, s->fsc, sizeof(w->fsc));
  w->zero = false;
  return FOMO_OK;
}

// rho_vx / rho_vz of padded cell (i, j) (0-based): from the host copies, or
// for a device-side setup from the unpadded rho with the setup's formulas.
static float model_rho(const fomo_model* m, bool vz, int i, int j) {
  const int nx = m->d.nx;
  if (!m->rho_vx.empty()) return (vz ? m->rho_vz : m->rho_vx)[(size_t)i + (size_t)j * nx];
  auto r = [&](int ii, int jj) {
    const int ci = std::min(std::max(ii - m->d.pad, 0), m->nxi - 1);
    const int cj = std::min(std::max(jj - m->d.pad, 0), m->nzi - 1);
    return m->rho_in[(size_t)ci + (size_t)cj * m->nxi];
  };
  if (!vz) return r(i, j);
  const int i2 = std::min(i, nx - 2), j2 = std::min(j, m->d.nz - 2);
  return 0.25f * (((r(i2, j2) + r(i2 + 1, j2)) + r(i2, j2 + 1)) + r(i2 + 1, j2 + 1));
}

// ------------------------------------------------------------------ shot
extern "C" int32_t fomo_shot_create(fomo_model* m, int32_t src_i, int32_t src_j, int32_t src_type,
                                    const float* wavelet, int32_t nw, int32_t nt,
                                    const int32_t* rec_i, const int32_t* rec_j,
                                    const int32_t* rec_type, int32_t n_rec, fomo_shot** out) {
  API_LOCK;
  if (!m || !out) return fail(m ? m->ctx : nullptr, FOMO_ERR_ARG, "NULL argument");
  fomo_ctx* c = m->ctx;
  const int nx = m->d.nx, nz = m->d.nz;
  if (src_i < 1 || src_i > nx || src_j < 1 || src_j > nz)
    return fail(c, FOMO_ERR_ARG, "source index outside the padded grid");
  if (src_type < FOMO_SRC_PRESSURE || src_type > FOMO_SRC_FORCE_X)
    return fail(c, FOMO_ERR_ARG, "unknown source type");
  if (nw < 0 || (nw > 0 && !wavelet)) return fail(c, FOMO_ERR_ARG, "bad wavelet");
  if (nt < 1) return fail(c, FOMO_ERR_ARG, "nt must be >= 1");
  if (n_rec < 0 || (n_rec > 0 && (!rec_i || !rec_j))) return fail(c, FOMO_ERR_ARG, "bad receivers");
  if ((int64_t)nt * std::max(n_rec, 1) > ((int64_t)1 << 31))
    return fail(c, FOMO_ERR_ARG, "gather too large");
  std::vector<int> rec(3 * (size_t)std::max(n_rec, 1), 0);
  for (int r = 0; r < n_rec; ++r) {
    const int ti = rec_type ? rec_type[r] : FOMO_REC_VZ;
    if (rec_i[r] < 1 || rec_i[r] > nx || rec_j[r] < 1 || rec_j[r] > nz)
      return fail(c, FOMO_ERR_ARG, "receiver index outside the padded grid");
    if (ti < FOMO_REC_VZ || ti > FOMO_REC_P) return fail(c, FOMO_ERR_ARG, "unknown receiver type");
    rec[3 * r] = rec_i[r] - 1;
    rec[3 * r + 1] = rec_j[r] - 1;
    rec[3 * r + 2] = ti;
  }
  auto s = std::make_unique<fomo_shot>();
  s->m = m;
  s->si = src_i - 1;
  s->sj = src_j - 1;
  s->stype = src_type;
  s->nw = nw;
  s->nt = nt;
  s->n_rec = n_rec;
  s->id = __atomic_fetch_add(&g_next_shot_id, 1, __ATOMIC_RELAXED);
  if (src_type == FOMO_SRC_FORCE_Z) s->src_scale = m->d.dt / model_rho(m, true, src_i - 1, src_j - 1);
  if (src_type == FOMO_SRC_FORCE_X) s->src_scale = m->d.dt / model_rho(m, false, src_i - 1, src_j - 1);
  if (m->fkind == ST_F16) {
    // Bound of the injected total: sum |w| (into stress) or sum |w| dt / rho
    // (into velocity); the other fields by the plane-wave relation
    // stress = impedance * velocity with the impedance extreme that makes it
    // larger.  Each bound lands in [2^13, 2^14): 4x below the FP16 maximum.
    double inj = 0;
    for (int k = 0; k < nw; ++k) inj += std::fabs((double)wavelet[k]);
    if (!(inj > 0)) inj = 1;
    double sb, vb;
    if (src_type == FOMO_SRC_PRESSURE) {
      sb = inj;
      vb = inj / m->zmin;
    } else {
      vb = inj * s->src_scale;
      sb = vb * m->zmax;
    }
    const float sv = pow2_scale(vb, 13), ss = pow2_scale(sb, 13);
    s->fsc[F_VX] = sv;
    s->fsc[F_VZ] = sv;
    s->fsc[F_TXX] = ss;
    s->fsc[F_TZZ] = ss;
    s->fsc[F_TXZ] = ss;
  }
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipMalloc(&s->dwav, std::max(nw, 1) * sizeof(float)));
  if (nw > 0) HIP_TRY(c, hipMemcpy(s->dwav, wavelet, nw * sizeof(float), hipMemcpyHostToDevice));
  s->hrec = rec;
  HIP_TRY(c, hipMalloc(&s->drec, rec.size() * sizeof(int)));
  HIP_TRY(c, hipMemcpy(s->drec, rec.data(), rec.size() * sizeof(int), hipMemcpyHostToDevice));
  const size_t gbytes = (size_t)nt * std::max(n_rec, 1) * sizeof(float);
  HIP_TRY(c, hipMalloc(&s->dgather, gbytes));
  HIP_TRY(c, hipMemsetAsync(s->dgather, 0, gbytes, c->stream));
  HIP_TRY(c, hipMalloc(&s->dk, sizeof(int)));
  HIP_TRY(c, hipMemsetAsync(s->dk, 0, sizeof(int), c->stream));
  *out = s.release();
  return FOMO_OK;
}

extern "C" int32_t fomo_shot_field_scales(const fomo_shot* s, float* out) {
  if (!s || !out) return fail(s ? s->m->ctx : nullptr, FOMO_ERR_ARG, "NULL argument");
  for (int k = 0; k < 5; ++k) out[k] = s->fsc[k];
  return FOMO_OK;
}

extern "C" int32_t fomo_shot_destroy(fomo_shot* s) {
  API_LOCK;
  if (!s) return FOMO_OK;
  if (s->pool_refs > 0) return fail(s->m->ctx, FOMO_ERR_STATE, "shot is in a pool: remove it first");
  (void)hipSetDevice(s->m->ctx->device);
  (void)hipStreamSynchronize(s->m->ctx->stream);
  (void)hipFree(s->dwav);
  (void)hipFree(s->drec);
  (void)hipFree(s->dgather);
  (void)hipFree(s->dk);
  delete s;
  return FOMO_OK;
}

extern "C" int32_t fomo_shot_reset(fomo_shot* s) {
  if (!s) return fail(nullptr, FOMO_ERR_ARG, "NULL shot");
  fomo_ctx* c = s->m->ctx;
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipMemsetAsync(s->dgather, 0, (size_t)s->nt * std::max(s->n_rec, 1) * sizeof(float),
                            c->stream));
  return FOMO_OK;
}

extern "C" int32_t fomo_shot_gather(fomo_shot* s, float* out) {
  if (!s || !out) return fail(s ? s->m->ctx : nullptr, FOMO_ERR_ARG, "NULL argument");
  fomo_ctx* c = s->m->ctx;
  if (s->n_rec == 0 || s->nt == 0) return FOMO_OK;
  // device [k][r] -> Julia [nt x n_rec] column-major (out[k + r*nt]), transposed
  // on the device into a staging buffer, then one copy
  const size_t n = (size_t)s->nt * s->n_rec;
  float* tposed = nullptr;
  {
    API_LOCK;  // allocations must not interleave with a graph capture on another thread
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMalloc(&tposed, n * sizeof(float)));
  }
  const dim3 g((unsigned)((s->n_rec + 63) / 64), (unsigned)((s->nt + 63) / 64));
  hipLaunchKernelGGL(k_transpose, g, dim3(256), 0, c->stream, s->dgather, tposed, s->nt, s->n_rec);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, tposed, n * sizeof(float), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  {
    API_LOCK;
    (void)hipFree(tposed);
  }
  if (e != hipSuccess) return fail(c, FOMO_ERR_HIP, std::string("shot_gather: ") + hipGetErrorString(e));
  return FOMO_OK;
}

// The same [nt x n_rec] download through the library's fixed device array:
// chunks of kc time steps are transposed into `scratch` ([r][k - k0], n_rec
// rows of kc floats) and copied with one strided copy each to out[k + r*nt].
// Stream order keeps a chunk's transpose behind the previous chunk's copy.
extern "C" int32_t fomo_shot_gather_noalloc(fomo_shot* s, float* out) {
  if (!s || !out) return fail(s ? s->m->ctx : nullptr, FOMO_ERR_ARG, "NULL argument");
  fomo_ctx* c = s->m->ctx;
  if (s->n_rec == 0 || s->nt == 0) return FOMO_OK;
  if (s->n_rec > NOALLOC_MAX_RECEIVERS)
    return fail(c, FOMO_ERR_ARG, "fomo_shot_gather_noalloc: " + std::to_string(s->n_rec) +
                                     " receivers exceed its maximum of " +
                                     std::to_string(NOALLOC_MAX_RECEIVERS) +
                                     " (64 steps of a 16 MiB device array); use fomo_shot_gather");
  if (!c->scratch) return fail(c, FOMO_ERR_STATE, "context has no device scratch array");
  HIP_TRY(c, hipSetDevice(c->device));
  const int nt = s->nt, nr = s->n_rec;
  const int kc = (int)std::min<size_t>((size_t)nt, NOALLOC_SCRATCH_FLOATS / (size_t)nr);
  for (int k0 = 0; k0 < nt; k0 += kc) {
    const int n = std::min(kc, nt - k0);
    const dim3 g((unsigned)((nr + 63) / 64), (unsigned)((n + 63) / 64));
    hipLaunchKernelGGL(k_transpose, g, dim3(256), 0, c->stream, s->dgather + (size_t)k0 * nr,
                       c->scratch, n, nr);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipMemcpy2DAsync(out + k0, (size_t)nt * sizeof(float), c->scratch,
                                (size_t)n * sizeof(float), (size_t)n * sizeof(float), (size_t)nr,
                                hipMemcpyDeviceToHost, c->stream));
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return FOMO_OK;
}

// ------------------------------------------------------------------ batches
static ShotDev make_shotdev(const fomo_model* m, const fomo_wavefield* w, const fomo_shot* s,
                            int cpt, int tz, int table, int band_rows, int tzi = 0, int tze = 16,
                            int tzd = 0) {
  ShotDev sd;
  std::memset(&sd, 0, sizeof(sd));
  const fomo_model_desc& d = m->d;
  sd.nx = d.nx;
  sd.nz = d.nz;
  sd.pitch = m->pitch;
  sd.nbc = d.nbc;
  sd.pad = d.pad;
  sd.fs = d.free_surface ? 1 : 0;
  sd.tz = tz;
  sd.tiles_x = (d.nx + 64 * cpt - 1) / (64 * cpt);
  sd.tiles_z = (d.nz + tz - 1) / tz;
  // HABC pieces (see k_habc): side-edge rows [fs ? n : n+2, nz-n-2), bottom/top
  // edge columns [n+2, nx-n-2); the rest belongs to the corner pieces.
  const int n = d.nbc;
  const int side_rows = (d.nz - n - 2) - (sd.fs ? n : n + 2);
  const int edge_cols = (d.nx - n - 2) - (n + 2);
  sd.n_l = (side_rows + 63) / 64;
  sd.n_r = sd.n_l;
  sd.n_b = (edge_cols + 63) / 64;
  sd.n_t = sd.fs ? 0 : sd.n_b;
  sd.dtx = d.dtx;
  sd.dtz = d.dtz;
  for (int k = 0; k < 5; ++k) {
    sd.fsc[k] = s->fsc[k];
    sd.fisc[k] = 1.0f / s->fsc[k];
  }
  for (int k = 0; k < 7; ++k) sd.csc[k] = m->csc[k];
  for (int k = 0; k < 5; ++k) {
    switch (table) {
      case TBL_IN0: sd.f[k] = w->buf[0][k]; sd.fo[k] = w->old[k]; break;
      case TBL_IN1: sd.f[k] = w->buf[1][k]; sd.fo[k] = w->old[k]; break;
      case TBL_FUSED0: sd.f[k] = w->buf[1][k]; sd.fo[k] = w->buf[0][k]; break;
      case TBL_IN2: sd.f[k] = w->buf2[k]; sd.fo[k] = w->old[k]; break;
      case TBL_DA0: sd.f[k] = w->buf2[k]; sd.fo[k] = w->buf[0][k]; break;
      case TBL_DA1: sd.f[k] = w->buf2[k]; sd.fo[k] = w->buf[1][k]; break;
      case TBL_DB0: sd.f[k] = w->buf[1][k]; sd.fo[k] = w->buf2[k]; break;
      case TBL_DB1: sd.f[k] = w->buf[0][k]; sd.fo[k] = w->buf2[k]; break;
      case TBL_DI0: sd.f[k] = w->buf[1][k]; sd.fo[k] = w->buf[0][k]; break;
      default: sd.f[k] = w->buf[0][k]; sd.fo[k] = w->buf[1][k]; break;  // TBL_FUSED1, TBL_DI1
    }
  }
  // fused kernel tiling: lanes cover 64*cpt columns, the outer halo lanes
  // (M rounded up to cpt) are computed but not written
  const int M = m->M;
  const int hl = (M + cpt - 1) / cpt * cpt;
  sd.ftw = 64 * cpt - 2 * hl;
  sd.ftiles_x = (d.nx + sd.ftw - 1) / sd.ftw;
  sd.ftiles_z = (d.nz + tz - 1) / tz;
  sd.band = d.nbc + M + 1;
  sd.band_rows = band_rows;
  sd.nb_side = ((sd.band + 63) / 64) * ((d.nz + band_rows - 1) / band_rows);
  sd.nb_mid = ((d.nx - 2 * sd.band + 63) / 64) * ((sd.band + band_rows - 1) / band_rows);
  // split step: the interior's written cells sit 2M + 2 cells inside the
  // band frame, so the cells its tiles read are outside the frame that the
  // strip kernels write (and the strip kernels read nothing it writes)
  {
    const int mg = sd.band + 2 * M + 2;
    int xi0 = (mg + 1) & ~1, xi1 = (d.nx - mg) & ~1, zi0 = mg, zi1 = d.nz - mg;
    // double step, first step: the edge tiles cover all but the interior
    // shrunk by 2M, so that the second step's edge tiles (all but the
    // interior) find every cell they read
    if ((table == TBL_DA0 || table == TBL_DA1) && xi1 > xi0 && zi1 > zi0) {
      xi0 += 2 * M;
      xi1 -= 2 * M;
      zi0 += 2 * M;
      zi1 -= 2 * M;
    }
    if (xi1 <= xi0 || zi1 <= zi0) {  // no interior: the edge tiles cover the grid
      xi0 = xi1 = 0;
      zi0 = zi1 = d.nz;
    }
    tzi = tzi > 0 ? tzi : tz;
    tze = std::max(1, tze);
    sd.xi0 = xi0;
    sd.xi1 = xi1;
    sd.zi0 = zi0;
    sd.zi1 = zi1;
    sd.tzi = tzi;
    sd.tze = tze;
    sd.nix = xi1 > xi0 ? (xi1 - xi0 + sd.ftw - 1) / sd.ftw : 0;
    sd.niz = zi1 > zi0 ? (zi1 - zi0 + tzi - 1) / tzi : 0;
    sd.e_top = sd.ftiles_x * ((zi0 + tze - 1) / tze);
    sd.e_bot = sd.ftiles_x * ((d.nz - zi1 + tze - 1) / tze);
    sd.e_mid = (zi1 - zi0 + tze - 1) / tze;
    sd.ntx_l = (xi0 + sd.ftw - 1) / sd.ftw;
    sd.ntx_r = (d.nx - xi1 + sd.ftw - 1) / sd.ftw;
    if (table == TBL_DI0 || table == TBL_DI1) {
      // two-step interior tiles (k_fused2: two columns per lane, 3M-column
      // halo per side) of tzd rows over the same interior rectangle
      sd.ftw = 128 - 2 * dbl_halo(M);
      sd.tzi = std::max(1, tzd);
      sd.nix = xi1 > xi0 ? (xi1 - xi0 + sd.ftw - 1) / sd.ftw : 0;
      sd.niz = zi1 > zi0 ? (zi1 - zi0 + sd.tzi - 1) / sd.tzi : 0;
    }
  }
  sd.bvx_x = m->coef[0];
  sd.bvx_z = m->coef[1];
  sd.bvz_x = m->coef[2];
  sd.bvz_z = m->coef[3];
  sd.lam = m->coef[4];
  sd.l2m = m->coef[5];
  sd.mu_txz = m->coef[6];
  sd.ramp = m->dramp;
  sd.ramp_len = d.nbc + 2;
  for (int l = 0; l < MAX_M; ++l) sd.a[l] = m->a[l];
  sd.qx = d.qx;
  sd.qz = d.qz;
  sd.qt_x = d.qt_x;
  sd.qt_z = d.qt_z;
  sd.qxt = d.qxt;
  sd.si = s->si;
  sd.sj = s->sj;
  sd.stype = s->stype;
  sd.wavelet = s->dwav;
  sd.nw = s->nw;
  sd.src_scale = s->src_scale;
  sd.rec = s->drec;
  sd.n_rec = s->n_rec;
  sd.gather = s->dgather;
  sd.nt = s->nt;
  sd.kp = s->dk;
  // fused step, buffer addressing (see ShotDev)
  const size_t org = (size_t)GUARD_ROWS * m->pitch + SLACK;
  const size_t fstr = m->fstride * m->fes, cstr = m->fstride * m->es;
  sd.fstr_b = (uint32_t)fstr;
  sd.forg_b = (uint32_t)(org * m->fes);
  sd.cstr_b = (uint32_t)cstr;
  sd.corg_b = (uint32_t)(org * m->es);
  sd.recF = (uint32_t)std::min<size_t>(5 * fstr, 0x7FFFFFFF);
  sd.recC = (uint32_t)std::min<size_t>((size_t)m->ncoef_blocks * cstr, 0x7FFFFFFF);
  sd.bufA = sd.fo[0] ? (const char*)sd.fo[0] - sd.forg_b : nullptr;
  sd.bufB = sd.f[0] ? (char*)sd.f[0] - sd.forg_b : nullptr;
  sd.bufC = m->dcoef;
  sd.loc = w->loc;
  return sd;
}

static int part_items(const ShotDev& sd, int part) {
  const int per_field = sd.n_l + sd.n_r + sd.n_b + sd.n_t + (sd.fs ? 2 : 4);
  switch (part) {
    case PART_STENCIL: return sd.tiles_x * sd.tiles_z;
    case PART_HABC_V: return 2 * per_field;
    case PART_HABC_S: return 3 * per_field;
    case PART_FSREC: return (6 * sd.nx + sd.n_rec + 63) / 64;
    case PART_BACKUP: return (int)(((int64_t)sd.nx * sd.nz + 63) / 64);
    case PART_SHOT: return 1;
    case PART_FUSED: return sd.ftiles_x * sd.ftiles_z;
    case PART_BAND: return 2 * sd.nb_side + (sd.fs ? 1 : 2) * sd.nb_mid;
    case PART_FUSED_EDGE: return sd.e_top + sd.e_bot + (sd.ntx_l + sd.ntx_r) * sd.e_mid;
    case PART_FUSED_INT: return sd.nix * sd.niz;
    // the double step's parts, counted on its TBL_DA / TBL_DB / TBL_DI rows
    case PART_DBL_EDGE_A:
    case PART_DBL_EDGE_B: return sd.e_top + sd.e_bot + (sd.ntx_l + sd.ntx_r) * sd.e_mid;
    case PART_DBL_INT: return sd.nix * sd.niz;
  }
  return 0;
}

constexpr int WPB = 4;  // wavefronts per block of the stencil kernels

// Resident wavefronts of the stencil kernels for (M, cpt) on the whole device.
template <int K, int M, int CPT>
static int stencil_slots_mc(int cus, int mode, bool compact) {
  int bv = 0, bs = 0;
  if (mode == 1 || K != ST_F32) {
    hipError_t e;
    if constexpr (K == ST_F32 || K == ST_F24)
      e = compact ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&bv, k_fused<M, CPT, WPB, K, true, true>,
                                                                 64 * WPB, 0)
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&bv, k_fused<M, CPT, WPB, K, true, false>,
                                                                 64 * WPB, 0);
    else
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&bv, k_fused<M, CPT, WPB, K, true, false>, 64 * WPB, 0);
    if (e != hipSuccess) bv = 1;
    bs = bv;
  } else if constexpr (K == ST_F32) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bv, k_velocity<M, CPT, WPB, true>, 64 * WPB, 0) !=
        hipSuccess)
      bv = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bs, k_stress<M, CPT, WPB>, 64 * WPB, 0) != hipSuccess)
      bs = 1;
  }
  return std::max(1, std::min(bv, bs)) * WPB * std::max(cus, 1);
}
template <int K, int M>
static int stencil_slots_m(int cpt, int cus, int mode, bool compact = false) {
  return cpt == 1   ? stencil_slots_mc<K, M, 1>(cus, mode, compact)
         : cpt == 2 ? stencil_slots_mc<K, M, 2>(cus, mode, compact)
                    : stencil_slots_mc<K, M, 4>(cus, mode, compact);
}
template <int K>
static int stencil_slots_reduced(int M, int cpt, int cus, bool compact = false) {
  switch (M) {
    case 2: return stencil_slots_m<K, 2>(cpt, cus, 1, compact);
    case 4: return stencil_slots_m<K, 4>(cpt, cus, 1, compact);
    default: return stencil_slots_m<K, 6>(cpt, cus, 1, compact);
  }
}
static int stencil_slots(int storage, int M, int cpt, int cus, int mode, bool compact = false) {
  if (storage == FOMO_STORAGE_BF16) return stencil_slots_reduced<ST_BF16>(M, cpt, cus);
  if (storage == FOMO_STORAGE_F16) return stencil_slots_reduced<ST_F16>(M, cpt, cus);
  if (storage == FOMO_STORAGE_F16C) return stencil_slots_reduced<ST_F16C>(M, cpt, cus);
  if (storage == FOMO_STORAGE_F24) return stencil_slots_reduced<ST_F24>(M, cpt, cus, compact);
  switch (M) {
    case 1: return stencil_slots_m<ST_F32, 1>(cpt, cus, mode, compact);
    case 2: return stencil_slots_m<ST_F32, 2>(cpt, cus, mode, compact);
    case 3: return stencil_slots_m<ST_F32, 3>(cpt, cus, mode, compact);
    case 4: return stencil_slots_m<ST_F32, 4>(cpt, cus, mode, compact);
    case 5: return stencil_slots_m<ST_F32, 5>(cpt, cus, mode, compact);
    case 6: return stencil_slots_m<ST_F32, 6>(cpt, cus, mode, compact);
    case 7: return stencil_slots_m<ST_F32, 7>(cpt, cus, mode, compact);
    default: return stencil_slots_m<ST_F32, 8>(cpt, cus, mode, compact);
  }
}

// Rows per stencil tile.  A launch takes `rounds` turns of the resident
// wavefront slots, each as long as one full tile's march, so the choice
// maximises stored rows / (rounds * slots * march rows per tile); a shot's
// short last tile counts only its own rows.  Ties go to the fewest marched
// rows in total (items * march), which the bandwidth pays for.  `block` > 0:
// the fused kernel marches tz + 2M rows in whole blocks of `block` rows, and
// heights up to 256 are tried (the 64-shot order-12 batch, 442 rows per shot:
// 221 rows, two tiles per column in one round, 6-10 % faster per step than 58
// rows in four rounds); the two-pass kernels re-read about 4 halo rows per
// tile and stay within [8, 64].
static int auto_tz(const std::vector<int>& tiles_x, const std::vector<int>& nz, int slots,
                   int M, int block, int halo = -1) {
  if (halo < 0) halo = 2 * M;  // rows a fused tile marches beyond its own
  int64_t rows = 0;
  for (size_t i = 0; i < nz.size(); ++i) rows += (int64_t)tiles_x[i] * nz[i];
  int best = 32;
  double best_eff = -1, best_work = 0;
  const int hi = block > 0 ? 256 : 64;
  for (int tz = 8; tz <= hi; ++tz) {
    int64_t it = 0;
    for (size_t i = 0; i < nz.size(); ++i) it += (int64_t)tiles_x[i] * ((nz[i] + tz - 1) / tz);
    const int64_t rounds = (it + slots - 1) / slots;
    const double march = block > 0 ? (double)((tz + halo + block - 1) / block * block) : tz + 4.0;
    const double eff = (double)rows / ((double)rounds * slots * march);
    const double work = (double)it * march;
    if (eff > best_eff + 1e-9 || (eff > best_eff - 1e-9 && work < best_work)) {
      best_eff = eff;
      best_work = work;
      best = tz;
    }
  }
  return best;
}

// Orders the double step runs at.  At fd_order 12 k_fused2 takes 492 of a
// SIMD lane's 512 registers, so no edge or strip kernel fits beside it and
// the chains run after it: the 64-shot order-12 batch took 0.997 ms per step
// that way against 0.790 whole-grid, and such runs take the split step.
static bool dbl_order(int M) { return M == 2 || M == 4; }
// The two-step kernel's coefficient form: read from the five stored arrays
// (0, every FP32 model holds them) or formed from rho_vx, lam, mu_txx (1,
// models that hold those too).  The stored form issues 17 % fewer
// instructions (no divisions) for 8 more bytes per cell and two steps: its
// launch takes 259 us against 309 us on the SEAM-shaped shot, and the double
// step 0.182 against 0.190 ms per step.
// Batches of two or more shots in the whole-grid step run as two halves on
// two streams (launch_alt_steps); 0 runs them on one stream.
#ifndef FOMO_ALT
#define FOMO_ALT 1
#endif
// Height of the double step's step-(k+1) edge tiles in edge rows (SEAM-shaped
// shot: 0.1829 ms per step with 1, 0.1815 with 2, 0.1781 with 3, 0.1822 with 4).
constexpr int DBL_B_ROWS = 3;
#ifndef DBL_DERIVED
#define DBL_DERIVED 0
#endif
static bool dbl_cmp(const fomo_batch* b) { return b->compact && DBL_DERIVED; }

// Resident wavefronts of k_fused2 on the whole device.
template <int M, bool CMP>
static int dbl_slots_m(int cus) {
  int bv = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bv, k_fused2<M, WPB, CMP>, 64 * WPB, 0) != hipSuccess)
    bv = 1;
  return std::max(1, bv) * WPB * std::max(cus, 1);
}
template <bool CMP>
static int dbl_slots_c(int M, int cus) {
  return M == 2 ? dbl_slots_m<2, CMP>(cus) : dbl_slots_m<4, CMP>(cus);
}
static int dbl_slots(const fomo_batch* b, int cus) {
  return dbl_cmp(b) ? dbl_slots_c<true>(b->M, cus) : dbl_slots_c<false>(b->M, cus);
}

// The third state buffer of w (the double step), laid out like the other two
// and zeroed once: its guard rows and padding are never written.
static int ensure_buf2(fomo_wavefield* w) {
  if (w->dbuf2) return FOMO_OK;
  const fomo_model* m = w->m;
  fomo_ctx* c = m->ctx;
  const size_t bytes = 5 * m->fstride * m->fes;
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipMalloc(&w->dbuf2, bytes));
  HIP_TRY(c, hipMemsetAsync(w->dbuf2, 0, bytes, c->stream));
  for (int k = 0; k < 5; ++k) w->buf2[k] = block_origin(w->dbuf2, m->fstride, k, m->pitch, m->fes);
  w->bytes2 = (int64_t)bytes;
  return FOMO_OK;
}

static void destroy_batch(fomo_batch* b) {
  if (!b) return;
  for (int q = 0; q < 2; ++q) {
    if (b->gexec[q]) (void)hipGraphExecDestroy(b->gexec[q]);
    if (b->graph[q]) (void)hipGraphDestroy(b->graph[q]);
    if (b->gexec2[q]) (void)hipGraphExecDestroy(b->gexec2[q]);
    if (b->graph2[q]) (void)hipGraphDestroy(b->graph2[q]);
  }
  (void)hipFree(b->ds);
  (void)hipFree(b->dn);
  (void)hipFree(b->dprefix);
  if (b->ev_edge) (void)hipEventDestroy(b->ev_edge);
  if (b->ev_chain) (void)hipEventDestroy(b->ev_chain);
  if (b->ev_dedge) (void)hipEventDestroy(b->ev_dedge);
  for (int h = 0; h < 2; ++h) {
    (void)hipFree(b->dprefix_h[h]);
    (void)hipFree(b->dn_h[h]);
  }
  for (hipEvent_t e : {b->ev_fork, b->ev_join})
    if (e) (void)hipEventDestroy(e);
  delete b;
}

static int build_batch(fomo_ctx* c, int n, fomo_model* const* ms, fomo_wavefield* const* ws,
                       fomo_shot* const* ss, fomo_batch** out) {
  if (n < 1) return fail(c, FOMO_ERR_ARG, "empty batch");
  auto b = std::unique_ptr<fomo_batch, void (*)(fomo_batch*)>(new fomo_batch, destroy_batch);
  b->ctx = c;
  b->n = n;
  b->cpt = c->cpt;
  b->tz = c->tz;
  b->tz_req = c->tz;
  b->graph_steps = c->graph_steps;
  b->split = c->step_mode >= 2;
  b->step_mode = c->step_mode >= 2 ? 1 : c->step_mode;
  b->step_mode_req = c->step_mode;
  b->tze = c->tze;
  b->int_bpc = c->int_bpc;
  if (!ms[0]) return fail(c, FOMO_ERR_ARG, "NULL batch member");
  b->M = ms[0]->M;
  b->storage = ms[0]->storage;
  if (b->storage != FOMO_STORAGE_F32 && b->step_mode != 1)
    return fail(c, FOMO_ERR_UNSUPPORTED, "BF16/FP16 storage runs the fused step only (step mode 1)");
  // The fused kernel addresses each state buffer (5 fields) and the
  // coefficient blocks as byte ranges below 2 GiB; larger grids take the
  // two-pass step (FP32) or are refused (reduced storage).
  for (int i = 0; i < n && b->step_mode == 1; ++i) {
    if (!ms[i]) break;
    const size_t fb = 5 * ms[i]->fstride * ms[i]->fes, cb = (size_t)ms[i]->ncoef_blocks * ms[i]->fstride * ms[i]->es;
    if (fb >= 0x7FFFFFFF || cb >= 0x7FFFFFFF) {
      if (b->storage != FOMO_STORAGE_F32)
        return fail(c, FOMO_ERR_UNSUPPORTED, "grid too large for the fused step with reduced storage");
      b->step_mode = 0;
      b->split = false;
    }
  }
  for (int i = 0; i < n; ++i) {
    if (!ms[i] || !ws[i] || !ss[i]) return fail(c, FOMO_ERR_ARG, "NULL batch member");
    if (ms[i]->ctx != c || ws[i]->m != ms[i] || ss[i]->m != ms[i])
      return fail(c, FOMO_ERR_STATE, "batch member belongs to another model or context");
    if (ms[i]->M != b->M) return fail(c, FOMO_ERR_UNSUPPORTED, "batch mixes FD orders");
    if (ms[i]->storage != b->storage)
      return fail(c, FOMO_ERR_UNSUPPORTED, "batch mixes storage types");
    for (int j = 0; j < i; ++j)
      if (ws[j] == ws[i]) return fail(c, FOMO_ERR_ARG, "a wavefield appears twice in the batch");
    b->ws.push_back(ws[i]);
    b->ss.push_back(ss[i]);
    if (ms[i]->d.dtx != ms[i]->d.dtz) b->same = false;
  }
  b->compact = b->same && (b->storage == FOMO_STORAGE_F32 || b->storage == FOMO_STORAGE_F24);
  for (int i = 0; i < n; ++i) {
    b->compact = b->compact && ms[i]->compact;
    b->habc_rows = std::max(b->habc_rows, ms[i]->d.nbc + 1);
  }
  if (b->tz == 0) {
    std::vector<int> tx, nzs;
    for (int i = 0; i < n; ++i) {
      const int M = ms[i]->M, hl = (M + b->cpt - 1) / b->cpt * b->cpt;
      const int w = b->step_mode == 1 ? 64 * b->cpt - 2 * hl : 64 * b->cpt;
      tx.push_back((ms[i]->d.nx + w - 1) / w);
      nzs.push_back(ms[i]->d.nz);
    }
    b->tz = auto_tz(tx, nzs, stencil_slots(b->storage, b->M, b->cpt, c->cus, b->step_mode, b->compact), b->M,
                    b->step_mode == 1 ? fused_block(b->M) : 0);
  }
  if (b->split) {
    // interior tile height: whole rounds of the slots the interior may take
    std::vector<int> tx, nzs;
    for (int i = 0; i < n; ++i) {
      const ShotDev g = make_shotdev(ms[i], ws[i], ss[i], b->cpt, b->tz, TBL_FUSED0, BAND_ROWS_SINGLE,
                                     b->tz, b->tze);
      tx.push_back(g.nix);
      nzs.push_back(g.zi1 - g.zi0);
    }
    int slots = stencil_slots(b->storage, b->M, b->cpt, c->cus, 1, b->compact);
    if (b->int_bpc > 0) slots = std::min(slots, b->int_bpc * WPB * std::max(c->cus, 1));
    b->tzi = c->tz > 0 ? c->tz : auto_tz(tx, nzs, slots, b->M, fused_block(b->M));
  }
  // Double step: FP32 with derived coefficients, an order k_fused2 is built
  // for, and per member nbc >= M (every cell within 3M of the interior is
  // then at least M from the grid's edge, which k_fused2 relies on) and an
  // interior that stays non-empty shrunk by 2M, with no receiver inside the
  // shrunk interior (its first step is recorded from the third buffer, which
  // holds that step outside it).  Otherwise the split step runs.
  if (b->split && b->step_mode == 1 && c->step_mode == 3 && b->storage == FOMO_STORAGE_F32 && b->same &&
      dbl_order(b->M)) {
    bool ok = true;
    std::vector<int> tx, nzs;
    for (int i = 0; i < n && ok; ++i) {
      const ShotDev g = make_shotdev(ms[i], ws[i], ss[i], b->ecpt, b->tz, TBL_DA0, BAND_ROWS_SINGLE,
                                     b->tzi, b->tze);
      if (g.xi1 <= g.xi0 || ms[i]->d.nbc < ms[i]->M) ok = false;
      const std::vector<int>& r = ss[i]->hrec;
      for (int q = 0; q < ss[i]->n_rec && ok; ++q) {
        const int ri = r[3 * q], rj = r[3 * q + 1];
        if (ri >= g.xi0 && ri < g.xi1 && rj >= g.zi0 && rj < g.zi1) ok = false;
      }
      const ShotDev h = make_shotdev(ms[i], ws[i], ss[i], b->cpt, b->tz, TBL_DI0, BAND_ROWS_SINGLE,
                                     b->tzi, b->tze, 1);
      tx.push_back(h.nix);
      nzs.push_back(h.zi1 - h.zi0);
    }
    if (ok) {
      for (int i = 0; i < n; ++i) {
        int rc = ensure_buf2(ws[i]);
        if (rc) return rc;
      }
      int slots = dbl_slots(b.get(), c->cus);
      if (b->int_bpc > 0) slots = std::min(slots, b->int_bpc * WPB * std::max(c->cus, 1));
      b->tzd = c->tz > 0 ? c->tz : auto_tz(tx, nzs, slots, b->M, dbl_block(b->M), 6 * b->M);
      b->dbl = true;
    }
  }
  b->hs.resize((size_t)TBL_COUNT * n);
  for (int t = 0; t < TBL_COUNT; ++t) {
    const bool dedge = t == TBL_DA0 || t == TBL_DA1 || t == TBL_DB0 || t == TBL_DB1;
    // the double step's step-(k+1) edge tiles run beside the interior launch,
    // where the bytes they move cost more than their latency: DBL_B_ROWS
    // times taller tiles (fewer halo rows re-read)
    const int tze = (t == TBL_DB0 || t == TBL_DB1) ? b->tze * DBL_B_ROWS : b->tze;
    for (int i = 0; i < n; ++i)
      b->hs[(size_t)t * n + i] = make_shotdev(ms[i], ws[i], ss[i], dedge ? b->ecpt : b->cpt, b->tz, t,
                                              n == 1 ? BAND_ROWS_SINGLE : BAND_ROWS_BATCH,
                                              b->split ? b->tzi : b->tz, tze, b->tzd);
  }
  if (b->split) {
    b->rec_in_frame = true;
    for (int i = 0; i < n; ++i) {
      const ShotDev& g = b->hs[(size_t)TBL_FUSED0 * n + i];
      const std::vector<int>& r = ss[i]->hrec;
      for (int q = 0; q < ss[i]->n_rec; ++q) {
        const int ri = r[3 * q], rj = r[3 * q + 1];
        if (ri >= g.xi0 && ri < g.xi1 && rj >= g.zi0 && rj < g.zi1) b->rec_in_frame = false;
      }
    }
    HIP_TRY(c, hipEventCreateWithFlags(&b->ev_edge, hipEventDisableTiming));
    HIP_TRY(c, hipEventCreateWithFlags(&b->ev_chain, hipEventDisableTiming));
    HIP_TRY(c, hipEventCreateWithFlags(&b->ev_dedge, hipEventDisableTiming));
  }
  b->hprefix.assign((size_t)PART_COUNT * (n + 1), 0);
  for (int p = 0; p < PART_COUNT; ++p) {
    int acc = 0;
    for (int i = 0; i < n; ++i) {
      b->hprefix[(size_t)p * (n + 1) + i] = acc;
      const int row = p == PART_DBL_EDGE_A ? TBL_DA0 : p == PART_DBL_EDGE_B ? TBL_DB0
                      : p == PART_DBL_INT ? TBL_DI0 : TBL_IN0;
      const int64_t next = (int64_t)acc + part_items(b->hs[(size_t)row * n + i], p);
      if (next > (1 << 30)) return fail(c, FOMO_ERR_ARG, "batch work list too large");
      acc = (int)next;
    }
    b->hprefix[(size_t)p * (n + 1) + n] = acc;
  }
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipMalloc(&b->ds, b->hs.size() * sizeof(ShotDev)));
  HIP_TRY(c, hipMemcpy(b->ds, b->hs.data(), b->hs.size() * sizeof(ShotDev), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMalloc(&b->dn, sizeof(int)));
  HIP_TRY(c, hipMemcpy(b->dn, &n, sizeof(int), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMalloc(&b->dprefix, b->hprefix.size() * sizeof(int)));
  HIP_TRY(c, hipMemcpy(b->dprefix, b->hprefix.data(), b->hprefix.size() * sizeof(int),
                       hipMemcpyHostToDevice));
  if (FOMO_ALT && n >= 2 && !b->split && b->step_mode == 1) {
    // halves of about equal fused work: half 0 ends at the first member
    // whose fused tiles reach half of the batch's
    const int* fp = b->hprefix.data() + (size_t)PART_FUSED * (n + 1);
    int nh = 1;
    while (nh < n - 1 && 2 * fp[nh] < fp[n]) ++nh;
    b->nh = nh;
    for (int h = 0; h < 2; ++h) {
      const int lo = h ? nh : 0, hi = h ? n : nh, cnt = hi - lo;
      std::vector<int>& pre = b->hprefix_h[h];
      pre.assign((size_t)PART_COUNT * (cnt + 1), 0);
      for (int q = 0; q < PART_COUNT; ++q) {
        const int* full = b->hprefix.data() + (size_t)q * (n + 1);
        for (int i = 0; i <= cnt; ++i) pre[(size_t)q * (cnt + 1) + i] = full[lo + i] - full[lo];
      }
      HIP_TRY(c, hipMalloc(&b->dprefix_h[h], pre.size() * sizeof(int)));
      HIP_TRY(c, hipMemcpy(b->dprefix_h[h], pre.data(), pre.size() * sizeof(int), hipMemcpyHostToDevice));
      HIP_TRY(c, hipMalloc(&b->dn_h[h], sizeof(int)));
      HIP_TRY(c, hipMemcpy(b->dn_h[h], &cnt, sizeof(int), hipMemcpyHostToDevice));
    }
    for (hipEvent_t* e : {&b->ev_fork, &b->ev_join})
      HIP_TRY(c, hipEventCreateWithFlags(e, hipEventDisableTiming));
    b->alt = true;
  }
  *out = b.release();
  return FOMO_OK;
}

static void drop_single(fomo_wavefield* w) {
  for (int q = 0; q < 2; ++q) {
    if (w->single[q]) destroy_batch(w->single[q]);
    w->single[q] = nullptr;
    w->single_id[q] = 0;
  }
  if (w->placeholder) fomo_shot_destroy(w->placeholder);
  w->placeholder = nullptr;
}

// The one-row batch for (m, w, s); s == NULL selects an idle placeholder shot
// (no source, no receivers, nt = 1) for the stages that take no shot.
static int single_batch(fomo_model* m, fomo_wavefield* w, fomo_shot* s, fomo_batch** out) {
  API_LOCK;  // may allocate (a new batch, the placeholder shot)
  fomo_ctx* c = m->ctx;
  if (w->m != m || (s && s->m != m)) return fail(c, FOMO_ERR_STATE, "wavefield/shot of another model");
  if (!s) {
    if (!w->placeholder) {
      int rc = fomo_shot_create(m, 1, 1, FOMO_SRC_PRESSURE, nullptr, 0, 1, nullptr, nullptr,
                                nullptr, 0, &w->placeholder);
      if (rc) return rc;
    }
    s = w->placeholder;
  }
  const int q = (s == w->placeholder) ? 0 : 1;
  fomo_batch* b = w->single[q];
  if (b && w->single_id[q] == s->id && b->cpt == c->cpt && b->tz_req == c->tz &&
      b->graph_steps == c->graph_steps && b->step_mode_req == c->step_mode && b->tze == c->tze &&
      b->int_bpc == c->int_bpc) {
    *out = b;
    return FOMO_OK;
  }
  if (b) destroy_batch(b);
  w->single[q] = nullptr;
  w->single_id[q] = 0;
  int rc = build_batch(c, 1, &m, &w, &s, &w->single[q]);
  if (rc) return rc;
  w->single_id[q] = s->id;
  *out = w->single[q];
  return FOMO_OK;
}

// ------------------------------------------------------------------ launches
// Work items of a part over the members launches cover (b->half).
static int part_total(const fomo_batch* b, int part) {
  if (b->half >= 0) {
    const int cnt = b->half ? b->n - b->nh : b->nh;
    return b->hprefix_h[b->half][(size_t)part * (cnt + 1) + cnt];
  }
  return b->hprefix[(size_t)part * (b->n + 1) + b->n];
}

static int blocks_for(const fomo_batch* b, int part, int wpb = WPB) {
  if (b->grid_fixed[part] > 0) return b->grid_fixed[part];
  return std::max(1, (part_total(b, part) + wpb - 1) / wpb);
}

static BatchArgs args_for(const fomo_batch* b, int table, int part, int k_off) {
  BatchArgs a;
  a.shots = b->ds + (size_t)table * b->n;
  a.prefix = b->dprefix + (size_t)part * (b->n + 1);
  a.n_shots_p = b->dn;
  if (b->half >= 0) {
    const int cnt = b->half ? b->n - b->nh : b->nh;
    a.shots += b->half ? b->nh : 0;
    a.prefix = b->dprefix_h[b->half] + (size_t)part * (cnt + 1);
    a.n_shots_p = b->dn_h[b->half];
  }
  a.k_off = k_off;
  return a;
}

// kind: 0 velocity, 1 stress (in-place two-pass, FP32 only), 2 fused, 3 band
template <int K, int M, int CPT>
static void launch_stencil_mc(const fomo_batch* b, int table, int kind, int flags, int k_off) {
  dim3 block(64 * WPB);
  hipStream_t st = lstream(b);
  if (kind == 2 || kind == 4 || kind == 5 || kind == 6 || kind == 8) {  // grid in whole XCD rounds (xcd_block)
    // 4: the split step's edge tiles, 5: its interior tiles (at most int_bpc
    // blocks per CU, so that the strip kernels beside them find free slots),
    // 6 / 8: the double step's edge tiles of step k / k + 1
    const int part = kind == 2   ? PART_FUSED
                     : kind == 4 ? PART_FUSED_EDGE
                     : kind == 6 ? PART_DBL_EDGE_A
                     : kind == 8 ? PART_DBL_EDGE_B
                                 : PART_FUSED_INT;
    if (kind == 4 || kind == 6 || kind == 8) flags |= FLAG_EDGE;
    if (kind == 5) flags |= FLAG_INT;
    int nb = blocks_for(b, part);
    if (kind == 5 && b->int_bpc > 0) nb = std::min(nb, b->int_bpc * std::max(b->ctx->cus, 1));
    const dim3 grid(round_up(nb, NUM_XCD));
    if (b->same && b->compact) {
      if constexpr (K == ST_F32 || K == ST_F24)
        hipLaunchKernelGGL((k_fused<M, CPT, WPB, K, true, true>), grid, block, 0, st,
                           args_for(b, table, part, k_off), flags);
    } else if (b->same) {
      hipLaunchKernelGGL((k_fused<M, CPT, WPB, K, true, false>), grid, block, 0, st,
                         args_for(b, table, part, k_off), flags);
    } else {
      hipLaunchKernelGGL((k_fused<M, CPT, WPB, K, false, false>), grid, block, 0, st,
                         args_for(b, table, part, k_off), flags);
    }
  } else if (kind == 3) {
    if constexpr (CPT == 1)  // one instance per (K, M)
      hipLaunchKernelGGL((k_band<M, WPB, K>), dim3(blocks_for(b, PART_BAND)), block, 0, st,
                         args_for(b, table, PART_BAND, k_off), flags);
  } else if constexpr (K == ST_F32) {
    if (kind == 0 && b->same)
      hipLaunchKernelGGL((k_velocity<M, CPT, WPB, true>), dim3(blocks_for(b, PART_STENCIL)), block, 0,
                         st, args_for(b, table, PART_STENCIL, k_off), flags);
    else if (kind == 0)
      hipLaunchKernelGGL((k_velocity<M, CPT, WPB, false>), dim3(blocks_for(b, PART_STENCIL)), block, 0,
                         st, args_for(b, table, PART_STENCIL, k_off), flags);
    else
      hipLaunchKernelGGL((k_stress<M, CPT, WPB>), dim3(blocks_for(b, PART_STENCIL)), block, 0, st,
                         args_for(b, table, PART_STENCIL, k_off), flags);
  }
}

template <int K, int M>
static void launch_stencil_m(const fomo_batch* b, int table, int kind, int flags, int k_off) {
  // the band kernel has one column per lane; the double step's edge tiles ecpt
  const int cpt = kind == 3 ? 1 : (kind == 6 || kind == 8) ? b->ecpt : b->cpt;
  switch (cpt) {
    case 1: launch_stencil_mc<K, M, 1>(b, table, kind, flags, k_off); break;
    case 2: launch_stencil_mc<K, M, 2>(b, table, kind, flags, k_off); break;
    default: launch_stencil_mc<K, M, 4>(b, table, kind, flags, k_off); break;
  }
}

template <int K>
static void launch_stencil_reduced(const fomo_batch* b, int table, int kind, int flags, int k_off) {
  switch (b->M) {  // the model accepted M in {2, 4, 6} only
    case 2: launch_stencil_m<K, 2>(b, table, kind, flags, k_off); break;
    case 4: launch_stencil_m<K, 4>(b, table, kind, flags, k_off); break;
    default: launch_stencil_m<K, 6>(b, table, kind, flags, k_off); break;
  }
}

static void launch_stencil(const fomo_batch* b, int table, int kind, int flags, int k_off) {
  if (b->storage == FOMO_STORAGE_BF16) return launch_stencil_reduced<ST_BF16>(b, table, kind, flags, k_off);
  if (b->storage == FOMO_STORAGE_F16) return launch_stencil_reduced<ST_F16>(b, table, kind, flags, k_off);
  if (b->storage == FOMO_STORAGE_F24) return launch_stencil_reduced<ST_F24>(b, table, kind, flags, k_off);
  if (b->storage == FOMO_STORAGE_F16C)
    return launch_stencil_reduced<ST_F16C>(b, table, kind, flags, k_off);
  switch (b->M) {
    case 1: launch_stencil_m<ST_F32, 1>(b, table, kind, flags, k_off); break;
    case 2: launch_stencil_m<ST_F32, 2>(b, table, kind, flags, k_off); break;
    case 3: launch_stencil_m<ST_F32, 3>(b, table, kind, flags, k_off); break;
    case 4: launch_stencil_m<ST_F32, 4>(b, table, kind, flags, k_off); break;
    case 5: launch_stencil_m<ST_F32, 5>(b, table, kind, flags, k_off); break;
    case 6: launch_stencil_m<ST_F32, 6>(b, table, kind, flags, k_off); break;
    case 7: launch_stencil_m<ST_F32, 7>(b, table, kind, flags, k_off); break;
    default: launch_stencil_m<ST_F32, 8>(b, table, kind, flags, k_off); break;
  }
}

// The double step's two-step interior tiles (FP32, derived coefficients),
// at most int_bpc blocks per CU like the split step's interior.
static void launch_dbl_interior(const fomo_batch* b, int table, int flags, int k_off) {
  int nb = blocks_for(b, PART_DBL_INT);
  if (b->int_bpc > 0) nb = std::min(nb, b->int_bpc * std::max(b->ctx->cus, 1));
  const dim3 grid(round_up(nb, NUM_XCD)), block(64 * WPB);
  const BatchArgs a = args_for(b, table, PART_DBL_INT, k_off);
  hipStream_t st = lstream(b);
  if (dbl_cmp(b)) {
    if (b->M == 2) hipLaunchKernelGGL((k_fused2<2, WPB, true>), grid, block, 0, st, a, flags);
    else hipLaunchKernelGGL((k_fused2<4, WPB, true>), grid, block, 0, st, a, flags);
  } else {
    if (b->M == 2) hipLaunchKernelGGL((k_fused2<2, WPB, false>), grid, block, 0, st, a, flags);
    else hipLaunchKernelGGL((k_fused2<4, WPB, false>), grid, block, 0, st, a, flags);
  }
}

static void launch_habc(const fomo_batch* b, int table, bool stress, int k_off) {
  const int part = stress ? PART_HABC_S : PART_HABC_V;
  const int items = part_total(b, part);
  const BatchArgs a = args_for(b, table, part, k_off);
  const dim3 g(b->grid_fixed[part] > 0 ? b->grid_fixed[part] : std::max(1, items));
  const int rows = b->habc_rows;
  // + the corner's padding row and the bottom-right sweep's read-ahead
  const size_t lds = ((size_t)2 * (rows + 1) * HS + habc_lds_slack(rows)) * sizeof(float);
  if (b->storage == FOMO_STORAGE_BF16)
    hipLaunchKernelGGL(k_habc<ST_BF16>, g, dim3(64), lds, lstream(b), a, stress ? 1 : 0, rows);
  else if (b->storage == FOMO_STORAGE_F16)
    hipLaunchKernelGGL(k_habc<ST_F16>, g, dim3(64), lds, lstream(b), a, stress ? 1 : 0, rows);
  else if (b->storage == FOMO_STORAGE_F24)
    hipLaunchKernelGGL(k_habc<ST_F24>, g, dim3(64), lds, lstream(b), a, stress ? 1 : 0, rows);
  else
    hipLaunchKernelGGL(k_habc<ST_F32>, g, dim3(64), lds, lstream(b), a, stress ? 1 : 0, rows);
}

static void launch_fsrec(const fomo_batch* b, int table, int mode, int k_off) {
  const BatchArgs a = args_for(b, table, PART_FSREC, k_off);
  const dim3 g(blocks_for(b, PART_FSREC)), blk(64 * WPB);
  if (b->storage == FOMO_STORAGE_BF16)
    hipLaunchKernelGGL((k_fsrec<WPB, ST_BF16>), g, blk, 0, lstream(b), a, mode);
  else if (b->storage == FOMO_STORAGE_F16)
    hipLaunchKernelGGL((k_fsrec<WPB, ST_F16>), g, blk, 0, lstream(b), a, mode);
  else if (b->storage == FOMO_STORAGE_F24)
    hipLaunchKernelGGL((k_fsrec<WPB, ST_F24>), g, blk, 0, lstream(b), a, mode);
  else
    hipLaunchKernelGGL((k_fsrec<WPB, ST_F32>), g, blk, 0, lstream(b), a, mode);
}

static void launch_frame(const fomo_batch* b, const FrameArgs& fa0, int table, int k_off, int slot) {
  FrameArgs fa = fa0;
  fa.slot = slot;
  const size_t total = (size_t)fa.nx_d * fa.nz_d * fa.nfields;
  const dim3 g((unsigned)std::max<size_t>(1, std::min<size_t>((total + 255) / 256, 2048)));
  const BatchArgs a = args_for(b, table, PART_FSREC, k_off);
  if (b->storage == FOMO_STORAGE_BF16)
    hipLaunchKernelGGL(k_frame<ST_BF16>, g, dim3(256), 0, lstream(b), a, fa);
  else if (b->storage == FOMO_STORAGE_F16)
    hipLaunchKernelGGL(k_frame<ST_F16>, g, dim3(256), 0, lstream(b), a, fa);
  else if (b->storage == FOMO_STORAGE_F24)
    hipLaunchKernelGGL(k_frame<ST_F24>, g, dim3(256), 0, lstream(b), a, fa);
  else
    hipLaunchKernelGGL(k_frame<ST_F32>, g, dim3(256), 0, lstream(b), a, fa);
}

// One step (time_stepper.jl:39-63) of every shot, starting from state buffer
// `parity`.  Fused: velocity+stress in one pass (backup = the read buffer),
// velocity HABC, band stress, stress HABC, free surface + receivers.
// Two-pass: in place, with backup and injection inside the stencil kernels.
static void launch_step(const fomo_batch* b, int parity, int k_off) {
  if (b->split) {
    // The interior tiles on the batch's stream, beside the edge tiles and
    // then the strip kernels on the aux stream: the interior and the edge
    // tiles both read only the previous state and write disjoint cells, and
    // the strip kernels read and write only the band frame, whose cells the
    // interior tiles neither read nor write (they lie 2M + 2 cells inside
    // it).  Both streams are joined before the next step.
    const int t = TBL_FUSED0 + parity;
    hipStream_t st = lstream(b), ax = b->ctx->aux;
    (void)hipEventRecord(b->ev_edge, st);
    (void)hipStreamWaitEvent(ax, b->ev_edge, 0);
    b->cur = ax;
    launch_stencil(b, t, 4, FLAG_INJECT, k_off);
    launch_habc(b, t, false, k_off);
    launch_stencil(b, t, 3, FLAG_INJECT, k_off);
    launch_habc(b, t, true, k_off);
    launch_fsrec(b, t, b->rec_in_frame ? (MODE_FS | MODE_RECORD) : MODE_FS, k_off);
    (void)hipEventRecord(b->ev_chain, ax);
    b->cur = nullptr;
    launch_stencil(b, t, 5, FLAG_INJECT, k_off);
    (void)hipStreamWaitEvent(st, b->ev_chain, 0);
    if (!b->rec_in_frame) launch_fsrec(b, t, MODE_RECORD, k_off);
    if (b->frames_id) launch_frame(b, b->fargs, t, k_off, -1);
    return;
  }
  if (b->step_mode == 1) {
    const int t = TBL_FUSED0 + parity;
    launch_stencil(b, t, 2, FLAG_INJECT, k_off);
    launch_habc(b, t, false, k_off);
    launch_stencil(b, t, 3, FLAG_INJECT, k_off);
    launch_habc(b, t, true, k_off);
    launch_fsrec(b, t, MODE_FS | MODE_RECORD, k_off);
    if (b->frames_id) launch_frame(b, b->fargs, t, k_off, -1);
  } else {
    const int t = TBL_IN0 + parity;
    launch_stencil(b, t, 0, FLAG_BACKUP | FLAG_INJECT, k_off);
    launch_habc(b, t, false, k_off);
    launch_stencil(b, t, 1, FLAG_BACKUP | FLAG_INJECT, k_off);
    launch_habc(b, t, true, k_off);
    launch_fsrec(b, t, MODE_FS | MODE_RECORD, k_off);
    if (b->frames_id) launch_frame(b, b->fargs, t, k_off, -1);
  }
}

// Steps k_off and k_off + 1 of every shot from state buffer `parity` into
// buffer 1 - parity (time_stepper.jl:39-63, twice).  On the aux stream, two
// split-step chains: step k outside the interior shrunk by 2M into the third
// buffer C, then step k + 1 outside the interior from C.  On the batch's
// stream, beside them, the interior's two steps in one launch (k_fused2),
// started once the step-k edge tiles are done: those latency-bound tiles
// then run without the interior's waves on their SIMDs, and the rest of the
// chains still fit beside the interior launch (0.182 ms per step against
// 0.193 with both started together; waiting also for the velocity HABC, or
// for the whole step-k chain, measured 0.196 and 0.202 ms on SEAM).
// Why this is race-free:
//   - A (buffer `parity`) is only read during the double step;
//   - C is written by the first chain and read by the second, in stream
//     order; the second chain reads C within 2M of the cells it writes, all
//     of which the first chain wrote (the shrunk interior is 2M smaller);
//   - B is written on disjoint cells: the interior by k_fused2, the rest by
//     the second chain, whose strip kernels read only cells at least 2M + 2
//     away from the interior;
//   - the interior's two steps read A within 4M of the interior, whose
//     step-k cells (within 2M) are outside every strip kernel's cells, so
//     no HABC, band or free-surface update is missed.
// Receivers are recorded from C for step k (none lies in the shrunk
// interior) and from B for step k + 1.
static void launch_double(const fomo_batch* b, int parity, int k_off) {
  const int ta = TBL_DA0 + parity, tb = TBL_DB0 + parity, ti = TBL_DI0 + parity;
  hipStream_t st = lstream(b), ax = b->ctx->aux;
  (void)hipEventRecord(b->ev_edge, st);
  (void)hipStreamWaitEvent(ax, b->ev_edge, 0);
  b->cur = ax;
  launch_stencil(b, ta, 6, FLAG_INJECT, k_off);
  (void)hipEventRecord(b->ev_dedge, ax);
  launch_habc(b, ta, false, k_off);
  launch_stencil(b, ta, 3, FLAG_INJECT, k_off);
  launch_habc(b, ta, true, k_off);
  launch_fsrec(b, ta, MODE_FS | MODE_RECORD, k_off);
  launch_stencil(b, tb, 8, FLAG_INJECT, k_off + 1);
  launch_habc(b, tb, false, k_off + 1);
  launch_stencil(b, tb, 3, FLAG_INJECT, k_off + 1);
  launch_habc(b, tb, true, k_off + 1);
  launch_fsrec(b, tb, b->rec_in_frame ? (MODE_FS | MODE_RECORD) : MODE_FS, k_off + 1);
  (void)hipEventRecord(b->ev_chain, ax);
  b->cur = nullptr;
  (void)hipStreamWaitEvent(st, b->ev_dedge, 0);
  launch_dbl_interior(b, ti, FLAG_INJECT, k_off);
  (void)hipStreamWaitEvent(st, b->ev_chain, 0);
  if (!b->rec_in_frame) launch_fsrec(b, tb, MODE_RECORD, k_off + 1);
  // frames: only step k + 1 is whole in memory; run_batch takes double steps
  // with a frame store only when every frame step is the second of a pair
  if (b->frames_id) launch_frame(b, b->fargs, tb, k_off + 1, -1);
}

static void launch_set_step(const fomo_batch* b, int k) {
  hipLaunchKernelGGL(k_set_step, dim3((b->n + 255) / 256), dim3(256), 0, lstream(b), b->ds,
                     b->dn, k);
}
// Reads the kp of rows [0, n) of table TBL_IN0 (filled for batches and pools).
// `count` whole-grid steps of a two-stream batch from state buffer `parity`:
// members [0, nh) on the batch's stream, [nh, n) on the context's aux stream,
// each half's steps in stream order and the halves unordered (they share no
// member).  A half's latency-bound strip kernels then overlap the other
// half's bandwidth-bound fused pass part of the time (64 Marmousi-shaped
// shots: 0.747 ms per step against 0.787 on one stream; ordering the fused
// passes so that they never overlap measured 0.944, and 0.824 at best with
// shorter tiles: a half's strip kernels are as long as the whole batch's,
// being latency-bound).  The aux stream joins at the end.
static void launch_alt_steps(const fomo_batch* b, int parity, int count) {
  hipStream_t st = lstream(b), ax = b->ctx->aux;
  (void)hipEventRecord(b->ev_fork, st);
  (void)hipStreamWaitEvent(ax, b->ev_fork, 0);
  for (int g = 0; g < count; ++g) {
    const int t = TBL_FUSED0 + ((parity + g) & 1);
    for (int h = 0; h < 2; ++h) {
      b->half = h;
      b->cur = h ? ax : nullptr;
      launch_stencil(b, t, 2, FLAG_INJECT, g);
      launch_habc(b, t, false, g);
      launch_stencil(b, t, 3, FLAG_INJECT, g);
      launch_habc(b, t, true, g);
      launch_fsrec(b, t, MODE_FS | MODE_RECORD, g);
    }
  }
  b->half = -1;
  b->cur = nullptr;
  (void)hipEventRecord(b->ev_join, ax);
  (void)hipStreamWaitEvent(st, b->ev_join, 0);
}

static void launch_advance(const fomo_batch* b, int by) {
  hipLaunchKernelGGL(k_advance, dim3((b->n + 255) / 256), dim3(256), 0, lstream(b), b->ds,
                     b->dn, by);
}

// Captures one chunk of G whole steps, starting from state buffer p, into
// b->gexec[p] (the step counters advance by G as the last node).
// dbl: G / 2 double steps instead, into b->gexec2[p] (the chunk ends on
// parity p + G / 2).
static int capture_chunk(fomo_batch* b, int p, bool dbl = false) {
  API_LOCK;
  fomo_ctx* c = b->ctx;
  const int G = b->graph_steps;
  const int flip = b->step_mode == 1 ? 1 : 0;
  hipStream_t st = lstream(b);
  HIP_TRY(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  if (dbl)
    for (int g = 0; g < G / 2; ++g) launch_double(b, (p + g) & 1, 2 * g);
  else if (b->alt)
    launch_alt_steps(b, p, G);
  else
    for (int g = 0; g < G; ++g) launch_step(b, (p + flip * g) & 1, g);
  launch_advance(b, G);
  const hipError_t le = hipGetLastError();
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &graph);
  if (le != hipSuccess || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return fail(c, FOMO_ERR_HIP, std::string("step capture: ") +
                                     hipGetErrorString(le != hipSuccess ? le : e));
  }
  (dbl ? b->graph2 : b->graph)[p] = graph;
  HIP_TRY(c, hipGraphInstantiate(&(dbl ? b->gexec2 : b->gexec)[p], graph, nullptr, nullptr, 0));
  return FOMO_OK;
}

static int check_launch(fomo_ctx* c, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(c, FOMO_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return FOMO_OK;
}

// After fused steps the strip backup lives in the previous state buffer; copy
// it into old[] (k_backup from that buffer) before anything reads old[].
static int refresh_backup(fomo_wavefield* w) {
  if (!w->backup_is_prev || w->m->storage != FOMO_STORAGE_F32) return FOMO_OK;
  fomo_batch* b = nullptr;
  int rc = single_batch(w->m, w, nullptr, &b);
  if (rc) return rc;
  HIP_TRY(w->m->ctx, hipSetDevice(w->m->ctx->device));
  launch_set_step(b, 1);
  const int t = w->prev_buf2 ? TBL_IN2 : TBL_IN0 + (1 - w->cur);
  hipLaunchKernelGGL(k_backup<WPB>, dim3(blocks_for(b, PART_BACKUP)), dim3(64 * WPB), 0,
                     w->m->ctx->stream, args_for(b, t, PART_BACKUP, 0));
  w->backup_is_prev = false;
  w->prev_buf2 = false;
  return check_launch(w->m->ctx, "refresh_backup");
}

static int run_batch(fomo_batch* b, int k_first, int n_steps) {
  fomo_ctx* c = b->ctx;
  if (n_steps < 0) return fail(c, FOMO_ERR_ARG, "n_steps < 0");
  if (n_steps == 0) return FOMO_OK;
  // A member steps while k <= its nt.  Members that step in this run must
  // hold their state in the same buffer; a member whose shot has ended is
  // skipped by every kernel and may sit in either.
  int parity = -1;
  for (int i = 0; i < b->n; ++i) {
    if (k_first > b->ss[i]->nt) continue;
    if (parity < 0) parity = b->ws[i]->cur;
    else if (b->ws[i]->cur != parity)
      return fail(c, FOMO_ERR_STATE, "batch wavefields are at different buffer parities");
  }
  if (parity < 0) parity = b->ws[0]->cur;
  if (b->step_mode == 0)
    for (fomo_wavefield* w : b->ws) {
      int rc = refresh_backup(w);
      if (rc) return rc;
    }
  for (int i = 0; i < b->n; ++i) {
    int rc = adopt_scales(b->ws[i], b->ss[i]);
    if (rc) return rc;
  }
  HIP_TRY(c, hipSetDevice(c->device));
  launch_set_step(b, k_first);
  // Double steps where every member is live for every step of the run (a
  // double step needs both of its steps); an odd last step is a split step.
  // With a frame store, every frame step (a multiple of its even skip) must
  // be the second step of a pair, which holds the whole state in memory.
  bool dbl = b->dbl && n_steps >= 2 && k_first >= 1 &&
             (!b->frames_id || (b->fargs.skip % 2 == 0 && k_first % 2 == 1));
  for (int i = 0; i < b->n && dbl; ++i)
    if ((int64_t)k_first + n_steps - 1 > b->ss[i]->nt) dbl = false;
  if (dbl) {
    const int G = b->graph_steps, H = G / 2;
    const int pairs = n_steps / 2, single = n_steps & 1;
    const int chunks = H > 0 ? pairs / H : 0, rest = pairs - chunks * H;
    int q = parity;
    for (int i = 0; i < chunks; ++i) {
      if (!b->gexec2[q]) {
        int rc = capture_chunk(b, q, true);
        if (rc) return rc;
      }
      HIP_TRY(c, hipGraphLaunch(b->gexec2[q], c->stream));
      q = (q + H) & 1;
    }
    for (int g = 0; g < rest; ++g, q ^= 1) launch_double(b, q, 2 * g);
    if (single) {
      launch_step(b, q, 2 * rest);
      q ^= 1;
    }
    if (2 * rest + single) launch_advance(b, 2 * rest + single);
    for (fomo_wavefield* w : b->ws) {
      w->cur = q;
      w->backup_is_prev = true;
      w->prev_buf2 = !single;  // the last step was a double step's second
    }
    return check_launch(c, "time loop");
  }
  const int G = b->graph_steps;  // even: a chunk ends on the parity it started from; 0: no graphs
  const int chunks = G > 0 ? n_steps / G : 0, rem = G > 0 ? n_steps % G : n_steps;
  const int flip = b->step_mode == 1 ? 1 : 0;
  if (chunks > 0 && !b->gexec[parity]) {
    int rc = capture_chunk(b, parity);
    if (rc) return rc;
  }
  for (int i = 0; i < chunks; ++i) HIP_TRY(c, hipGraphLaunch(b->gexec[parity], c->stream));
  if (b->alt && rem)
    launch_alt_steps(b, parity, rem);
  else
    for (int g = 0; g < rem; ++g) launch_step(b, (parity + flip * g) & 1, g);
  if (rem) launch_advance(b, rem);
  if (flip) {
    for (int i = 0; i < b->n; ++i) {
      // the state of a member whose shot ends inside the run is in the
      // buffer of its own last step
      const int64_t live = std::min<int64_t>(n_steps, std::max<int64_t>(0, (int64_t)b->ss[i]->nt - k_first + 1));
      if (live == 0) continue;
      fomo_wavefield* w = b->ws[i];
      w->cur = (parity + (int)live) & 1;
      w->backup_is_prev = true;
      w->prev_buf2 = false;
    }
  }
  return check_launch(c, "time loop");
}

// Attaches frame store f (null: none) to a one-shot batch; its captured
// chunks are dropped when the attachment changes.
static void set_frames(fomo_batch* b, const fomo_frames* f) {
  const uint64_t id = f ? f->id : 0;
  if (b->frames_id == id) return;
  for (int q = 0; q < 2; ++q) {
    if (b->gexec[q]) (void)hipGraphExecDestroy(b->gexec[q]);
    if (b->graph[q]) (void)hipGraphDestroy(b->graph[q]);
    if (b->gexec2[q]) (void)hipGraphExecDestroy(b->gexec2[q]);
    if (b->graph2[q]) (void)hipGraphDestroy(b->graph2[q]);
    b->gexec[q] = b->gexec2[q] = nullptr;
    b->graph[q] = b->graph2[q] = nullptr;
  }
  b->frames_id = id;
  b->fargs = f ? f->a : FrameArgs{};
}

extern "C" int32_t fomo_run_time_loop(fomo_model* m, fomo_wavefield* w, fomo_shot* s,
                                      int32_t k_first, int32_t n_steps) {
  if (!m || !w || !s) return fail(m ? m->ctx : nullptr, FOMO_ERR_ARG, "NULL argument");
  fomo_batch* b = nullptr;
  int rc = single_batch(m, w, s, &b);
  if (rc) return rc;
  set_frames(b, nullptr);
  return run_batch(b, k_first, n_steps);
}

// Captures the step graphs fomo_run_time_loop replays for (m, w, s), for
// both buffer parities of the fused step, without running a step: the first
// run then starts replaying at once instead of capturing inside it.
extern "C" int32_t fomo_run_time_loop_prepare(fomo_model* m, fomo_wavefield* w, fomo_shot* s) {
  if (!m || !w || !s) return fail(m ? m->ctx : nullptr, FOMO_ERR_ARG, "NULL argument");
  fomo_batch* b = nullptr;
  int rc = single_batch(m, w, s, &b);
  if (rc) return rc;
  set_frames(b, nullptr);
  HIP_TRY(m->ctx, hipSetDevice(m->ctx->device));
  const int parities = b->graph_steps == 0 ? 0 : b->step_mode == 1 ? 2 : 1;
  for (int q = 0; q < parities; ++q) {
    const int par = (w->cur + q) & 1;
    if (!b->gexec[par]) {
      rc = capture_chunk(b, par);
      if (rc) return rc;
    }
    if (b->dbl && !b->gexec2[par]) {  // double-step chunks start from either parity
      rc = capture_chunk(b, par, true);
      if (rc) return rc;
    }
  }
  return FOMO_OK;
}

// ------------------------------------------------------------------ frames
extern "C" int32_t fomo_frames_create(fomo_ctx* c, int32_t nx, int32_t nz, const int32_t* fields,
                                      int32_t n_fields, int32_t skip, int32_t downsample,
                                      int32_t capacity, fomo_frames** out) {
  API_LOCK;
  if (!c || !fields || !out) return fail(c, FOMO_ERR_ARG, "NULL argument");
  if (n_fields < 1 || n_fields > 4) return fail(c, FOMO_ERR_ARG, "n_fields must be 1..4");
  if (nx < 1 || nz < 1 || skip < 1 || downsample < 1 || capacity < 1)
    return fail(c, FOMO_ERR_ARG, "nx, nz, skip, downsample and capacity must be >= 1");
  auto f = std::make_unique<fomo_frames>();
  f->ctx = c;
  f->nx = nx;
  f->nz = nz;
  FrameArgs& a = f->a;
  for (int q = 0; q < n_fields; ++q) {
    if (fields[q] < FOMO_FRAME_P || fields[q] > FOMO_FRAME_VZ)
      return fail(c, FOMO_ERR_ARG, "unknown frame field (p, vel, vx, vz)");
    a.field[q] = fields[q];
  }
  a.nfields = n_fields;
  a.skip = skip;
  a.ds = downsample;
  a.nx_d = (nx + downsample - 1) / downsample;  // cld(nx, downsample)
  a.nz_d = (nz + downsample - 1) / downsample;
  a.capacity = capacity;
  a.slot = -1;
  const size_t bytes = (size_t)capacity * n_fields * a.nx_d * a.nz_d * sizeof(float);
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipMalloc(&a.out, bytes));
  HIP_TRY(c, hipMemsetAsync(a.out, 0, bytes, c->stream));
  f->bytes = (int64_t)bytes;
  f->id = g_next_frames_id++;
  *out = f.release();
  return FOMO_OK;
}

extern "C" int32_t fomo_frames_destroy(fomo_frames* f) {
  API_LOCK;
  if (!f) return FOMO_OK;
  (void)hipSetDevice(f->ctx->device);
  (void)hipDeviceSynchronize();  // batches of any stream may still write it
  (void)hipFree(f->a.out);
  delete f;
  return FOMO_OK;
}

extern "C" int32_t fomo_frames_shape(const fomo_frames* f, int32_t* nx_d, int32_t* nz_d) {
  if (!f || !nx_d || !nz_d) return FOMO_ERR_ARG;
  *nx_d = f->a.nx_d;
  *nz_d = f->a.nz_d;
  return FOMO_OK;
}

extern "C" int64_t fomo_frames_device_bytes(const fomo_frames* f) { return f ? f->bytes : 0; }

extern "C" int32_t fomo_frames_download(fomo_frames* f, int32_t first, int32_t count, float* host) {
  if (!f || !host) return fail(f ? f->ctx : nullptr, FOMO_ERR_ARG, "NULL argument");
  if (first < 0 || count < 0 || (int64_t)first + count > f->a.capacity)
    return fail(f->ctx, FOMO_ERR_ARG, "frame range outside the store");
  fomo_ctx* c = f->ctx;
  const size_t per = (size_t)f->a.nfields * f->a.nx_d * f->a.nz_d;
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipMemcpyAsync(host, f->a.out + (size_t)first * per, (size_t)count * per * sizeof(float),
                            hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return FOMO_OK;
}

static int check_frames_ctx(fomo_ctx* c, const fomo_frames* f) {
  if (f->ctx != c) return fail(c, FOMO_ERR_STATE, "frame store of another context");
  return FOMO_OK;
}

extern "C" int32_t fomo_run_time_loop_frames(fomo_model* m, fomo_wavefield* w, fomo_shot* s,
                                             int32_t k_first, int32_t n_steps, fomo_frames* f) {
  if (!m || !w || !s || !f) return fail(m ? m->ctx : nullptr, FOMO_ERR_ARG, "NULL argument");
  int rc = check_frames_ctx(m->ctx, f);
  if (rc) return rc;
  fomo_batch* b = nullptr;
  rc = single_batch(m, w, s, &b);
  if (rc) return rc;
  set_frames(b, f);
  return run_batch(b, k_first, n_steps);
}

extern "C" int32_t fomo_wavefield_frame(fomo_wavefield* w, fomo_frames* f, int32_t slot) {
  if (!w || !f) return fail(w ? w->m->ctx : nullptr, FOMO_ERR_ARG, "NULL argument");
  fomo_model* m = w->m;
  int rc = check_frames_ctx(m->ctx, f);
  if (rc) return rc;
  if (slot < 0 || slot >= f->a.capacity) return fail(m->ctx, FOMO_ERR_ARG, "slot outside the store");
  fomo_batch* b = nullptr;
  rc = single_batch(m, w, nullptr, &b);
  if (rc) return rc;
  FrameArgs a = f->a;
  a.own_sc = 1;  // the placeholder shot's row does not carry the state's scales
  for (int c = 0; c < 5; ++c) a.isc[c] = 1.0f / w->fsc[c];
  HIP_TRY(m->ctx, hipSetDevice(m->ctx->device));
  launch_frame(b, a, TBL_IN0 + w->cur, 0, slot);
  return check_launch(m->ctx, "wavefield_frame");
}

extern "C" int32_t fomo_time_step(fomo_model* m, fomo_wavefield* w, fomo_shot* s, int32_t k) {
  if (!m || !w || !s) return fail(m ? m->ctx : nullptr, FOMO_ERR_ARG, "NULL argument");
  fomo_batch* b = nullptr;
  int rc = single_batch(m, w, s, &b);
  if (rc) return rc;
  if (b->step_mode == 0) {
    rc = refresh_backup(w);
    if (rc) return rc;
  }
  rc = adopt_scales(w, s);
  if (rc) return rc;
  set_frames(b, nullptr);
  HIP_TRY(m->ctx, hipSetDevice(m->ctx->device));
  launch_set_step(b, k);
  launch_step(b, w->cur, 0);
  if (b->step_mode == 1) {
    w->cur ^= 1;
    w->backup_is_prev = true;
    w->prev_buf2 = false;
  }
  return check_launch(m->ctx, "time_step");
}

// Per-stage entry points: the unfused in-place kernels, one stage each, on
// the current buffer with old[] as the strip backup.
static int stage(fomo_model* m, fomo_wavefield* w, fomo_shot* s, int k, int which) {
  if (!m || !w) return fail(m ? m->ctx : nullptr, FOMO_ERR_ARG, "NULL argument");
  if (m->storage != FOMO_STORAGE_F32)
    return fail(m->ctx, FOMO_ERR_UNSUPPORTED, "per-stage calls need FP32 storage");
  int rc = refresh_backup(w);
  if (rc) return rc;
  fomo_batch* b = nullptr;
  rc = single_batch(m, w, s, &b);
  if (rc) return rc;
  fomo_ctx* c = m->ctx;
  const int t = TBL_IN0 + w->cur;
  HIP_TRY(c, hipSetDevice(c->device));
  launch_set_step(b, s ? k : 1);
  switch (which) {
    case 0:
      hipLaunchKernelGGL(k_backup<WPB>, dim3(blocks_for(b, PART_BACKUP)), dim3(64 * WPB), 0,
                         c->stream, args_for(b, t, PART_BACKUP, 0));
      break;
    case 1:
      hipLaunchKernelGGL(k_inject, dim3(1), dim3(64), 0, c->stream, args_for(b, t, PART_SHOT, 0));
      break;
    case 2: launch_stencil(b, t, 0, 0, 0); break;
    case 3: launch_habc(b, t, false, 0); break;
    case 4: launch_stencil(b, t, 1, 0, 0); break;
    case 5: launch_habc(b, t, true, 0); break;
    case 6: launch_fsrec(b, t, MODE_FS, 0); break;
    case 7: launch_fsrec(b, t, MODE_RECORD, 0); break;
  }
  return check_launch(c, "stage");
}

extern "C" int32_t fomo_backup_boundary(fomo_model* m, fomo_wavefield* w) { return stage(m, w, nullptr, 1, 0); }
extern "C" int32_t fomo_inject_source(fomo_model* m, fomo_wavefield* w, fomo_shot* s, int32_t k) {
  if (!s) return fail(m ? m->ctx : nullptr, FOMO_ERR_ARG, "NULL shot");
  return stage(m, w, s, k, 1);
}
extern "C" int32_t fomo_update_velocity(fomo_model* m, fomo_wavefield* w) { return stage(m, w, nullptr, 1, 2); }
extern "C" int32_t fomo_apply_habc_velocity(fomo_model* m, fomo_wavefield* w) { return stage(m, w, nullptr, 1, 3); }
extern "C" int32_t fomo_update_stress(fomo_model* m, fomo_wavefield* w) { return stage(m, w, nullptr, 1, 4); }
extern "C" int32_t fomo_apply_habc_stress(fomo_model* m, fomo_wavefield* w) { return stage(m, w, nullptr, 1, 5); }
extern "C" int32_t fomo_apply_free_surface(fomo_model* m, fomo_wavefield* w) { return stage(m, w, nullptr, 1, 6); }
extern "C" int32_t fomo_record_receivers(fomo_model* m, fomo_wavefield* w, fomo_shot* s, int32_t k) {
  if (!s) return fail(m ? m->ctx : nullptr, FOMO_ERR_ARG, "NULL shot");
  return stage(m, w, s, k, 7);
}

// Timing hook: `reps` launches of one kernel with the step's configuration.
// Stages 0-4 as declared in the header; 5 = fused velocity+stress kernel,
// 6 = band stress kernel, 7 = HABC of velocity, 8 = HABC of stress, the last
// three in the fused step's buffer roles (they read one state buffer and
// write the other, so repeated launches leave the state unchanged).
extern "C" int32_t fomo_launch_stage(fomo_model* m, fomo_wavefield* w, fomo_shot* s, int32_t which,
                                     int32_t k, int32_t reps) {
  if (!m || !w || !s) return fail(m ? m->ctx : nullptr, FOMO_ERR_ARG, "NULL argument");
  if (which < 0 || which > 13 || reps < 0) return fail(m->ctx, FOMO_ERR_ARG, "bad stage or reps");
  if (which < 4 && m->storage != FOMO_STORAGE_F32)
    return fail(m->ctx, FOMO_ERR_UNSUPPORTED, "the in-place stages need FP32 storage");
  fomo_batch* b = nullptr;
  int rc = single_batch(m, w, s, &b);
  if (rc) return rc;
  HIP_TRY(m->ctx, hipSetDevice(m->ctx->device));
  launch_set_step(b, k);
  if (which >= 11 && !b->dbl)
    return fail(m->ctx, FOMO_ERR_UNSUPPORTED, "stages 11-13 need a double-step run (step mode 3)");
  const int ti = TBL_IN0 + w->cur, tf = TBL_FUSED0 + w->cur;
  for (int r = 0; r < reps; ++r) {
    switch (which) {
      case 0: launch_stencil(b, ti, 0, FLAG_BACKUP | FLAG_INJECT, 0); break;
      case 1: launch_stencil(b, ti, 1, FLAG_BACKUP | FLAG_INJECT, 0); break;
      case 2: launch_habc(b, ti, false, 0); break;
      case 3: launch_habc(b, ti, true, 0); break;
      case 4: launch_fsrec(b, ti, MODE_FS | MODE_RECORD, 0); break;
      case 5: launch_stencil(b, tf, 2, FLAG_INJECT, 0); break;
      case 6: launch_stencil(b, tf, 3, FLAG_INJECT, 0); break;
      case 7: launch_habc(b, tf, false, 0); break;
      case 8: launch_habc(b, tf, true, 0); break;
      case 9: launch_stencil(b, tf, 4, FLAG_INJECT, 0); break;
      case 10: launch_stencil(b, tf, 5, FLAG_INJECT, 0); break;
      case 11: launch_dbl_interior(b, TBL_DI0 + w->cur, FLAG_INJECT, 0); break;
      case 12: launch_stencil(b, TBL_DA0 + w->cur, 6, FLAG_INJECT, 0); break;
      case 13: launch_stencil(b, TBL_DB0 + w->cur, 8, FLAG_INJECT, 1); break;
    }
  }
  return check_launch(m->ctx, "launch_stage");
}

// ------------------------------------------------------------------ batch ABI
extern "C" int32_t fomo_batch_create(fomo_ctx* c, int32_t n, fomo_model* const* ms,
                                     fomo_wavefield* const* ws, fomo_shot* const* ss,
                                     fomo_batch** out) {
  API_LOCK;
  if (!c || !ms || !ws || !ss || !out) return fail(c, FOMO_ERR_ARG, "NULL argument");
  HIP_TRY(c, hipSetDevice(c->device));
  return build_batch(c, n, ms, ws, ss, out);
}

extern "C" int32_t fomo_batch_destroy(fomo_batch* b) {
  API_LOCK;
  if (!b) return FOMO_OK;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  destroy_batch(b);
  return FOMO_OK;
}

extern "C" int32_t fomo_batch_run(fomo_batch* b, int32_t k_first, int32_t n_steps) {
  if (!b) return fail(nullptr, FOMO_ERR_ARG, "NULL batch");
  return run_batch(b, k_first, n_steps);
}

// ------------------------------------------------------------------ pools
// A pool is a batch whose membership changes between runs: the serving
// loop's continuous batching.  Its shot table has `capacity` rows, the live
// members packed at the front; kernels run on fixed persistent grids and
// read the live count and work lists from device memory, so one captured
// graph per pool serves every membership.  Members join at step 1 with their
// own step counter and leave when done; each member's rows in the two fused
// tables are laid out from its own current buffer, so members need not share
// a buffer parity.  Table updates are uploaded on the context's stream
// between runs (never inside a captured graph).
struct fomo_pool {
  fomo_ctx* ctx = nullptr;
  fomo_batch* b = nullptr;
  int capacity = 0;
  std::vector<fomo_model*> ms;
  std::vector<fomo_wavefield*> ws;
  std::vector<fomo_shot*> ss;
  std::vector<int> c0;      // member's current buffer when it joined
  std::vector<int> k_next;  // member's next step (1-based), host view
  hipStream_t stream = nullptr;  // the pool's own stream: pools run beside each other
  ShotDev* h_rows = nullptr;  // pinned staging: [2][capacity] fused-table rows
  int* h_prefix = nullptr;    // pinned staging: [PART_COUNT][capacity+1], then the live count
  hipEvent_t uploaded = nullptr;
  hipEvent_t joined = nullptr;  // the context's stream at the last join
  bool dirty = false;
};

static void pool_free(fomo_pool* p) {
  if (!p) return;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  for (size_t i = 0; i < p->ss.size(); ++i) {  // members still in the pool leave it
    --p->ms[i]->pool_refs;
    --p->ws[i]->pool_refs;
    --p->ss[i]->pool_refs;
  }
  if (p->b) destroy_batch(p->b);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  if (p->h_rows) (void)hipHostFree(p->h_rows);
  if (p->h_prefix) (void)hipHostFree(p->h_prefix);
  if (p->uploaded) (void)hipEventDestroy(p->uploaded);
  if (p->joined) (void)hipEventDestroy(p->joined);
  delete p;
}

extern "C" int32_t fomo_pool_create(fomo_ctx* c, int32_t fd_order, int32_t storage, int32_t capacity,
                                    fomo_pool** out) {
  API_LOCK;
  if (!c || !out) return fail(c, FOMO_ERR_ARG, "NULL argument");
  if (fd_order < 2 || fd_order > 2 * MAX_M || fd_order % 2)
    return fail(c, FOMO_ERR_UNSUPPORTED, "fd_order must be even, 2..16");
  if (storage < FOMO_STORAGE_F32 || storage > FOMO_STORAGE_F24)
    return fail(c, FOMO_ERR_ARG, "unknown storage type");
  const int M = fd_order / 2;
  if (storage != FOMO_STORAGE_F32 && !reduced_m_ok(M))
    return fail(c, FOMO_ERR_UNSUPPORTED, "reduced-precision storage supports fd_order 4, 8 and 12");
  if (capacity < 1 || capacity > 4096) return fail(c, FOMO_ERR_ARG, "capacity must be 1..4096");
  HIP_TRY(c, hipSetDevice(c->device));
  auto p = std::unique_ptr<fomo_pool, void (*)(fomo_pool*)>(new fomo_pool, pool_free);
  p->ctx = c;
  p->capacity = capacity;
  fomo_batch* b = new fomo_batch;
  p->b = b;
  b->ctx = c;
  b->n = capacity;
  b->M = M;
  b->storage = storage;
  b->same = true;
  // until a member without rho_vx / mu_txx blocks joins
  b->compact = storage == FOMO_STORAGE_F32 || storage == FOMO_STORAGE_F24;
  b->cpt = c->cpt;
  b->graph_steps = c->graph_steps;
  b->step_mode = 1;
  b->step_mode_req = 1;
  // rows per tile: about 12 with tz + 2M a whole number of march blocks (14,
  // 12, 16 rows at orders 4, 8, 12).  The served grids are small, and short
  // tiles give more items to fill the GPU: on a 40-request stream of mixed
  // orders this served 16 % more shots per hour than about 32 rows (20-30
  // rows: 11 %; 60-90 rows: 17-30 % fewer), with lower latency.
  const int L = fused_block(M);
  b->tz = c->tz > 0 ? c->tz : std::max(8, round_up(12 + 2 * M, L) - 2 * M);
  b->tz_req = c->tz;
  // fixed persistent grids: one round of resident wavefronts for the stencil
  // kernels, a few blocks per CU for the strip kernels
  const int cus = std::max(1, c->cus);
  b->grid_fixed[PART_FUSED] = round_up(stencil_slots(storage, M, b->cpt, cus, 1, b->compact) / WPB, NUM_XCD);
  b->grid_fixed[PART_BAND] = 2 * cus;
  b->grid_fixed[PART_HABC_V] = 4 * cus;
  b->grid_fixed[PART_HABC_S] = 4 * cus;
  b->grid_fixed[PART_FSREC] = cus;
  b->hs.assign((size_t)TBL_COUNT * capacity, ShotDev{});
  b->hprefix.assign((size_t)PART_COUNT * (capacity + 1), 0);
  HIP_TRY(c, hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
  b->ls = p->stream;
  HIP_TRY(c, hipMalloc(&b->ds, b->hs.size() * sizeof(ShotDev)));
  HIP_TRY(c, hipMemsetAsync(b->ds, 0, b->hs.size() * sizeof(ShotDev), p->stream));
  HIP_TRY(c, hipMalloc(&b->dn, sizeof(int)));
  HIP_TRY(c, hipMemsetAsync(b->dn, 0, sizeof(int), p->stream));
  HIP_TRY(c, hipMalloc(&b->dprefix, b->hprefix.size() * sizeof(int)));
  HIP_TRY(c, hipMemsetAsync(b->dprefix, 0, b->hprefix.size() * sizeof(int), p->stream));
  HIP_TRY(c, hipHostMalloc((void**)&p->h_rows, 2 * (size_t)capacity * sizeof(ShotDev)));
  HIP_TRY(c, hipHostMalloc((void**)&p->h_prefix, (b->hprefix.size() + 1) * sizeof(int)));
  HIP_TRY(c, hipEventCreateWithFlags(&p->uploaded, hipEventDisableTiming));
  HIP_TRY(c, hipEventCreateWithFlags(&p->joined, hipEventDisableTiming));
  HIP_TRY(c, hipEventRecord(p->uploaded, p->stream));
  *out = p.release();
  return FOMO_OK;
}

extern "C" int32_t fomo_pool_destroy(fomo_pool* p) {
  API_LOCK;
  if (!p) return FOMO_OK;
  (void)hipSetDevice(p->ctx->device);
  (void)hipStreamSynchronize(p->ctx->stream);
  pool_free(p);
  return FOMO_OK;
}

static int pool_find(const fomo_pool* p, const fomo_shot* s) {
  for (size_t i = 0; i < p->ss.size(); ++i)
    if (p->ss[i] == s) return (int)i;
  return -1;
}

extern "C" int32_t fomo_pool_add(fomo_pool* p, fomo_model* m, fomo_wavefield* w, fomo_shot* s) {
  API_LOCK;
  if (!p || !m || !w || !s) return fail(p ? p->ctx : nullptr, FOMO_ERR_ARG, "NULL argument");
  fomo_ctx* c = p->ctx;
  if (m->ctx != c || w->m != m || s->m != m)
    return fail(c, FOMO_ERR_STATE, "pool member belongs to another model or context");
  if (m->M != p->b->M || m->storage != p->b->storage)
    return fail(c, FOMO_ERR_UNSUPPORTED, "member's fd_order or storage differs from the pool's");
  if (m->d.dtx != m->d.dtz) return fail(c, FOMO_ERR_UNSUPPORTED, "pool members need dtx == dtz");
  const size_t fb = 5 * m->fstride * m->fes, cb = (size_t)m->ncoef_blocks * m->fstride * m->es;
  if (fb >= 0x7FFFFFFF || cb >= 0x7FFFFFFF)
    return fail(c, FOMO_ERR_UNSUPPORTED, "grid too large for the fused step");
  if ((int)p->ss.size() >= p->capacity) return fail(c, FOMO_ERR_ARG, "pool is full");
  if (w->pool_refs > 0 || s->pool_refs > 0)
    return fail(c, FOMO_ERR_ARG, "wavefield or shot already in a pool");
  int rc = adopt_scales(w, s);
  if (rc) return rc;
  HIP_TRY(c, hipSetDevice(c->device));
  // A member without rho_vx / mu_txx blocks needs the kernel that reads the
  // five stored coefficient arrays; a wider HABC strip needs more LDS rows.
  // Either way the captured step is rebuilt at the next run.
  const bool to_stored = p->b->compact && !m->compact;
  const bool wider = m->d.nbc + 1 > p->b->habc_rows;
  if (to_stored || wider) {
    HIP_TRY(c, hipStreamSynchronize(p->stream));
    fomo_batch* b = p->b;
    for (int q = 0; q < 2; ++q) {
      if (b->gexec[q]) (void)hipGraphExecDestroy(b->gexec[q]);
      if (b->graph[q]) (void)hipGraphDestroy(b->graph[q]);
      b->gexec[q] = nullptr;
      b->graph[q] = nullptr;
    }
    if (to_stored) b->compact = false;
    if (wider) b->habc_rows = m->d.nbc + 1;
  }
  // The member's zeroing (creation, fomo_wavefield_reset, fomo_shot_reset)
  // and any earlier run were queued on the context's stream: the pool's
  // stream waits for them before the member's first step.
  HIP_TRY(c, hipEventRecord(p->joined, c->stream));
  HIP_TRY(c, hipStreamWaitEvent(p->stream, p->joined, 0));
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, p->stream, s->dk, 1);
  ++m->pool_refs;
  ++w->pool_refs;
  ++s->pool_refs;
  p->ms.push_back(m);
  p->ws.push_back(w);
  p->ss.push_back(s);
  p->c0.push_back(w->cur);
  p->k_next.push_back(1);
  p->dirty = true;
  return check_launch(c, "pool_add");
}

extern "C" int32_t fomo_pool_remove(fomo_pool* p, fomo_shot* s) {
  API_LOCK;
  if (!p || !s) return fail(p ? p->ctx : nullptr, FOMO_ERR_ARG, "NULL argument");
  const int i = pool_find(p, s);
  if (i < 0) return fail(p->ctx, FOMO_ERR_ARG, "shot is not in the pool");
  HIP_TRY(p->ctx, hipSetDevice(p->ctx->device));
  HIP_TRY(p->ctx, hipStreamSynchronize(p->stream));  // its last steps have run
  const int done = std::min(p->k_next[i] - 1, s->nt);
  fomo_wavefield* w = p->ws[i];
  w->cur = p->c0[i] ^ (done & 1);
  w->backup_is_prev = done > 0;
  w->prev_buf2 = false;
  --p->ms[i]->pool_refs;
  --w->pool_refs;
  --s->pool_refs;
  p->ms.erase(p->ms.begin() + i);
  p->ws.erase(p->ws.begin() + i);
  p->ss.erase(p->ss.begin() + i);
  p->c0.erase(p->c0.begin() + i);
  p->k_next.erase(p->k_next.begin() + i);
  p->dirty = true;
  return FOMO_OK;
}

extern "C" int32_t fomo_pool_size(const fomo_pool* p) {
  API_LOCK;
  return p ? (int32_t)p->ss.size() : 0;
}

extern "C" int32_t fomo_pool_synchronize(fomo_pool* p) {
  if (!p) return fail(nullptr, FOMO_ERR_ARG, "NULL pool");
  HIP_TRY(p->ctx, hipSetDevice(p->ctx->device));
  HIP_TRY(p->ctx, hipStreamSynchronize(p->stream));
  return FOMO_OK;
}

extern "C" int32_t fomo_pool_next_step(const fomo_pool* p, const fomo_shot* s, int32_t* k_next) {
  API_LOCK;
  if (!p || !s || !k_next) return fail(p ? p->ctx : nullptr, FOMO_ERR_ARG, "NULL argument");
  const int i = pool_find(p, s);
  if (i < 0) return fail(p->ctx, FOMO_ERR_ARG, "shot is not in the pool");
  *k_next = p->k_next[i];
  return FOMO_OK;
}

// Rebuilds the live rows of the two fused tables and the work lists, and
// queues their upload (staging is reused once the previous upload is done).
static int pool_upload(fomo_pool* p) {
  fomo_ctx* c = p->ctx;
  fomo_batch* b = p->b;
  const int C = p->capacity, n = (int)p->ss.size();
  HIP_TRY(c, hipEventSynchronize(p->uploaded));
  for (int t = 0; t < 2; ++t)
    for (int i = 0; i < n; ++i) {
      // table t of the member: reads buffer c0 ^ t, writes the other
      const int table = TBL_FUSED0 + (t ^ p->c0[i]);
      p->h_rows[(size_t)t * C + i] =
          make_shotdev(p->ms[i], p->ws[i], p->ss[i], b->cpt, b->tz, table, BAND_ROWS_BATCH);
    }
  for (int q = 0; q < PART_COUNT; ++q) {
    int acc = 0;
    int* pre = p->h_prefix + (size_t)q * (C + 1);
    for (int i = 0; i < n; ++i) {
      pre[i] = acc;
      const int64_t next = (int64_t)acc + part_items(p->h_rows[i], q);
      if (next > (1 << 30)) return fail(c, FOMO_ERR_ARG, "pool work list too large");
      acc = (int)next;
    }
    for (int i = n; i <= C; ++i) pre[i] = acc;
  }
  p->h_prefix[(size_t)PART_COUNT * (C + 1)] = n;
  std::copy(p->h_prefix, p->h_prefix + (size_t)PART_COUNT * (C + 1), b->hprefix.begin());
  if (n > 0) {
    // table TBL_IN0 only serves the step counters (k_advance reads its kp)
    HIP_TRY(c, hipMemcpyAsync(b->ds + (size_t)TBL_IN0 * C, p->h_rows, (size_t)n * sizeof(ShotDev),
                              hipMemcpyHostToDevice, p->stream));
    HIP_TRY(c, hipMemcpyAsync(b->ds + (size_t)TBL_FUSED0 * C, p->h_rows, (size_t)n * sizeof(ShotDev),
                              hipMemcpyHostToDevice, p->stream));
    HIP_TRY(c, hipMemcpyAsync(b->ds + (size_t)TBL_FUSED1 * C, p->h_rows + C, (size_t)n * sizeof(ShotDev),
                              hipMemcpyHostToDevice, p->stream));
  }
  HIP_TRY(c, hipMemcpyAsync(b->dprefix, p->h_prefix, b->hprefix.size() * sizeof(int),
                            hipMemcpyHostToDevice, p->stream));
  HIP_TRY(c, hipMemcpyAsync(b->dn, p->h_prefix + (size_t)PART_COUNT * (C + 1), sizeof(int),
                            hipMemcpyHostToDevice, p->stream));
  HIP_TRY(c, hipEventRecord(p->uploaded, p->stream));
  p->dirty = false;
  return FOMO_OK;
}

extern "C" int32_t fomo_pool_run(fomo_pool* p, int32_t n_steps) {
  API_LOCK;
  if (!p) return fail(nullptr, FOMO_ERR_ARG, "NULL pool");
  fomo_ctx* c = p->ctx;
  fomo_batch* b = p->b;
  if (n_steps < 0 || n_steps % 2) return fail(c, FOMO_ERR_ARG, "n_steps must be even and >= 0");
  HIP_TRY(c, hipSetDevice(c->device));
  if (p->dirty) {
    int rc = pool_upload(p);
    if (rc) return rc;
  }
  if (p->ss.empty() || n_steps == 0) return FOMO_OK;
  const int G = b->graph_steps;  // even: every chunk starts on table 0; 0: no graphs
  const int chunks = G > 0 ? n_steps / G : 0, rem = G > 0 ? n_steps % G : n_steps;
  if (chunks > 0 && !b->gexec[0]) {
    int rc = capture_chunk(b, 0);
    if (rc) return rc;
  }
  for (int i = 0; i < chunks; ++i) HIP_TRY(c, hipGraphLaunch(b->gexec[0], p->stream));
  for (int g = 0; g < rem; ++g) launch_step(b, g & 1, g);
  if (rem) launch_advance(b, rem);
  for (int& k : p->k_next) k += n_steps;
  return check_launch(c, "pool_run");
}
