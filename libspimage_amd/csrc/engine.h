// Device state and operations of the HIP phasing engine (host side).
//
// One Engine holds everything that `nrec` same-shape reconstructions keep
// resident in HBM: the two iterates, the work buffers the fused passes stream
// through, the encoded amplitudes, the pixel flags and the scratch of the
// reductions and of the shrinkwrap select.  Reconstruction r occupies
// elements [r N, (r+1) N) of every array, so a row pass sees nrec*ny*nz lines
// and a column pass nrec outer blocks: a batch costs the same launches as one
// reconstruction.  A phaser (sp_phaser_*) is an Engine with nrec = 1.
//
// All work is enqueued on the engine's own stream; only the explicit
// download / metric calls synchronise.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>
#include <utility>

#include "passes.h"

namespace spx {

// error reporting shared by the whole library
void set_error(const char* fmt, ...);
const char* last_error();

#define SPX_HIP(call)                                                                                  \
  do {                                                                                                 \
    hipError_t e_ = (call);                                                                            \
    if (e_ != hipSuccess) {                                                                            \
      ::spx::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__);     \
      return -20;                                                                                      \
    }                                                                                                  \
  } while (0)

bool device_available();
// Queues a host-to-device copy on stream s through a ring of pinned host
// memory and returns without waiting for the stream (arrays above 16 MiB are
// copied with a wait); the host array may be reused at once.
int upload_h2d(void* dst, const void* src, size_t bytes, hipStream_t s);
// make_shape chooses, once per process and array shape, the plan variant
// (elements per thread) of each axis by timing the compiled variants on the
// shape's geometry.  SPX_FFT_E=4|8|16 forces a variant and SPX_NO_TUNE=1
// keeps the default (8).

struct IterParams {
  int alg;            // SpPhasingAlgorithmType
  int constraints;    // SpPhasingConstraints
  int objective;      // SpPhasingObjective
  float gamma1, gamma2;
};

// Lengths and line geometry of one (nz, ny, nx) shape stacked nrec times.
struct Shape {
  int nx = 0, ny = 0, nz = 0, nrec = 0;
  AxisOps ex, ey, ez;  // per-axis passes with the variants chosen for this shape
  LineGeom gy, gz;
  bool t4 = false;    // 2-D: W, Wf, amplitudes, margins, phased amplitudes in the 4-column tiled layout
  bool colw = false;  // 2-D row-major: the projection pass runs with the wider work-groups
  int nt = 0;         // non-temporal hints: W loads / stores (bits 0, 1; 2-D only), g_k loads /
                      // g_{k+1} stores (bits 2, 3)
  LineGeom gyt;  // y lines of the tiled layout
  // 2-D: the Fourier-space pass of plain iterations (the phase objective
  // without margins, not DM) runs on a transposed copy of W (Wt, columns
  // contiguous; gytr), made and undone by tiled transposes around it
  bool tr = false;
  LineGeom gytr;
};
int make_shape(int nx, int ny, int nz, int nrec, Shape* sh);
// elements per thread (4, 8 or 16; 0 for a trivial axis) of the x, y and z
// transform variants chosen for arrays of shape (nz, ny, nx), tuning the
// shape on its first use
int shape_variant_elements(int nx, int ny, int nz, int* e);
// unnormalised transform of every reconstruction of `sh` (dir < 0 forward,
// e^{-2 pi i k.n/N}); in == out is allowed
void enqueue_fft(const Shape& sh, const float2* in, float2* out, int dir, hipStream_t s);
// forward transform of |in| (as real complex values), the magnitude taken in
// the x pass; needs an x axis longer than 1 (sh.ex)
void enqueue_abs_fft(const Shape& sh, const float2* in, float2* out, hipStream_t s);

class Engine {
 public:
  // Returns null (with the last error set) if the shape is unsupported or
  // device memory is short.
  static Engine* create(int nx, int ny, int nz, int nrec = 1);
  ~Engine();

  int nx = 0, ny = 0, nz = 0, nrec = 1;
  long N = 0;   // pixels of one reconstruction
  long NT = 0;  // pixels of all reconstructions
  hipStream_t s = nullptr;

  // ---- host <-> device (rec < 0: all reconstructions, host arrays of NT) ----
  // uploads return without waiting for the stream (upload_h2d)
  int upload_amplitudes(const float* amp, const uint8_t* flags, int rec = -1);  // also uploads flags
  int upload_flags(const uint8_t* flags, int rec = -1);
  int download_flags(uint8_t* flags, int rec = -1);
  int upload_margins(const float* amp_min, const float* amp_max);
  int upload_phased(const float2* phased, const uint8_t* flags);
  int upload_model(const float2* model, int rec = -1);  // into g1
  int download(const float2* dev, float2* host, int rec = -1);
  bool tiled() const { return sh.t4; }  // Fourier-space arrays in the 4-column tiled layout
  int layout() const { return sh.tr ? 2 : sh.t4 ? 1 : 0; }  // 0 row-major, 1 tiled, 2 transposed pass
  int nt_hints() const { return sh.nt; }
  float2* g1() { return g[cur]; }
  float2* g0() { return g[cur ^ 1]; }
  float2* gp_buf() { return gp; }
  float2* scratch() { return S1; }
  // per-reconstruction algorithm / beta tables (null: uniform)
  int set_rec_params(const int* alg, const float* beta);
  // per-reconstruction activity (nrec entries, 0 = idle: iterations leave
  // that reconstruction unchanged); null or all non-zero: every one active
  int set_rec_active(const int* active);

  // ---- model initialisation (reference src/phasing.c:606-724) ----
  // rec < 0: every reconstruction, amp_raw holds NT values and reconstruction
  // r draws from seed + r; rec >= 0: that reconstruction only, amp_raw holds
  // its N values and it draws from seed.  The generator is indexed by the
  // pixel's position inside its reconstruction, so a reconstruction seeded
  // with s starts the same in any batch and alone.
  int init_model_from_amplitudes(const float* amp_raw, int flags, uint64_t seed, int rec = -1);
  int init_model_user(const float2* model, int flags);
  int finish_model_init(int flags, int rec = -1);  // masked-out zeroing, g0 = gp = 0

  // ---- iterations ----
  // Runs n iterations with betas[i] (ignored when per-reconstruction betas
  // are set); writes gp on the last one when want_gp.
  int iterate(const IterParams& p, const float* betas, int n, bool want_gp);
  // n iterations (no gp) replayed from HIP graphs of one iteration, one per
  // parity of the ping-pong iterates; recaptured when p or beta change
  int iterate_graph(const IterParams& p, float beta, int n);
  // W (= FFT_x of the current iterate) must be recomputed: for every
  // reconstruction (rec < 0) or only for `rec`, whose iterate was replaced.
  // The others keep the W the fused update kernel wrote, so a reconstruction
  // that shares a batch with newly started ones follows bit for bit the
  // trajectory it follows alone.
  void invalidate_R(int rec = -1) {
    Rf_valid = false;
    if (rec < 0) R_valid = false;
    else R_stale.push_back(rec);
  }

  // ---- metrics (reference src/phasing.c:803-893), nrec results each ----
  int efourier(double* out /* 4 per rec */, float2* fmodel_out /* device, may be null */);
  // the same sums, enqueued behind the work on the engine stream into pinned
  // host memory; efourier_fetch returns them once the stream has reached them
  int efourier_enqueue();
  int efourier_fetch(double* out /* 4 per rec */);
  int fmodel(float2* fmodel_out);
  int ereal(double* out /* 2 per rec */);
  int support_count(double* out /* 1 per rec */);
  int model_change(float2* out_dev);

  // ---- support updates (reference src/support_update.c:136-179) ----
  int shrinkwrap_area(float sigma, float area);
  int shrinkwrap_threshold(float sigma, float rel);
  int patterson_support(const float* amp_raw, float value);
  // closing of the support: `size` dilations, then `size` erosions, with the
  // 6-neighbour cross (reference src/support_update.c:213-280)
  int support_close(int size);
  // centre image: the support moved cyclically so that its centre of mass
  // sits at the origin (reference src/support_update.c:282-389), for every
  // reconstruction, on the device
  int support_centre_image();
  // support = tmpl_abs > t (reference src/support_update.c:181-207).  The N
  // host values of tmpl_abs are uploaded on the first call with `key` and
  // kept on the device for the engine's lifetime.
  int support_template(unsigned long long key, const float* tmpl_abs, float t);
  // threshold used by the last area/threshold/Patterson update (per rec)
  int last_threshold(float* t);

  // ---- plain transforms ----
  int fft_nd(const float2* in, float2* out, int dir);

  // timing of `n` iterations with HIP events on the engine stream
  int time_iterations(const IterParams& p, const float* betas, int n, float* ms);
  // the same iterations with events around each pass: total milliseconds of
  // the Fourier-space (column) passes and of the real-space (row) passes
  int time_passes(const IterParams& p, const float* betas, int n, float* ms_fourier, float* ms_real);

  unsigned flags_version = 0;  // bumped whenever the support bits change
  uint64_t seed = 0x5EED5EEDull;
  // an axis has no Stockham plan: single reconstructions only, iterations
  // unfused over Bluestein transforms (engine.hip, "Transforms of any axis length")
  bool any_mode = false;

 private:
  Engine() = default;
  int alloc_all();
  int iterate_impl(const IterParams& p, const float* betas, int n, bool want_gp, hipEvent_t* marks);
  int ensure_R();
  int ensure_Rf(float gamma1);
  int reduce_per_rec(int parts_per_rec, int nv, double* host_out, int count = -1);
  int blur_to_abs(float sigma, bool hist);
  // the full transform of one or all reconstructions (shp: sh or one)
  int fft_full(const Shape& shp, const float2* in, float2* out, int dir);
  int iterate_any(const IterParams& p, const float* betas, int n, bool want_gp, hipEvent_t* marks);
  float2* Sany = nullptr;  // Bluestein scratch of the any-length mode
  // the transposed Fourier-space pass (sh.tr): Wt and the amplitudes in
  // its layout, refreshed when the amplitudes change
  int ensure_transposed();
  float2* Wt = nullptr;
  float* ampt = nullptr;
  unsigned amp_version = 0, ampt_version = ~0u;

  Shape sh{};
  Shape one{};  // the same shape with one reconstruction (per-slot transforms)
  float2* g[2] = {nullptr, nullptr};
  int cur = 0;
  float2 *W = nullptr, *Wf = nullptr, *gp = nullptr, *S1 = nullptr;
  float *ampx = nullptr, *amp_min = nullptr, *amp_max = nullptr, *absb = nullptr;
  float2* phased = nullptr;
  uint8_t* flags = nullptr;
  double* partials = nullptr;
  double* red = nullptr;
  double* red_host = nullptr;  // pinned, 4 per rec (efourier_enqueue)
  bool red_host_pending = false;
  unsigned* hist = nullptr;   // kHistStride per rec
  unsigned* ctl = nullptr;    // kCtlStride per rec
  int* alg_rec = nullptr;     // optional per-rec tables
  float* beta_rec = nullptr;
  uint8_t* rec_active = nullptr;
  std::vector<std::pair<unsigned long long, float*>> templates;  // support templates by key
  bool some_idle = false;
  bool R_valid = false, Rf_valid = false;
  std::vector<int> R_stale;  // reconstructions whose W is out of date
  unsigned Rf_version = 0;
  float Rf_gamma1 = 0.f;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  IterParams gkey{};
  float gbeta = 0.f;
  bool gvalid = false;
  void drop_graphs();
};

// one-shot transforms for the Image-level API (host buffers)
int fft_host(const float2* in, float2* out, int nx, int ny, int nz, int dir);
int blur_host(const float2* in, float2* out, int nx, int ny, int nz, float sigma);
// transform of device data on the library's stream (synchronous)
int fft_device(const float2* in, float2* out, int nx, int ny, int nz, int dir);
// The one-shot and device transforms accept any axis length up to
// kBluesteinMaxL besides the Stockham lengths (Bluestein's method over a
// power-of-two Stockham transform of length <= 8192); single phasers run
// such shapes with unfused iterations, batches need Stockham lengths.
constexpr int kBluesteinMaxL = 4096;
bool fft_length_supported(int L);
// fft_host staged through a fixed device array instead of an allocation:
// at most kFixedElems elements (-12 beyond), untuned default plans
constexpr long kFixedElems = 1L << 21;
int fft_fixed(const float2* in, float2* out, int nx, int ny, int nz, int dir);

}  // namespace spx
