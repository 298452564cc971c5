// Device state and operations of the HIP phasing engine (host side).
//
// One Engine holds everything a phaser keeps resident in HBM for one array
// shape: the two iterates, the work buffers the fused passes stream through,
// the encoded amplitudes, the pixel flags and the scratch of the reductions
// and of the shrinkwrap select.  All work is enqueued on the engine's own
// stream; only the explicit download/metric calls synchronise.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <vector>
#include "passes.h"

namespace spx {

// error reporting shared by the whole library
void set_error(const char* fmt, ...);
const char* last_error();

#define SPX_HIP(call)                                                      \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess) {                                                \
      ::spx::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
      return -20;                                                          \
    }                                                                      \
  } while (0)

bool device_available();
// device twiddle table of a supported length (built once per process)
const float2* twiddles(int L);

struct IterParams {
  int alg;            // SpPhasingAlgorithmType
  int constraints;    // SpPhasingConstraints
  int objective;      // SpPhasingObjective
  float gamma1, gamma2;
};

// Radix-select control block kept on the device
enum CtlSlot { kCtlK = 0, kCtlPrefix = 1, kCtlThresh = 2, kCtlMax = 3, kCtlWords = 8 };

class Engine {
 public:
  // Returns null (with spx_last_error set) if the shape is unsupported or
  // device memory is short.
  static Engine* create(int nx, int ny, int nz);
  ~Engine();

  int nx = 0, ny = 0, nz = 0, ndim = 0;
  long N = 0;
  hipStream_t s = nullptr;

  // ---- host <-> device ----
  int upload_amplitudes(const float* amp, const uint8_t* flags);  // also uploads flags
  int upload_flags(const uint8_t* flags);
  int download_flags(uint8_t* flags);
  int upload_margins(const float* amp_min, const float* amp_max);
  int upload_phased(const float2* phased, const uint8_t* flags);
  int upload_model(const float2* model);       // into g1
  int download(const float2* dev, float2* host);
  float2* g1() { return g[cur]; }
  float2* g0() { return g[cur ^ 1]; }
  float2* gp_buf() { return gp; }

  // ---- model initialisation (reference src/phasing.c:606-724) ----
  int init_model_from_amplitudes(const float* amp_raw, int flags, uint64_t seed);
  int init_model_user(const float2* model, int flags);
  int finish_model_init(int flags);  // masked-out zeroing, g0 = gp = 0

  // ---- iterations ----
  // Runs len(betas) iterations; writes gp on the last one when want_gp.
  int iterate(const IterParams& p, const float* betas, int n, bool want_gp);
  void invalidate_R() { R_valid = false; Rf_valid = false; }

  // ---- metrics (reference src/phasing.c:803-893) ----
  int efourier(double out[4], float2* fmodel_out /* device, may be null */);
  int fmodel(float2* fmodel_out);
  int ereal(double out[2]);
  int support_count(double* out);
  int model_change(float2* out_dev);

  // ---- support updates (reference src/support_update.c:136-179) ----
  int shrinkwrap_area(float sigma, float area);
  int shrinkwrap_threshold(float sigma, float rel);
  int patterson_support(const float* amp_raw, float value);
  // threshold used by the last area/threshold/Patterson support update
  int last_threshold(float* t);

  // ---- plain transforms ----
  int fft_nd(const float2* in, float2* out, int dir);
  int blur(const float2* in, float2* out, float sigma);  // sp_gaussian_blur on device data

  // timing of `n` iterations with HIP events on the engine stream
  int time_iterations(const IterParams& p, const float* betas, int n, float* ms);

  unsigned flags_version = 0;  // bumped whenever the support bits change
  uint64_t seed = 0x5EED5EEDull;

 private:
  Engine() = default;
  int alloc_all();
  int ensure_R();
  int ensure_Rf(float gamma1);
  int fourier_pass(int mode_last, const float2* in, float2* out, ColArgs base);
  int reduce_partials(int nparts, int nv, double* host_out);
  int blur_to_abs(float sigma, unsigned* hist, unsigned* maxbits);

  const LengthEntry *ex = nullptr, *ey = nullptr, *ez = nullptr;
  const float2 *twx = nullptr, *twy = nullptr, *twz = nullptr;
  LineGeom gy{}, gz{};

  float2* g[2] = {nullptr, nullptr};
  int cur = 0;
  float2 *W = nullptr, *Wf = nullptr, *gp = nullptr, *S1 = nullptr;
  float *ampx = nullptr, *amp_min = nullptr, *amp_max = nullptr, *absb = nullptr;
  float2* phased = nullptr;
  uint8_t* flags = nullptr;
  double* partials = nullptr;
  int max_partials = 0;
  double* red = nullptr;
  unsigned* hist = nullptr;   // 3 x 2048 bins
  unsigned* ctl = nullptr;    // kCtlWords
  bool have_margins = false, have_phased = false;
  bool R_valid = false, Rf_valid = false;
  unsigned Rf_version = 0;
  float Rf_gamma1 = 0.f;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

// process-wide helper stream/engine cache for the Image-level FFT calls
int fft_host(const float2* in, float2* out, int nx, int ny, int nz, int dir);
int blur_host(const float2* in, float2* out, int nx, int ny, int nz, float sigma);
// transform of device data on the library's stream (synchronous)
int fft_device(const float2* in, float2* out, int nx, int ny, int nz, int dir);

}  // namespace spx
