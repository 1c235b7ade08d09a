// Host-side context of libspd_hip.so (one simulation: configuration,
// geometry, operator matrices, scratch buffers), shared by the step and its
// adjoint translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <string>
#include <vector>

#include "../../include/spd.h"
#include "spd_device.hpp"
#include "spd_host_ops.hpp"
#include "spd_layout.hpp"

// Record `msg` as the calling thread's last error and return `code`.
int fail(int code, const std::string& msg);

#define SPD_HIP_CHECK(expr)                                                          \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    if (_e != hipSuccess)                                                            \
      return fail(SPD_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));     \
  } while (0)

inline unsigned grid_for(long long n, int bs = 256) {
  return (unsigned)((n + bs - 1) / bs);
}

namespace spd {

// Device-side time loop (src/sdader_simulator.py:394-413): time, dt, step
// count and the stop condition live on the device so that chunks of steps can
// be captured in a graph and replayed without host synchronisation.
struct RunState {
  double time, t_end, dt;
  long long n_step;
  long long n_left;  // steps still allowed (< 0: unlimited)
  int done;
  int clamp;         // clamp dt to reach t_end and stop when dt < 1e-14
};

}  // namespace spd

using namespace spd;

// Kernel classes timed by the optional HIP-event profiler.
enum KernelClass : int {
  KC_FACE_SD = 0, KC_PRED, KC_HOFLUX, KC_TENSOR, KC_APPLY, KC_DETECT, KC_THETA1, KC_RING,
  KC_FBFLUX, KC_CV2SP, KC_HANCOCK, KC_OTHER, KC_COUNT
};
struct spd_ctx {
  spd_config cfg{};
  // HIP-event profiler: one event pair per recorded launch
  bool prof_on = false;
  std::vector<hipEvent_t> prof_ev;
  std::vector<int> prof_cls;
  size_t prof_used = 0;
  int ndim = 0, p = 0, n = 0, nf = 0, T = 0, V = 0, NL = 0, NLP = 0;
  Geom g{};
  Params prm{};
  std::unique_ptr<host::OpsHost> ops;
  double h_dt = 0.0;  // h_min / (p+1)
  int device = 0;
  std::vector<void*> allocs;
  size_t scratch_bytes = 0;
  double* geom_buf = nullptr;
  // bulk-type buffers (FP64, or FP32 in the FP32-bulk form)
  void* Uad = nullptr;
  void* Uad2 = nullptr;       // second predictor buffer of the fused 2-D iteration
  void* Uad_cur = nullptr;    // predictor state of the current step's last iteration
  void* Fface[3] = {nullptr, nullptr, nullptr};
  void* Fho[3] = {nullptr, nullptr, nullptr};
  double* Ucv = nullptr;
  double* Unew = nullptr;
  uint8_t* flags = nullptr;
  double* theta1 = nullptr;
  double* theta = nullptr;
  void* HS = nullptr;         // Hancock states and half slopes, (1+ndim)*V per cell (bulk)
  double* rp = nullptr;       // density and pressure of the candidate (1-D/3-D detection)
  double* Wold = nullptr;     // primitives of the state entering the substep (1-D/3-D)
  int* fbt_counters = nullptr;  // compacted 2-D fallback: list length, next entry, path
  int* fbt_list = nullptr;      // active tiles of the compacted 2-D fallback
  uint8_t* fbt_marks = nullptr; // tiles theta marked active (cleared by the list kernel)
  uint8_t* theta_dirty = nullptr;  // 2-D theta tiles that may hold non-zero values (k_theta2d)
  long long theta_tiles = 0;
  int* any_flag = nullptr;
  int* zero_flag = nullptr;   // always 0: the skip flag of ordinary calls
  int* run_skip = nullptr;    // skip flag of device-loop steps
  RunState* run = nullptr;
  double* run_dt = nullptr;
  double* cmax = nullptr;     // max wave speed of the last state (for dt)
  double* partial = nullptr;  // per-block partial maxima
  long long npartial = 0;
  bool need_gate = false;     // some side is not periodic (global NAD gate needed)
  bool bc_ready = false;      // the stored boundary values of ic / pressure sides are set
  int launch[SPD_LAUNCH_COUNT] = {0, 0, 0};  // tuned launch shapes (0 = default)
  // debug tap
  int dbg_sub = -1;
  double* dbg_tr = nullptr;
  double* dbg_th = nullptr;
  double* dbg_aff[3] = {nullptr, nullptr, nullptr};
  // tape of one step for the reverse-mode derivative (allocated by the adjoint
  // on first use): with tape_on the step writes its m predictor iterates
  // (m, T, V, E, L; bulk type), the FV state before each substep (T, V, Nc)
  // and each substep's blending coefficients on the ghosted grid (T, Ngc);
  // with tape_replay the coefficients are read from the tape instead of
  // being detected.
  int tape_on = 0;
  int tape_replay = 0;
  void* tape_uad = nullptr;
  double* tape_ucv = nullptr;
  double* tape_theta = nullptr;
  struct AdjScratch* adj = nullptr;

  ~spd_ctx() {
    adj_release();
    for (hipEvent_t e : prof_ev) (void)hipEventDestroy(e);
    for (void* p : allocs) (void)hipFree(p);
  }
  // Launch `fn` on stream s, bracketed by an event pair when profiling.
  template <typename F> void timed(int cls, hipStream_t s, F&& fn) {
    const bool rec = prof_on && prof_used < prof_cls.size();
    if (rec) (void)hipEventRecord(prof_ev[2 * prof_used], s);
    fn();
    if (rec) {
      (void)hipEventRecord(prof_ev[2 * prof_used + 1], s);
      prof_cls[prof_used] = cls;
      ++prof_used;
    }
  }
  void adj_release();  // frees the adjoint's host-side state (spd_adjoint.hip)
  template <typename X> int alloc(X** ptr, size_t count) {
    void* p = nullptr;
    const size_t bytes = std::max<size_t>(count * sizeof(X), 256);
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess)
      return fail(SPD_E_NOMEM, std::string("hipMalloc(") + std::to_string(bytes) +
                                   "): " + hipGetErrorString(e));
    allocs.push_back(p);
    scratch_bytes += bytes;
    *ptr = (X*)p;
    return SPD_OK;
  }
};

namespace spd {

template <int P, typename R = double> Ops<P, R> make_ops(const host::OpsHost& h) {
  Ops<P, R> o{};
  constexpr int n = P + 1, nf = P + 2, T = P + 1;
  for (int a = 0; a < nf; ++a)
    for (int b = 0; b < n; ++b) o.s2f[a][b] = (R)h.sp_to_fp[a][b];
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < nf; ++b) o.d2s[a][b] = (R)h.dfp_to_sp[a][b];
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) {
      o.s2c[a][b] = (R)h.sp_to_cv[a][b];
      o.c2s[a][b] = (R)h.cv_to_sp[a][b];
      o.s2r[a][b] = (R)h.sp_to_reg[a][b];
    }
  for (int a = 0; a < T; ++a) {
    for (int b = 0; b < T; ++b) o.inv[a][b] = (R)h.invader[a][b];
    o.w[a] = h.w_tp[a];
  }
  return o;
}

}  // namespace spd
