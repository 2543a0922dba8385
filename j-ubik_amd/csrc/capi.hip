// C ABI (include/jubik_hip.h): plans, argument validation, error text and
// the fused gradient evaluation / metric product built from the kernels.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/jubik_hip.h"
#include "cg.hpp"
#include "gp.hpp"
#include "points.hpp"
#include "linpatch.hpp"
#include "dft128.hpp"
#include "primitives.hpp"
#include "psf.hpp"
#include "response.hpp"

namespace jubik {

static thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }
int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}
int check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return kOk;
  g_error = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? kOutOfMemory : kRuntimeError;
}

namespace {
struct TimingRec {
  const char* name;
  hipEvent_t a, b;
};
std::atomic<bool> g_timing{false};
std::mutex g_timing_mu;
std::vector<TimingRec> g_timing_recs;
std::string g_timing_only;  // empty: every kernel
}  // namespace

bool timing_enabled() { return g_timing.load(std::memory_order_relaxed); }
bool timing_wanted(const char* name) {
  if (!timing_enabled()) return false;
  std::lock_guard<std::mutex> lk(g_timing_mu);
  return g_timing_only.empty() || g_timing_only == name;
}
void timing_record(const char* name, hipEvent_t a, hipEvent_t b) {
  std::lock_guard<std::mutex> lk(g_timing_mu);
  g_timing_recs.push_back({name, a, b});
}

namespace {

// Keeps the plan's device current for the duration of a call.
struct DeviceGuard {
  int prev = -1;
  bool ok = true;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) ok = hipSetDevice(dev) == hipSuccess;
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

size_t elem_size(int dtype) { return dtype == 0 ? sizeof(float) : sizeof(double); }

// Fault injection (jubik_debug_fail_alloc): the n-th next device allocation
// fails as if the device were out of memory; 0 = off.
std::atomic<int64_t> g_fail_alloc{0};

int dev_alloc(void** p, size_t bytes, const char* what) {
  *p = nullptr;
  if (g_fail_alloc.load() > 0 && g_fail_alloc.fetch_sub(1) == 1)
    return fail(kOutOfMemory, std::string("allocating ") + what + ": injected failure");
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(e == hipErrorOutOfMemory ? kOutOfMemory : kRuntimeError,
                std::string("allocating ") + what + " (" + std::to_string(bytes) +
                    " bytes): " + hipGetErrorString(e));
  }
  return kOk;
}

void dev_free(void*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// Twiddles exp(-2 pi i e / L) in the plan's precision.
int upload_twiddles(void** dst, int dtype, int L) {
  const size_t es = elem_size(dtype);
  std::vector<unsigned char> host(2 * es * L);
  for (int e = 0; e < L; ++e) {
    const long double ang = -2.0L * 3.141592653589793238462643383279502884L * e / L;
    const double c = (double)cosl(ang), s = (double)sinl(ang);
    if (dtype == 0) {
      float* h = reinterpret_cast<float*>(host.data());
      h[2 * e] = (float)c;
      h[2 * e + 1] = (float)s;
    } else {
      double* h = reinterpret_cast<double*>(host.data());
      h[2 * e] = c;
      h[2 * e + 1] = s;
    }
  }
  int st = dev_alloc(dst, host.size(), "twiddles");
  if (st) return st;
  JUBIK_HIP(hipMemcpy(*dst, host.data(), host.size(), hipMemcpyHostToDevice));
  return kOk;
}

// Tent weights of jubik/convolve.py:18-25: numpy linspace(0, 1, d) (i * step,
// last entry exactly 1) followed by its mirror image.
int upload_tent(void** dst, int dtype, int d) {
  std::vector<double> b(2 * d);
  const double step = d > 1 ? 1.0 / (d - 1) : 0.0;
  for (int i = 0; i < d; ++i) {
    double v = (d > 1 && i == d - 1) ? 1.0 : i * step;
    b[i] = v;
    b[2 * d - 1 - i] = v;
  }
  const size_t es = elem_size(dtype);
  std::vector<unsigned char> host(es * 2 * d);
  for (int i = 0; i < 2 * d; ++i) {
    if (dtype == 0) reinterpret_cast<float*>(host.data())[i] = (float)b[i];
    else reinterpret_cast<double*>(host.data())[i] = b[i];
  }
  int st = dev_alloc(dst, host.size(), "tent weights");
  if (st) return st;
  JUBIK_HIP(hipMemcpy(*dst, host.data(), host.size(), hipMemcpyHostToDevice));
  return kOk;
}

}  // namespace
}  // namespace jubik

using namespace jubik;

struct jubik_linpatch {
  int dtype = 0;
  int device = 0;
  LinGeom g{};
  void* tw = nullptr;
  void* tent = nullptr;
  void* KT = nullptr;
  void* R1 = nullptr;
  void* R2 = nullptr;
  void* KB = nullptr;
  size_t ws_bytes = 0, spec_bytes = 0;
  bool ready = false;
  int col_div = 1;  // column- and row-pass tile divisors (set by the autotuner)
  int row_div = 1;
  int transform = 0;       // column-pass transform: 0 FFT, 1 split-BF16 matrix-core DFT
  void* frag = nullptr;    // DFT fragments of transform 1 (allocated on first use)
  LinBuffers bufs() const {
    return LinBuffers{tw, tent, KT, R1, R2, KB, col_div, row_div, transform == 1 ? frag : nullptr};
  }
  void release() {
    dev_free(frag);
    dev_free(tw);
    dev_free(tent);
    dev_free(KT);
    dev_free(R1);
    dev_free(R2);
    dev_free(KB);
  }
};

struct jubik_gp {
  int dtype = 0;
  int device = 0;
  GpGeom g{};
  void* tw = nullptr;
  void* work = nullptr;
  void* u = nullptr;
  GpBuffers bufs() const { return GpBuffers{tw, work}; }
  void release() {
    dev_free(tw);
    dev_free(work);
    dev_free(u);
  }
};

// A model owns the per-evaluation buffers.  Slot 0 of sky / lam holds the
// base point of the metric (the last jubik_model_grad); batched gradients
// use slots 1 .. S.  With max_samples > 1 the model also owns linpatch and GP
// workspaces for S jobs sets; with max_samples = 1 it borrows the plans'.
struct jubik_model {
  jubik_gp* gp = nullptr;
  jubik_linpatch* lp = nullptr;
  double pixel_area = 1.0;
  int max_samples = 1;
  const void* expo = nullptr;
  const int32_t* counts = nullptr;
  void* sky = nullptr;      // (1 + max_samples) x E x H x W
  void* lam = nullptr;      // (1 + max_samples) x J x H x W
  void* dsky = nullptr;     // max_samples x E x H x W
  void* resid = nullptr;    // max_samples x J x H x W
  void* partial = nullptr;  // 2 x max_samples x J x H doubles (NLL terms, lam <= 0 counts)
  void* nbad = nullptr;     // max_samples doubles: read-out pixels with lam <= 0
  int last_S = 0;           // samples of the last gradient evaluation
  int batch_S = 0;          // samples of the last batched gradient (slots 1 .. S)
  void* cg_vec = nullptr;   // solver vectors r, p, Ap, Mv: 4 x max_samples x 2 x N^2
  void* cg_work = nullptr;  // solver scalars and partial sums (doubles)
  void* u = nullptr;        // max_samples x E x H x W (cotangent planes of log s)
  void* R1 = nullptr;       // own linpatch workspaces (max_samples > 1)
  void* R2 = nullptr;
  void* gwork = nullptr;    // own GP workspace (max_samples > 1)
  bool have_point = false;
  int col_div = 0;  // column- and row-pass tile divisors; 0 = the plan's
  int row_div = 0;
  // point sources (jubik_model_set_point_sources)
  bool ps_on = false;
  PointSources ps{};
  void* skyt = nullptr;     // max_samples x E x H x W: diffuse + point sources (the PSF input)
  void* ct = nullptr;       // max_samples x E x H x W: cotangent of that total sky
  void* psw = nullptr;      // max_samples x 3 x H x W: log I, d log I / d xi, alpha
  // geoVI nonlinear sample updates (jubik_model_geovi_*): the sample point x
  // lives in slot 1; geo holds sqrt(lam_x lam_p) and eta = 2 (sqrt lam_x -
  // sqrt lam_p) (2 x J x H x W); geo_vec three latent vectors (2 x N^2 each)
  void* geo = nullptr;
  void* geo_vec = nullptr;
  bool geo_ready = false;
  void release() {
    dev_free(geo);
    dev_free(geo_vec);
    dev_free(skyt);
    dev_free(ct);
    dev_free(psw);
    dev_free(sky);
    dev_free(dsky);
    dev_free(resid);
    dev_free(lam);
    dev_free(partial);
    dev_free(nbad);
    dev_free(cg_vec);
    dev_free(cg_work);
    dev_free(u);
    dev_free(R1);
    dev_free(R2);
    dev_free(gwork);
  }
  LinBuffers lbufs() const {
    LinBuffers b = lp->bufs();
    if (R1) b.R1 = R1;
    if (R2) b.R2 = R2;
    if (col_div) b.col_div = col_div;
    if (row_div) b.row_div = row_div;
    return b;
  }
  GpBuffers gbufs() const {
    GpBuffers b = gp->bufs();
    if (gwork) b.work = gwork;
    return b;
  }
  size_t img_bytes() const {
    return (size_t)lp->g.H * lp->g.W * (lp->dtype == 0 ? sizeof(float) : sizeof(double));
  }
  void* sky_slot(int k) const { return (char*)sky + (size_t)k * lp->g.E * img_bytes(); }
  void* lam_slot(int k) const { return (char*)lam + (size_t)k * lp->g.J * img_bytes(); }
};

static hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// An executable HIP graph of one fused gradient evaluation with fixed
// buffers; ``single`` graphs refresh the model's metric base point.
struct jubik_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<jubik_model*> points;  // models whose single evaluation sets the base point
  int device = 0;
};

extern "C" {

const char* jubik_last_error(void) { return g_error.c_str(); }
int jubik_abi_version(void) { return 1; }

int jubik_timing_enable(int on) {
  g_timing.store(on != 0);
  return kOk;
}

int jubik_debug_fail_alloc(int64_t nth) {
  if (nth < 0) return fail(kInvalidArgument, "nth must be non-negative");
  g_fail_alloc.store(nth);
  return kOk;
}

int jubik_timing_select(const char* name) {
  std::lock_guard<std::mutex> lk(g_timing_mu);
  g_timing_only = name ? name : "";
  return kOk;
}

int jubik_timing_report(char* buf, int64_t len) {
  if (!buf || len <= 0) return fail(kInvalidArgument, "NULL or empty buffer");
  std::vector<TimingRec> recs;
  {
    std::lock_guard<std::mutex> lk(g_timing_mu);
    recs.swap(g_timing_recs);
  }
  struct Agg {
    std::string name;
    int64_t count = 0;
    double total = 0, lo = 1e300, hi = 0;
  };
  std::vector<Agg> aggs;
  int st = kOk;
  for (auto& r : recs) {
    float ms = 0.f;
    if (st == kOk) {
      hipError_t e = hipEventSynchronize(r.b);
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, r.a, r.b);
      if (e != hipSuccess) st = check_hip(e, "timing report");
    }
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
    if (st != kOk) continue;
    Agg* a = nullptr;
    for (auto& x : aggs)
      if (x.name == r.name) a = &x;
    if (!a) {
      aggs.push_back(Agg{r.name});
      a = &aggs.back();
    }
    a->count += 1;
    a->total += ms;
    a->lo = std::min(a->lo, (double)ms);
    a->hi = std::max(a->hi, (double)ms);
  }
  if (st != kOk) return st;
  std::string out;
  for (auto& a : aggs)
    out += a.name + " " + std::to_string(a.count) + " " + std::to_string(a.total) + " " +
           std::to_string(a.lo) + " " + std::to_string(a.hi) + "\n";
  if ((int64_t)out.size() + 1 > len) return fail(kInvalidArgument, "report buffer too small");
  std::memcpy(buf, out.c_str(), out.size() + 1);
  return kOk;
}

// ------------------------------------------------------------------ linpatch
int jubik_linpatch_create(jubik_linpatch** plan, int dtype, int64_t H, int64_t W, int64_t E,
                          int64_t M, int64_t n, int64_t margin, int kernel_per_energy,
                          int x_shared) {
  if (!plan) return fail(kInvalidArgument, "plan pointer is NULL");
  *plan = nullptr;
  if (dtype != 0 && dtype != 1) return fail(kInvalidArgument, "dtype must be 0 or 1");
  if (H < 2 || W < 2 || E < 1 || M < 1 || n < 1)
    return fail(kInvalidArgument, "H, W >= 2 and E, M, n_patches_per_axis >= 1 required");
  if (margin < 1) return fail(kInvalidArgument, "margin must be positive");
  if (n > 64) return fail(kUnsupported, "at most 64 patches per axis are supported");
  if (W > 4096) return fail(kUnsupported, "grids wider than 4096 pixels are not supported");
  const int64_t dx = H / n, dy = W / n;
  if (dx < 1) return fail(kInvalidArgument, "more patches per axis than pixels");
  if (dx != dy)
    return fail(kInvalidArgument,
                "patch strides differ along the two axes (square windows are required)");
  LinGeom g{};
  g.H = (int)H;
  g.W = (int)W;
  g.n = (int)n;
  g.d = (int)dx;
  g.h = (int)(dx / 2);
  g.w = (int)(2 * dx);
  g.m = (int)margin;
  g.Lw = g.w + 2 * g.m;
  g.L = linpatch_fft_length(g.w, g.m);
  g.mf = (g.L - g.w) / 2;
  g.L2 = g.L / 2;
  g.limx = g.H + 2 * g.h - g.w;
  g.limy = g.W + 2 * g.h - g.w;
  g.P = g.n * g.n;
  g.E = (int)E;
  g.M = (int)M;
  g.J = g.M * g.E;
  g.x_shared = x_shared ? 1 : 0;
  g.Ek = kernel_per_energy ? g.E : 1;
  g.S = 1;
  g.sky_shared = 1;
  if (g.limx < 0 || g.limy < 0)
    return fail(kInvalidArgument, "patch window of " + std::to_string(g.w) +
                                      " does not fit the padded grid");
  if (g.L == 0)
    return fail(kUnsupported, "no compiled FFT length for window " + std::to_string(g.w) +
                                  " + 2 * margin " + std::to_string(g.m) +
                                  " (at most 512, or twice the window in fold mode)");
  auto* p = new jubik_linpatch();
  p->dtype = dtype;
  p->g = g;
  if (hipGetDevice(&p->device) != hipSuccess) {
    delete p;
    return fail(kRuntimeError, "no HIP device available");
  }
  const size_t es = elem_size(dtype);
  const size_t cs = 2 * es;
  const size_t nspec = (size_t)g.M * g.P * g.Ek;
  const size_t l2p = (size_t)half_stride(g.L);
  p->spec_bytes = nspec * l2p * g.L * cs;
  p->ws_bytes = (size_t)g.J * g.P * l2p * g.L * cs;
  int st = upload_twiddles(&p->tw, dtype, g.L);
  if (!st) st = upload_tent(&p->tent, dtype, g.d);
  if (!st) st = dev_alloc(&p->KT, p->spec_bytes, "kernel spectra");
  if (!st) st = dev_alloc(&p->R1, std::max(p->ws_bytes, nspec * l2p * g.w * cs), "workspace R1");
  if (!st) st = dev_alloc(&p->R2, p->ws_bytes, "workspace R2");
  if (!st) st = dev_alloc(&p->KB, nspec * g.L * cs, "packed-column kernel factors");
  if (st) {
    p->release();
    delete p;
    return st;
  }
  *plan = p;
  return kOk;
}

int jubik_linpatch_destroy(jubik_linpatch* plan) {
  if (!plan) return kOk;
  DeviceGuard dg(plan->device);
  plan->release();
  delete plan;
  return kOk;
}

int jubik_linpatch_info(const jubik_linpatch* plan, int64_t* out) {
  if (!plan || !out) return fail(kInvalidArgument, "NULL argument");
  const LinGeom& g = plan->g;
  out[0] = g.d;
  out[1] = g.w;
  out[2] = g.L;
  out[3] = g.P;
  out[4] = g.h;
  out[5] = g.J;
  out[6] = (int64_t)(2 * plan->ws_bytes);
  out[7] = (int64_t)plan->spec_bytes;
  return kOk;
}

int jubik_linpatch_set_kernel(jubik_linpatch* plan, const void* kernel, int layout,
                              void* stream) {
  if (!plan || !kernel) return fail(kInvalidArgument, "NULL argument");
  if (layout != 0 && layout != 1) return fail(kInvalidArgument, "layout must be 0 or 1");
  const LinGeom& g = plan->g;
  if (layout == 0) {
    // convolve.py:107-111: the centred cut is empty when (H - 2dx) // 2 == 0
    // and has the wrong size for odd frames; the reference fails there too.
    if ((g.H - g.w) / 2 == 0 || (g.W - g.w) / 2 == 0)
      return fail(kInvalidArgument, "kernel cut is empty (frame of twice the patch stride)");
    if (g.H % 2 || g.W % 2)
      return fail(kInvalidArgument, "full-frame kernels need an even grid");
  }
  DeviceGuard dg(plan->device);
  int st = linpatch_prepare(plan->dtype, g, plan->bufs(), kernel, layout, as_stream(stream));
  if (st == kOk) plan->ready = true;
  return st;
}

int jubik_linpatch_forward(jubik_linpatch* plan, const void* x, void* y, double scale,
                           void* stream) {
  if (!plan || !x || !y) return fail(kInvalidArgument, "NULL argument");
  if (!plan->ready) return fail(kNotReady, "no kernel set (call jubik_linpatch_set_kernel)");
  DeviceGuard dg(plan->device);
  return linpatch_forward(plan->dtype, plan->g, plan->bufs(), x, y, scale, as_stream(stream));
}

int jubik_linpatch_adjoint(jubik_linpatch* plan, const void* gy, void* gx, double scale,
                           void* stream) {
  if (!plan || !gy || !gx) return fail(kInvalidArgument, "NULL argument");
  if (!plan->ready) return fail(kNotReady, "no kernel set (call jubik_linpatch_set_kernel)");
  DeviceGuard dg(plan->device);
  return linpatch_adjoint(plan->dtype, plan->g, plan->bufs(), gy, gx, scale, as_stream(stream));
}

// Times one linpatch forward + adjoint (S samples, the given workspaces) on
// zero inputs for every column-tile divisor compiled in for L (row tile 1),
// then for every row-tile divisor with the fastest column tile, and returns
// the fastest pair; ms[0..2] receive the times of column divisors 1, 2, 4
// and ms[3..5] those of row divisors 1, 2, 4 (-1 where not compiled in).
// Set-up work: allocates and synchronises.
static int autotune_tiles(int dtype, const LinGeom& g0, LinBuffers b, int S, int reps,
                          int* best_col, int* best_row, double* ms, hipStream_t s) {
  LinGeom g = g0;
  g.S = S;
  const size_t bytes = (size_t)S * g.J * g.H * g.W * elem_size(dtype);
  void *x = nullptr, *y = nullptr;
  int st = dev_alloc(&x, bytes, "autotune input");
  if (!st) st = dev_alloc(&y, bytes, "autotune output");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!st && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    st = fail(kRuntimeError, "creating timing events");
  if (!st && (hipMemsetAsync(x, 0, bytes, s) != hipSuccess ||
              hipMemsetAsync(y, 0, bytes, s) != hipSuccess))
    st = fail(kRuntimeError, "clearing autotune buffers");
  auto time_tiles = [&](double& out) {
    auto once = [&]() {
      int r = linpatch_forward(dtype, g, b, x, y, 1.0, s);
      if (!r) r = linpatch_adjoint(dtype, g, b, y, x, 1.0, s);
      return r;
    };
    int r = 0;
    for (int w = 0; w < 2 && !r; ++w) r = once();
    if (r) return r;
    if (hipEventRecord(e0, s) != hipSuccess) return fail(kRuntimeError, "timing a tile");
    for (int k = 0; k < reps && !r; ++k) r = once();
    if (r) return r;
    float t = 0.f;
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&t, e0, e1) != hipSuccess)
      return fail(kRuntimeError, "timing a tile");
    out = (double)t / reps;
    return 0;
  };
  const int divs[3] = {1, 2, 4};
  *best_col = 1;
  *best_row = 1;
  for (int k = 0; k < 6; ++k) ms[k] = -1.0;
  double best_ms = -1.0;
  b.row_div = 1;
  const int ndiv = g.Lw == g.L ? 3 : 1;  // fold mode runs the full tiles only
  for (int k = 0; k < ndiv && !st; ++k) {
    if (!linpatch_col_tile(dtype, g.L, divs[k])) continue;
    b.col_div = divs[k];
    st = time_tiles(ms[k]);
    if (!st && (best_ms < 0.0 || ms[k] < best_ms)) {
      best_ms = ms[k];
      *best_col = divs[k];
    }
  }
  b.col_div = *best_col;
  best_ms = -1.0;
  for (int k = 0; k < ndiv && !st; ++k) {
    if (!linpatch_row_tile(dtype, g.L, divs[k])) continue;
    b.row_div = divs[k];
    st = time_tiles(ms[3 + k]);
    if (!st && (best_ms < 0.0 || ms[3 + k] < best_ms)) {
      best_ms = ms[3 + k];
      *best_row = divs[k];
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (x || y) (void)hipStreamSynchronize(s);
  dev_free(x);
  dev_free(y);
  return st;
}

// 0 if (col_div, row_div) is compiled in for the plan's length; col_div = 0
// or row_div = 0 is accepted where `follow` allows it (the plan's value).
static int check_tiles(const jubik_linpatch* plan, int col_div, int row_div, bool follow) {
  const int L = plan->g.L;
  if (plan->g.Lw != L && (col_div > 1 || row_div > 1))
    return fail(kInvalidArgument, "window length " + std::to_string(plan->g.Lw) +
                                      " runs in fold mode at L = " + std::to_string(L) +
                                      ", which has the full tiles only");
  if (!(follow && col_div == 0) && !linpatch_col_tile(plan->dtype, L, col_div))
    return fail(kInvalidArgument, "column tile divisor " + std::to_string(col_div) +
                                      " is not compiled in for L = " + std::to_string(L));
  if (!(follow && row_div == 0) && !linpatch_row_tile(plan->dtype, L, row_div))
    return fail(kInvalidArgument, "row tile divisor " + std::to_string(row_div) +
                                      " is not compiled in for L = " + std::to_string(L));
  return kOk;
}

int jubik_linpatch_set_tile(jubik_linpatch* plan, int col_div, int row_div) {
  if (!plan) return fail(kInvalidArgument, "NULL argument");
  const int st = check_tiles(plan, col_div, row_div, false);
  if (st) return st;
  plan->col_div = col_div;
  plan->row_div = row_div;
  return kOk;
}

int jubik_linpatch_autotune(jubik_linpatch* plan, int reps, int* col_div, int* row_div,
                            double* ms, void* stream) {
  if (!plan || !col_div || !row_div || !ms) return fail(kInvalidArgument, "NULL argument");
  if (reps < 1) return fail(kInvalidArgument, "reps must be positive");
  if (!plan->ready) return fail(kNotReady, "no kernel set (call jubik_linpatch_set_kernel)");
  DeviceGuard dg(plan->device);
  int bc = 1, br = 1;
  const int st = autotune_tiles(plan->dtype, plan->g, plan->bufs(), 1, reps, &bc, &br, ms,
                                as_stream(stream));
  if (st) return st;
  plan->col_div = bc;
  plan->row_div = br;
  *col_div = bc;
  *row_div = br;
  return kOk;
}

int jubik_linpatch_set_transform(jubik_linpatch* plan, int kind) {
  if (!plan) return fail(kInvalidArgument, "NULL argument");
  if (kind != 0 && kind != 1) return fail(kInvalidArgument, "transform must be 0 or 1");
  if (kind == 1) {
    if (!linpatch_mfma_supported(plan->dtype, plan->g))
      return fail(kUnsupported, "the split-BF16 matrix-core transform needs a float32 plan with "
                                "FFT length 384 (window + 2 * margin = 384); this plan has " +
                                    std::string(plan->dtype ? "float64" : "float32") +
                                    " and length " + std::to_string(plan->g.Lw));
    if (!plan->frag) {
      DeviceGuard dg(plan->device);
      std::vector<unsigned short> host;
      dft128_fragments(host);
      const size_t bytes = host.size() * sizeof(unsigned short);
      const int st = dev_alloc(&plan->frag, bytes, "matrix-core DFT fragments");
      if (st) return st;
      JUBIK_HIP(hipMemcpy(plan->frag, host.data(), bytes, hipMemcpyHostToDevice));
    }
  }
  plan->transform = kind;
  return kOk;
}

// ------------------------------------------------------------------ primitives
int jubik_slice_patches(int dtype, const void* x, int64_t B, int64_t H, int64_t W, int64_t n,
                        int64_t margin, void* out, void* stream) {
  if (dtype != 0 && dtype != 1) return fail(kInvalidArgument, "dtype must be 0 or 1");
  if (!x || !out) return fail(kInvalidArgument, "NULL argument");
  return slice_patches(dtype, x, B, H, W, n, margin, out, as_stream(stream));
}

int jubik_slice_patches_adjoint(int dtype, const void* patches, int64_t B, int64_t H, int64_t W,
                                int64_t n, int64_t margin, void* x, void* stream) {
  if (dtype != 0 && dtype != 1) return fail(kInvalidArgument, "dtype must be 0 or 1");
  if (!patches || !x) return fail(kInvalidArgument, "NULL argument");
  return slice_patches_adjoint(dtype, patches, B, H, W, n, margin, x, as_stream(stream));
}

int64_t jubik_convolve_work_bytes(int dtype, int64_t nby, int64_t H, int64_t W) {
  const int64_t LH = convolve_fft_length(H), LW = convolve_fft_length(W);
  if (nby < 1 || W < 1 || !LH || !LW) return -1;
  return (nby * LH * W + LH + LW) * 2 * (int64_t)elem_size(dtype);
}

int jubik_convolve(int dtype, const void* x, const void* y, int64_t nbx, int64_t nby, int64_t H,
                   int64_t W, double dvol, void* work, void* out, void* stream) {
  if (dtype != 0 && dtype != 1) return fail(kInvalidArgument, "dtype must be 0 or 1");
  if (!x || !y || !work || !out) return fail(kInvalidArgument, "NULL argument");
  if (H < 1 || W < 1) return fail(kInvalidArgument, "convolve axes must not be empty");
  const int64_t LH = convolve_fft_length(H), LW = convolve_fft_length(W);
  if (!LH || !LW)
    return fail(kUnsupported, "convolve axis of length " + std::to_string(!LH ? H : W) +
                                  " exceeds 2048");
  // twiddles of both FFT lengths go behind the complex work planes
  const size_t es = elem_size(dtype);
  static thread_local std::vector<unsigned char> host;
  host.resize(2 * es * (LH + LW));
  int64_t at = 0;
  for (int64_t L : {LH, LW})
    for (int64_t e = 0; e < L; ++e, ++at) {
      const long double ang = -2.0L * 3.141592653589793238462643383279502884L * e / L;
      const double c = (double)cosl(ang), sn = (double)sinl(ang);
      if (dtype == 0) {
        reinterpret_cast<float*>(host.data())[2 * at] = (float)c;
        reinterpret_cast<float*>(host.data())[2 * at + 1] = (float)sn;
      } else {
        reinterpret_cast<double*>(host.data())[2 * at] = c;
        reinterpret_cast<double*>(host.data())[2 * at + 1] = sn;
      }
    }
  char* twh = reinterpret_cast<char*>(work) + (size_t)nby * LH * W * 2 * es;
  char* tww = twh + (size_t)LH * 2 * es;
  JUBIK_HIP(hipMemcpyAsync(twh, host.data(), host.size(), hipMemcpyHostToDevice,
                           as_stream(stream)));
  return convolve2d(dtype, x, y, nbx, nby, H, W, dvol, twh, tww, work, out, as_stream(stream));
}

// ------------------------------------------------------------------ response
int jubik_exposure_apply(int dtype, const void* exposure, int64_t n_e, const void* x,
                         int64_t n_x, void* out, int64_t n_out, void* stream) {
  if (dtype != 0 && dtype != 1) return fail(kInvalidArgument, "dtype must be 0 or 1");
  if (n_out < 0 || n_x <= 0 || n_e <= 0 || (n_out % n_x) != 0 || (n_out % n_e) != 0)
    return fail(kInvalidArgument, "operand sizes must divide the output size");
  if (n_out && (!exposure || !x || !out)) return fail(kInvalidArgument, "NULL argument");
  return exposure_apply(dtype, exposure, n_e, x, n_x, out, n_out, as_stream(stream));
}

int jubik_readout_gather(int dtype, const void* x, const int64_t* index, int64_t nnz, void* out,
                         void* stream) {
  if (dtype != 0 && dtype != 1) return fail(kInvalidArgument, "dtype must be 0 or 1");
  if (nnz < 0) return fail(kInvalidArgument, "negative size");
  if (nnz && (!x || !index || !out)) return fail(kInvalidArgument, "NULL argument");
  return readout_gather(dtype, x, index, nnz, out, as_stream(stream));
}

int jubik_readout_scatter(int dtype, const void* v, const int64_t* index, int64_t nnz, int64_t n,
                          void* out, void* stream) {
  if (dtype != 0 && dtype != 1) return fail(kInvalidArgument, "dtype must be 0 or 1");
  if (nnz < 0 || n < 0 || nnz > n) return fail(kInvalidArgument, "invalid sizes");
  if (n && !out) return fail(kInvalidArgument, "NULL argument");
  if (nnz && (!v || !index)) return fail(kInvalidArgument, "NULL argument");
  return readout_scatter(dtype, v, index, nnz, n, out, as_stream(stream));
}

int jubik_poisson_sample(int dtype, const void* lam, int64_t n, uint64_t seed, int32_t* counts,
                         void* stream) {
  if (dtype != 0 && dtype != 1) return fail(kInvalidArgument, "dtype must be 0 or 1");
  if (n < 0) return fail(kInvalidArgument, "negative size");
  if (n && (!lam || !counts)) return fail(kInvalidArgument, "NULL argument");
  return poisson_sample(dtype, lam, n, seed, counts, as_stream(stream));
}

int jubik_poisson_nll_grad(int dtype, const void* lam, const void* data, int64_t n, double* nll,
                           void* grad, void* work, void* stream) {
  if (dtype != 0 && dtype != 1) return fail(kInvalidArgument, "dtype must be 0 or 1");
  if (n < 0) return fail(kInvalidArgument, "negative size");
  if (!nll || !work || (n && (!lam || !data))) return fail(kInvalidArgument, "NULL argument");
  return poisson_nll_grad(dtype, lam, data, n, nll, grad, (double*)work, as_stream(stream));
}

int jubik_poisson_nll_grad_noalloc(int dtype, const void* lam, const void* data, int64_t n,
                                   double* nll, void* grad, void* stream) {
  if (dtype != 0 && dtype != 1) return fail(kInvalidArgument, "dtype must be 0 or 1");
  if (n < 0) return fail(kInvalidArgument, "negative size");
  if (!nll || (n && (!lam || !data))) return fail(kInvalidArgument, "NULL argument");
  return poisson_nll_grad_noalloc(dtype, lam, data, n, nll, grad, as_stream(stream));
}

int jubik_poisson_metric(int dtype, const void* lam, const void* t, int64_t n, int sqrt_form,
                         void* out, void* stream) {
  if (dtype != 0 && dtype != 1) return fail(kInvalidArgument, "dtype must be 0 or 1");
  if (sqrt_form != 0 && sqrt_form != 1) return fail(kInvalidArgument, "sqrt_form must be 0 or 1");
  if (n < 0) return fail(kInvalidArgument, "negative size");
  if (n && (!lam || !t || !out)) return fail(kInvalidArgument, "NULL argument");
  return poisson_metric(dtype, lam, t, n, sqrt_form, out, as_stream(stream));
}

// ------------------------------------------------------------------ DFT comparison
struct jubik_dft128 {
  void* frag = nullptr;
  void* tw = nullptr;
  int device = 0;
  int nblocks = 0;
};

int jubik_dft128_create(jubik_dft128** plan) {
  if (!plan) return fail(kInvalidArgument, "NULL argument");
  *plan = nullptr;
  auto* p = new jubik_dft128();
  int st = hipGetDevice(&p->device) == hipSuccess ? kOk : fail(kRuntimeError, "no HIP device");
  hipDeviceProp_t prop;
  if (!st) st = check_hip(hipGetDeviceProperties(&prop, p->device), "hipGetDeviceProperties");
  if (!st) p->nblocks = prop.multiProcessorCount;
  std::vector<unsigned short> host;
  dft128_fragments(host);
  if (!st) st = dev_alloc(&p->frag, host.size() * 2, "DFT fragments");
  if (!st) st = check_hip(hipMemcpy(p->frag, host.data(), host.size() * 2, hipMemcpyHostToDevice),
                          "DFT fragments");
  if (!st) st = upload_twiddles(&p->tw, 0, 128);
  if (st) {
    dev_free(p->frag);
    dev_free(p->tw);
    delete p;
    return st;
  }
  *plan = p;
  return kOk;
}

int jubik_dft128_run(jubik_dft128* plan, int method, const void* x, void* y, int64_t B,
                     void* stream) {
  if (!plan || !x || !y) return fail(kInvalidArgument, "NULL argument");
  if (method != 0 && method != 1) return fail(kInvalidArgument, "method must be 0 or 1");
  if (B <= 0 || B % kDft128TileVectors)
    return fail(kInvalidArgument, "batch must be a positive multiple of 256 vectors");
  DeviceGuard dg(plan->device);
  return dft128_run(method, plan->frag, plan->tw, x, y, B, plan->nblocks, as_stream(stream));
}

int jubik_dft128_destroy(jubik_dft128* plan) {
  if (!plan) return kOk;
  DeviceGuard dg(plan->device);
  dev_free(plan->frag);
  dev_free(plan->tw);
  delete plan;
  return kOk;
}

// ------------------------------------------------------------------ PSF
int jubik_psf_interpolate(int dtype, const void* psfs, int64_t nrad, int64_t ph, int64_t pw,
                          const double* rs, const double* patch_center_ids,
                          const double* patch_deltas, const double* pointing_center, int64_t H,
                          int64_t W, double dH, double dW, int64_t npatch, int64_t e, int64_t E,
                          int layout, void* out, void* stream) {
  if (dtype != 0 && dtype != 1) return fail(kInvalidArgument, "dtype must be 0 or 1");
  if (!psfs || !rs || !patch_center_ids || !patch_deltas || !pointing_center || !out)
    return fail(kInvalidArgument, "NULL argument");
  if (nrad < 1 || ph < 1 || pw < 1 || npatch < 1 || H < 2 || W < 2 || E < 1 || e < 0 || e >= E)
    return fail(kInvalidArgument, "invalid sizes");
  if (layout != 0 && layout != 1) return fail(kInvalidArgument, "layout must be 0 or 1");
  // erosita_psf_utils.py:196-201 and :143-147
  if (H % npatch || W % npatch) return fail(kInvalidArgument, "grid not divisible by npatch");
  if (H % 2 || W % 2) return fail(kInvalidArgument, "grid sides must be even");
  for (int64_t k = 0; k < nrad; ++k) {
    const double c0 = patch_center_ids[2 * k], c1 = patch_center_ids[2 * k + 1];
    if (c0 > (double)ph || c1 > (double)pw || c0 < 0 || c1 < 0)
      return fail(kInvalidArgument, "PSF centre index outside its grid");
  }
  if (!(patch_deltas[0] > 0) || !(patch_deltas[1] > 0))
    return fail(kInvalidArgument, "PSF pixel sizes must be positive");
  PsfGeom g{};
  g.H = (int)H;
  g.W = (int)W;
  g.dH = dH;
  g.dW = dW;
  g.npatch = (int)npatch;
  g.nrad = (int)nrad;
  g.ph = (int)ph;
  g.pw = (int)pw;
  g.d0 = patch_deltas[0];
  g.d1 = patch_deltas[1];
  g.layout = layout;
  g.w = (int)(2 * (H / npatch));
  g.e = (int)e;
  g.E = (int)E;
  if (layout == 1 && (g.w > H || g.w > W))
    return fail(kInvalidArgument, "window wider than the grid (npatch must be >= 2)");
  std::vector<int> order;
  std::vector<double> phi, wgt;
  psf_patch_table(rs, (int)nrad, pointing_center, g.H, g.W, dH, dW, g.npatch, order, phi, wgt);
  std::vector<double> cid(2 * nrad);
  for (int64_t k = 0; k < nrad; ++k) {
    cid[2 * k] = patch_center_ids[2 * order[k]];
    cid[2 * k + 1] = patch_center_ids[2 * order[k] + 1];
  }
  // one set-up table on the device: order | cid | phi | wgt
  const size_t b_order = (size_t)nrad * sizeof(int);
  const size_t off_cid = (b_order + 15) / 16 * 16;
  const size_t off_phi = off_cid + cid.size() * sizeof(double);
  const size_t off_wgt = off_phi + phi.size() * sizeof(double);
  const size_t bytes = off_wgt + wgt.size() * sizeof(double);
  std::vector<unsigned char> host(bytes, 0);
  std::memcpy(host.data(), order.data(), b_order);
  std::memcpy(host.data() + off_cid, cid.data(), cid.size() * sizeof(double));
  std::memcpy(host.data() + off_phi, phi.data(), phi.size() * sizeof(double));
  std::memcpy(host.data() + off_wgt, wgt.data(), wgt.size() * sizeof(double));
  void* table = nullptr;
  int st = dev_alloc(&table, bytes, "PSF patch table");
  if (st) return st;
  hipStream_t s = as_stream(stream);
  hipError_t he = hipMemcpyAsync(table, host.data(), bytes, hipMemcpyHostToDevice, s);
  if (he == hipSuccess) {
    char* t = reinterpret_cast<char*>(table);
    st = psf_interpolate(dtype, g, psfs, reinterpret_cast<const int*>(t),
                         reinterpret_cast<const double*>(t + off_cid),
                         reinterpret_cast<const double*>(t + off_phi),
                         reinterpret_cast<const double*>(t + off_wgt), out, s);
    he = hipStreamSynchronize(s);
  }
  if (!st && he != hipSuccess) st = check_hip(he, "psf_interpolate");
  dev_free(table);
  return st;
}

// ------------------------------------------------------------------ GP
int jubik_gp_create(jubik_gp** plan, int dtype, int64_t N, int64_t H, int64_t W, int64_t E,
                    int has_alpha) {
  if (!plan) return fail(kInvalidArgument, "plan pointer is NULL");
  *plan = nullptr;
  if (dtype != 0 && dtype != 1) return fail(kInvalidArgument, "dtype must be 0 or 1");
  if (E < 1 || H < 1 || W < 1 || H > N || W > N)
    return fail(kInvalidArgument, "sky grid must fit in the GP grid");
  if (!gp_length_supported((int)N))
    return fail(kUnsupported, "GP grid size " + std::to_string(N) + " is not compiled in");
  if (!has_alpha && E != 1)
    return fail(kInvalidArgument, "several energies need the spectral-index field");
  auto* p = new jubik_gp();
  p->dtype = dtype;
  p->g = GpGeom{(int)N, (int)H, (int)W, (int)E, has_alpha ? 1 : 0, 0.0, 0.0, 1, 1, 1};
  if (hipGetDevice(&p->device) != hipSuccess) {
    delete p;
    return fail(kRuntimeError, "no HIP device available");
  }
  const size_t cs = 2 * elem_size(dtype);
  int st = upload_twiddles(&p->tw, dtype, (int)N);
  if (!st) st = dev_alloc(&p->work, (size_t)N * N * cs, "GP workspace");
  if (!st) st = dev_alloc(&p->u, (size_t)E * H * W * elem_size(dtype), "GP cotangent planes");
  if (st) {
    p->release();
    delete p;
    return st;
  }
  *plan = p;
  return kOk;
}

int jubik_gp_destroy(jubik_gp* plan) {
  if (!plan) return kOk;
  DeviceGuard dg(plan->device);
  plan->release();
  delete plan;
  return kOk;
}

int jubik_gp_forward(jubik_gp* plan, const void* amp_ref, const void* amp_alpha,
                     const void* xi_ref, const void* xi_alpha, double zm, double mu,
                     const void* dnu, void* sky, void* stream) {
  if (!plan || !amp_ref || !xi_ref || !sky) return fail(kInvalidArgument, "NULL argument");
  if (plan->g.has_alpha && (!amp_alpha || !xi_alpha || !dnu))
    return fail(kInvalidArgument, "spectral-index inputs are NULL");
  DeviceGuard dg(plan->device);
  GpGeom g = plan->g;
  g.zm = zm;
  g.mu = mu;
  return gp_forward(plan->dtype, g, plan->bufs(), amp_ref, amp_alpha, xi_ref, xi_alpha, dnu, sky,
                    as_stream(stream));
}

int jubik_gp_vjp(jubik_gp* plan, const void* amp_ref, const void* amp_alpha, const void* dnu,
                 const void* sky, const void* cot_sky, void* g_ref, void* g_alpha, void* stream) {
  if (!plan || !amp_ref || !sky || !cot_sky || !g_ref)
    return fail(kInvalidArgument, "NULL argument");
  if (plan->g.has_alpha && (!amp_alpha || !dnu || !g_alpha))
    return fail(kInvalidArgument, "spectral-index inputs are NULL");
  DeviceGuard dg(plan->device);
  int st = gp_fold(plan->dtype, plan->g, sky, cot_sky, plan->u, as_stream(stream));
  if (st) return st;
  GpGeom gg = plan->g;
  gg.uplanes = gg.E;
  return gp_adjoint_u(plan->dtype, gg, plan->bufs(), amp_ref, amp_alpha, dnu, plan->u, g_ref,
                      g_alpha, as_stream(stream));
}

// ------------------------------------------------------------------ model
int jubik_model_create_batched(jubik_model** model, jubik_gp* gp, jubik_linpatch* lp,
                               double pixel_area, int64_t max_samples) {
  if (!model || !gp || !lp) return fail(kInvalidArgument, "NULL argument");
  *model = nullptr;
  if (max_samples < 1 || max_samples > 4096)
    return fail(kInvalidArgument, "max_samples must be in [1, 4096]");
  if (gp->dtype != lp->dtype) return fail(kInvalidArgument, "gp and linpatch dtypes differ");
  if (gp->device != lp->device) return fail(kInvalidArgument, "gp and linpatch devices differ");
  const LinGeom& g = lp->g;
  if (gp->g.H != g.H || gp->g.W != g.W || gp->g.E != g.E)
    return fail(kInvalidArgument, "gp sky grid and linpatch grid differ");
  if (!g.x_shared) return fail(kInvalidArgument, "the model needs a linpatch plan with x_shared");
  auto* m = new jubik_model();
  m->gp = gp;
  m->lp = lp;
  m->pixel_area = pixel_area;
  m->max_samples = (int)max_samples;
  DeviceGuard dg(lp->device);
  const size_t S = (size_t)max_samples;
  const size_t es = elem_size(lp->dtype);
  const size_t img = (size_t)g.H * g.W;
  int st = dev_alloc(&m->sky, img * g.E * (S + 1) * es, "sky");
  if (!st) st = dev_alloc(&m->lam, img * g.J * (S + 1) * es, "lambda");
  if (!st) st = dev_alloc(&m->dsky, img * g.E * S * es, "sky tangent");
  if (!st) st = dev_alloc(&m->resid, img * g.J * S * es, "residual");
  if (!st) st = dev_alloc(&m->partial, 2 * (size_t)g.J * g.H * S * sizeof(double), "NLL partials");
  if (!st) st = dev_alloc(&m->nbad, S * sizeof(double), "lambda domain counts");
  // GP cotangent: E real planes per sample, or one complex plane (energies
  // folded in) per sample
  if (!st) st = dev_alloc(&m->u, img * std::max(g.E, 2) * S * es, "GP cotangent planes");
  if (!st && S > 1) {
    st = dev_alloc(&m->R1, lp->ws_bytes * S, "batched workspace R1");
    if (!st) st = dev_alloc(&m->R2, lp->ws_bytes * S, "batched workspace R2");
    if (!st) st = dev_alloc(&m->gwork, (size_t)gp->g.N * gp->g.N * 2 * es * S, "batched GP workspace");
  }
  if (st) {
    m->release();
    delete m;
    return st;
  }
  *model = m;
  return kOk;
}

int jubik_model_create(jubik_model** model, jubik_gp* gp, jubik_linpatch* lp, double pixel_area) {
  return jubik_model_create_batched(model, gp, lp, pixel_area, 1);
}

int jubik_model_destroy(jubik_model* model) {
  if (!model) return kOk;
  DeviceGuard dg(model->lp->device);
  model->release();
  delete model;
  return kOk;
}

int jubik_model_set_data(jubik_model* model, const void* exposure, const int32_t* counts) {
  if (!model || !exposure || !counts) return fail(kInvalidArgument, "NULL argument");
  model->expo = exposure;
  model->counts = counts;
  model->have_point = false;
  return kOk;
}

namespace {

// S gradient evaluations at the positions xi[s]; sky and lam go to slots
// slot0 .. slot0 + S - 1, nll[s] is the NLL of position s.
// With point sources (xi_ps set; the model's point-source prior) the PSF
// sees the diffuse sky plus p_e = I(xi_ps) exp(alpha dnu_e), and g_ps,
// g_ps_alpha receive the gradients with respect to xi_ps and xi_ps_alpha.
struct PsArgs {
  const void* xi = nullptr;
  const void* xa = nullptr;
  void* g_xi = nullptr;
  void* g_xa = nullptr;
};

int model_grad(jubik_model* model, int S, int slot0, const void* amp_ref, const void* amp_alpha,
               const void* xi_ref, const void* xi_alpha, double zm, double mu, const void* dnu,
               double* nll, void* g_ref, void* g_alpha, hipStream_t s, GpHyper hy = GpHyper{},
               PsArgs pa = PsArgs{}) {
  jubik_linpatch* lp = model->lp;
  jubik_gp* gp = model->gp;
  model->geo_ready = false;  // slots 0 / 1 change
  GpGeom gg = gp->g;
  gg.zm = zm;
  gg.mu = mu;
  gg.S = S;
  gg.sky_shared = 0;
  LinGeom g = lp->g;
  g.S = S;
  g.sky_shared = 0;
  void* sky = model->sky_slot(slot0);
  int st = gp_forward(gp->dtype, gg, model->gbufs(), amp_ref, amp_alpha, xi_ref, xi_alpha, dnu,
                      sky, s);
  if (st) return st;
  const bool ps = pa.xi != nullptr;
  const int64_t npix = (int64_t)g.H * g.W;
  const void* psf_in = sky;
  if (ps) {
    st = ps_prepare(lp->dtype, model->ps, S, npix, pa.xi, pa.xa, model->psw, s);
    if (!st)
      st = ps_add(lp->dtype, S, g.E, npix, sky, model->psw, gp->g.has_alpha ? dnu : nullptr,
                  model->skyt, s);
    if (st) return st;
    psf_in = model->skyt;
  }
  st = linpatch_forward_poisson(lp->dtype, g, model->lbufs(), psf_in, model->pixel_area,
                                model->expo, model->counts, model->resid, model->lam_slot(slot0),
                                (double*)model->partial, s);
  if (st) return st;
  NllFinal fin;
  fin.partial = (const double*)model->partial;
  fin.seglen = (int64_t)g.J * g.H;
  fin.nseg = S;
  fin.out = nll;
  fin.bad_out = (double*)model->nbad;
  model->last_S = S;
  if (slot0 == 1) model->batch_S = S;
  st = linpatch_adjoint_gp(lp->dtype, g, model->lbufs(), model->resid, model->pixel_area, sky,
                           gp->g.has_alpha ? dnu : nullptr, model->u, s, fin,
                           ps ? model->ct : nullptr);
  if (st) return st;
  // without point sources the adjoint folded the energies into one complex
  // plane per sample
  gg.uplanes = ps ? g.E : 1;
  gg.ucomplex = ps ? 0 : 1;
  st = gp_adjoint_u(gp->dtype, gg, model->gbufs(), amp_ref, amp_alpha, dnu, model->u, g_ref,
                    g_alpha, s, hy);
  if (st || !ps) return st;
  return ps_grad(lp->dtype, model->ps, S, g.E, npix, model->ct, model->psw,
                 gp->g.has_alpha ? dnu : nullptr, pa.g_xi, gp->g.has_alpha ? pa.g_xa : nullptr,
                 s);
}

int check_model(jubik_model* model) {
  if (!model->expo) return fail(kNotReady, "no data set (call jubik_model_set_data)");
  if (!model->lp->ready) return fail(kNotReady, "no kernel set (call jubik_linpatch_set_kernel)");
  return kOk;
}

int check_alpha(const jubik_model* model, const void* amp_alpha, const void* a, const void* b,
                const void* dnu) {
  if (model->gp->g.has_alpha && (!amp_alpha || !a || !b || !dnu))
    return fail(kInvalidArgument, "spectral-index inputs are NULL");
  return kOk;
}

}  // namespace

int jubik_model_set_point_sources(jubik_model* model, int enable, double a, double q,
                                  double alpha_mean, double alpha_std) {
  if (!model) return fail(kInvalidArgument, "NULL argument");
  if (!enable) {
    model->ps_on = false;
    return kOk;
  }
  if (!(a > 0.0) || !(q > 0.0))
    return fail(kInvalidArgument, "the inverse-gamma prior needs a > 0 and q > 0");
  if (!(alpha_std > 0.0))
    return fail(kInvalidArgument, "the spectral-index prior needs std > 0");
  DeviceGuard dg(model->lp->device);
  const size_t plane = model->img_bytes();
  const size_t S = (size_t)model->max_samples;
  int st = kOk;
  if (!model->skyt) st = dev_alloc(&model->skyt, S * model->lp->g.E * plane, "point-source sky");
  if (!st && !model->ct) st = dev_alloc(&model->ct, S * model->lp->g.E * plane, "sky cotangent");
  if (!st && !model->psw) st = dev_alloc(&model->psw, S * 3 * plane, "point-source parameters");
  if (st) {
    dev_free(model->skyt);
    dev_free(model->ct);
    dev_free(model->psw);
    return st;
  }
  model->ps = PointSources{a, q, alpha_mean, alpha_std};
  model->ps_on = true;
  model->have_point = false;
  return kOk;
}

int jubik_model_grad_points(jubik_model* model, int64_t S, const void* amp_ref,
                            const void* amp_alpha, const void* xi_ref, const void* xi_alpha,
                            double zm, double mu, const void* dnu, const void* xi_ps,
                            const void* xi_ps_alpha, double* nll, void* g_ref, void* g_alpha,
                            void* g_ps, void* g_ps_alpha, void* stream) {
  if (!model || !nll || !g_ref || !amp_ref || !xi_ref || !xi_ps || !g_ps)
    return fail(kInvalidArgument, "NULL argument");
  if (!model->ps_on)
    return fail(kNotReady, "no point-source prior set (call jubik_model_set_point_sources)");
  int st = check_model(model);
  if (!st) st = check_alpha(model, amp_alpha, xi_alpha, g_alpha, dnu);
  if (!st && model->gp->g.has_alpha && (!xi_ps_alpha || !g_ps_alpha))
    st = fail(kInvalidArgument, "spectral-index inputs are NULL");
  if (st) return st;
  if (S < 1 || S > model->max_samples)
    return fail(kInvalidArgument, "samples " + std::to_string(S) + " outside [1, " +
                                      std::to_string(model->max_samples) + "]");
  DeviceGuard dg(model->lp->device);
  PsArgs pa;
  pa.xi = xi_ps;
  pa.xa = xi_ps_alpha;
  pa.g_xi = g_ps;
  pa.g_xa = g_ps_alpha;
  // one position: the metric base point (slot 0); a batch: slots 1 .. S
  const int slot0 = S == 1 ? 0 : 1;
  if (S == 1) model->have_point = false;
  st = model_grad(model, (int)S, slot0, amp_ref, amp_alpha, xi_ref, xi_alpha, zm, mu, dnu, nll,
                  g_ref, g_alpha, as_stream(stream), GpHyper{}, pa);
  if (st == kOk && S == 1) model->have_point = true;
  return st;
}

int jubik_invgamma_prior(int dtype, const void* xi, int64_t n, double a, double q, void* log_i,
                         void* dlog_i, void* stream) {
  if (dtype != 0 && dtype != 1) return fail(kInvalidArgument, "dtype must be 0 or 1");
  if (n < 0 || (n > 0 && (!xi || !log_i))) return fail(kInvalidArgument, "NULL argument");
  if (!(a > 0.0) || !(q > 0.0))
    return fail(kInvalidArgument, "the inverse-gamma prior needs a > 0 and q > 0");
  return invgamma_prior(dtype, xi, n, a, q, log_i, dlog_i, as_stream(stream));
}

int jubik_model_grad_hyper(jubik_model* model, const void* amp_ref, const void* amp_alpha,
                           const void* xi_ref, const void* xi_alpha, double zm, double mu,
                           const void* dnu, double* nll, void* g_ref, void* g_alpha,
                           void* g_amp_ref, void* g_amp_alpha, void* g_offsets, void* stream) {
  if (!model || !nll || !g_ref || !amp_ref || !xi_ref || !g_amp_ref || !g_offsets)
    return fail(kInvalidArgument, "NULL argument");
  int st = check_model(model);
  if (!st) st = check_alpha(model, amp_alpha, xi_alpha, g_alpha, dnu);
  if (!st && model->gp->g.has_alpha && !g_amp_alpha)
    st = fail(kInvalidArgument, "spectral-index inputs are NULL");
  if (st) return st;
  DeviceGuard dg(model->lp->device);
  model->have_point = false;
  GpHyper hy;
  hy.xi_ref = xi_ref;
  hy.xi_alpha = xi_alpha;
  hy.g_amp_ref = g_amp_ref;
  hy.g_amp_alpha = g_amp_alpha;
  hy.g_off = g_offsets;
  st = model_grad(model, 1, 0, amp_ref, amp_alpha, xi_ref, xi_alpha, zm, mu, dnu, nll, g_ref,
                  g_alpha, as_stream(stream), hy);
  if (st == kOk) model->have_point = true;
  return st;
}

int jubik_model_grad(jubik_model* model, const void* amp_ref, const void* amp_alpha,
                     const void* xi_ref, const void* xi_alpha, double zm, double mu,
                     const void* dnu, double* nll, void* g_ref, void* g_alpha, void* stream) {
  if (!model || !nll || !g_ref || !amp_ref || !xi_ref)
    return fail(kInvalidArgument, "NULL argument");
  int st = check_model(model);
  if (!st) st = check_alpha(model, amp_alpha, xi_alpha, g_alpha, dnu);
  if (st) return st;
  DeviceGuard dg(model->lp->device);
  model->have_point = false;
  st = model_grad(model, 1, 0, amp_ref, amp_alpha, xi_ref, xi_alpha, zm, mu, dnu, nll, g_ref,
                  g_alpha, as_stream(stream));
  if (st == kOk) model->have_point = true;
  return st;
}

int jubik_model_grad_batch(jubik_model* model, int64_t S, const void* amp_ref,
                           const void* amp_alpha, const void* xi_ref, const void* xi_alpha,
                           double zm, double mu, const void* dnu, double* nll, void* g_ref,
                           void* g_alpha, void* stream) {
  if (!model || !nll || !g_ref || !amp_ref || !xi_ref)
    return fail(kInvalidArgument, "NULL argument");
  if (S < 1 || S > model->max_samples)
    return fail(kInvalidArgument, "batch of " + std::to_string(S) +
                                      " samples outside [1, max_samples = " +
                                      std::to_string(model->max_samples) + "]");
  int st = check_model(model);
  if (!st) st = check_alpha(model, amp_alpha, xi_alpha, g_alpha, dnu);
  if (st) return st;
  DeviceGuard dg(model->lp->device);
  return model_grad(model, (int)S, 1, amp_ref, amp_alpha, xi_ref, xi_alpha, zm, mu, dnu, nll,
                    g_ref, g_alpha, as_stream(stream));
}

int jubik_model_set_tile(jubik_model* model, int col_div, int row_div) {
  if (!model || !model->lp) return fail(kInvalidArgument, "NULL argument");
  const int st = check_tiles(model->lp, col_div, row_div, true);
  if (st) return st;
  model->col_div = col_div;
  model->row_div = row_div;
  return kOk;
}

int jubik_model_autotune(jubik_model* model, int64_t S, int reps, int* col_div, int* row_div,
                         double* ms, void* stream) {
  if (!model || !col_div || !row_div || !ms) return fail(kInvalidArgument, "NULL argument");
  if (reps < 1) return fail(kInvalidArgument, "reps must be positive");
  if (S < 1 || S > model->max_samples)
    return fail(kInvalidArgument, "batch of " + std::to_string(S) +
                                      " samples outside [1, max_samples = " +
                                      std::to_string(model->max_samples) + "]");
  int st = check_model(model);
  if (st) return st;
  DeviceGuard dg(model->lp->device);
  int bc = 1, br = 1;
  st = autotune_tiles(model->lp->dtype, model->lp->g, model->lbufs(), (int)S, reps, &bc, &br, ms,
                      as_stream(stream));
  if (st) return st;
  model->col_div = bc;
  model->row_div = br;
  *col_div = bc;
  *row_div = br;
  return kOk;
}

int jubik_model_grad_graph(jubik_graph** out, jubik_model* model, int64_t S,
                           const void* amp_ref, const void* amp_alpha, const void* xi_ref,
                           const void* xi_alpha, double zm, double mu, const void* dnu,
                           double* nll, void* g_ref, void* g_alpha) {
  if (!out || !model || !nll || !g_ref || !amp_ref || !xi_ref)
    return fail(kInvalidArgument, "NULL argument");
  *out = nullptr;
  if (S < 0 || S > model->max_samples)
    return fail(kInvalidArgument, "batch of " + std::to_string(S) +
                                      " samples outside [0, max_samples = " +
                                      std::to_string(model->max_samples) + "]");
  if (timing_enabled())
    return fail(kInvalidArgument, "disable kernel timing before capturing a graph");
  int st = check_model(model);
  if (!st) st = check_alpha(model, amp_alpha, xi_alpha, g_alpha, dnu);
  if (st) return st;
  DeviceGuard dg(model->lp->device);
  hipStream_t cs = nullptr;
  JUBIK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  auto* gr = new jubik_graph();
  if (S == 0) gr->points.push_back(model);
  gr->device = model->lp->device;
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    st = S == 0 ? model_grad(model, 1, 0, amp_ref, amp_alpha, xi_ref, xi_alpha, zm, mu, dnu,
                                 nll, g_ref, g_alpha, cs)
                    : model_grad(model, (int)S, 1, amp_ref, amp_alpha, xi_ref, xi_alpha, zm, mu,
                                 dnu, nll, g_ref, g_alpha, cs);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(cs, &g);
    gr->graph = g;
    if (!st && e2 != hipSuccess) st = check_hip(e2, "hipStreamEndCapture");
    if (!st) st = check_hip(hipGraphInstantiate(&gr->exec, gr->graph, nullptr, nullptr, 0),
                            "hipGraphInstantiate");
  } else {
    st = check_hip(e, "hipStreamBeginCapture");
  }
  (void)hipStreamDestroy(cs);
  if (st) {
    if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
    if (gr->graph) (void)hipGraphDestroy(gr->graph);
    delete gr;
    return st;
  }
  *out = gr;
  return kOk;
}

int jubik_graph_launch(jubik_graph* graph, void* stream) {
  if (!graph || !graph->exec) return fail(kInvalidArgument, "NULL graph");
  DeviceGuard dg(graph->device);
  JUBIK_HIP(hipGraphLaunch(graph->exec, as_stream(stream)));
  for (jubik_model* m : graph->points) m->have_point = true;
  return kOk;
}

int jubik_graph_capture_batch(jubik_graph** out, const jubik_grad_desc* descs, int64_t count) {
  if (!out || !descs) return fail(kInvalidArgument, "NULL argument");
  *out = nullptr;
  if (count < 1 || count > 64) return fail(kInvalidArgument, "count must be in [1, 64]");
  if (timing_enabled())
    return fail(kInvalidArgument, "disable kernel timing before capturing a graph");
  for (int64_t i = 0; i < count; ++i) {
    const jubik_grad_desc& d = descs[i];
    const std::string at = "descriptor " + std::to_string(i) + ": ";
    if (!d.model || !d.nll || !d.g_ref || !d.amp_ref || !d.xi_ref)
      return fail(kInvalidArgument, at + "NULL argument");
    if (d.samples < 0 || d.samples > d.model->max_samples)
      return fail(kInvalidArgument, at + "batch of " + std::to_string(d.samples) +
                                        " samples outside [0, max_samples = " +
                                        std::to_string(d.model->max_samples) + "]");
    int st = check_model(d.model);
    if (!st) st = check_alpha(d.model, d.amp_alpha, d.xi_alpha, d.g_alpha, d.dnu);
    if (st) return st;
    if (d.model->lp->device != descs[0].model->lp->device)
      return fail(kInvalidArgument, at + "models on different devices");
    // the evaluations run as parallel branches: no two may share workspaces
    for (int64_t k = 0; k < i; ++k) {
      const jubik_model* o = descs[k].model;
      if (o == d.model || o->lp == d.model->lp || o->gp == d.model->gp)
        return fail(kInvalidArgument, at + "shares its model or a plan with descriptor " +
                                          std::to_string(k));
    }
  }
  DeviceGuard dg(descs[0].model->lp->device);
  const int n = (int)count;
  hipStream_t cs = nullptr;
  std::vector<hipStream_t> br(n, nullptr);
  std::vector<hipEvent_t> ev(n + 1, nullptr);
  int st = check_hip(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreate");
  for (int i = 0; i < n && !st; ++i)
    st = check_hip(hipStreamCreateWithFlags(&br[i], hipStreamNonBlocking), "hipStreamCreate");
  for (int i = 0; i <= n && !st; ++i)
    st = check_hip(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming), "hipEventCreate");
  auto* gr = new jubik_graph();
  gr->device = descs[0].model->lp->device;
  if (!st) st = check_hip(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal),
                          "hipStreamBeginCapture");
  if (!st) {
    // fork: every branch waits on the origin stream, runs its evaluation
    // and is joined back before the capture ends
    int cst = check_hip(hipEventRecord(ev[0], cs), "hipEventRecord");
    for (int i = 0; i < n && !cst; ++i) {
      const jubik_grad_desc& d = descs[i];
      cst = check_hip(hipStreamWaitEvent(br[i], ev[0], 0), "hipStreamWaitEvent");
      if (cst) break;
      cst = d.samples == 0
                ? model_grad(d.model, 1, 0, d.amp_ref, d.amp_alpha, d.xi_ref, d.xi_alpha, d.zm,
                             d.mu, d.dnu, d.nll, d.g_ref, d.g_alpha, br[i])
                : model_grad(d.model, (int)d.samples, 1, d.amp_ref, d.amp_alpha, d.xi_ref,
                             d.xi_alpha, d.zm, d.mu, d.dnu, d.nll, d.g_ref, d.g_alpha, br[i]);
      const int jst = check_hip(hipEventRecord(ev[i + 1], br[i]), "hipEventRecord");
      const int wst = jst ? jst : check_hip(hipStreamWaitEvent(cs, ev[i + 1], 0),
                                            "hipStreamWaitEvent");
      if (!cst) cst = wst;
      if (d.samples == 0) gr->points.push_back(d.model);
    }
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(cs, &g);
    gr->graph = g;
    st = cst;
    if (!st && e2 != hipSuccess) st = check_hip(e2, "hipStreamEndCapture");
    if (!st) st = check_hip(hipGraphInstantiate(&gr->exec, gr->graph, nullptr, nullptr, 0),
                            "hipGraphInstantiate");
  }
  for (hipEvent_t e : ev)
    if (e) (void)hipEventDestroy(e);
  for (hipStream_t b : br)
    if (b) (void)hipStreamDestroy(b);
  if (cs) (void)hipStreamDestroy(cs);
  if (st) {
    if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
    if (gr->graph) (void)hipGraphDestroy(gr->graph);
    delete gr;
    return st;
  }
  *out = gr;
  return kOk;
}

int jubik_graph_destroy(jubik_graph* graph) {
  if (!graph) return kOk;
  DeviceGuard dg(graph->device);
  if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
  if (graph->graph) (void)hipGraphDestroy(graph->graph);
  delete graph;
  return kOk;
}

namespace {

// Metric products of the model.  per_sample = false: S tangents v[s] at the
// base point of the last jubik_model_grad (out[s] = J^T diag(1/lam) J v[s]).
// per_sample = true: one tangent v at each of the S positions of the last
// jubik_model_grad_batch, summed: out = sum_s J_s^T diag(1/lam_s) J_s v.
int model_metric(jubik_model* model, int S, bool per_sample, const void* amp_ref,
                 const void* amp_alpha, const void* dnu, const void* v_ref, const void* v_alpha,
                 void* out_ref, void* out_alpha, hipStream_t s) {
  jubik_linpatch* lp = model->lp;
  jubik_gp* gp = model->gp;
  GpGeom gg = gp->g;
  gg.S = S;
  gg.sky_shared = per_sample ? 0 : 1;
  gg.work_shared = per_sample ? 1 : 0;
  LinGeom g = lp->g;
  g.S = S;
  g.sky_shared = per_sample ? 0 : 1;
  const int slot = per_sample ? 1 : 0;
  const void* sky = model->sky_slot(slot);
  int st = gp_tangent(gp->dtype, gg, model->gbufs(), amp_ref, amp_alpha, v_ref, v_alpha, dnu, sky,
                      model->dsky, s);
  if (st) return st;
  st = linpatch_forward_metric(lp->dtype, g, model->lbufs(), model->dsky, model->pixel_area,
                               model->expo, model->counts, model->lam_slot(slot), model->resid, s);
  if (st) return st;
  st = linpatch_adjoint_gp(lp->dtype, g, model->lbufs(), model->resid, model->pixel_area, sky,
                           gp->g.has_alpha ? dnu : nullptr, model->u, s);
  if (st) return st;
  gg.ucomplex = 1;
  if (per_sample) {
    // one GP adjoint of the cotangent planes of every sample (it is linear)
    gg.S = 1;
    gg.uplanes = S;
  } else {
    gg.uplanes = 1;
  }
  return gp_adjoint_u(gp->dtype, gg, model->gbufs(), amp_ref, amp_alpha, dnu, model->u, out_ref,
                      out_alpha, s);
}

// The cross operators of geoVI at the base point p (slot 0) and the sample
// point x (slot 1): to_p = true: K v = J(p)^T D_p D_x J(x) v, to_p = false:
// K^T v = J(x)^T D_x D_p J(p) v, D = diag(lam^-1/2) on read-out pixels: the
// metric pass with the tangent taken at one sky, the lam scaling
// sqrt(lam_x lam_p) and the adjoint at the other sky.
int model_cross(jubik_model* model, bool to_p, const void* amp_ref, const void* amp_alpha,
                const void* dnu, const void* v_ref, const void* v_alpha, void* out_ref,
                void* out_alpha, hipStream_t s) {
  jubik_linpatch* lp = model->lp;
  jubik_gp* gp = model->gp;
  GpGeom gg = gp->g;
  gg.S = 1;
  gg.sky_shared = 1;
  LinGeom g = lp->g;
  g.S = 1;
  g.sky_shared = 1;
  const void* sky_t = model->sky_slot(to_p ? 1 : 0);
  const void* sky_a = model->sky_slot(to_p ? 0 : 1);
  int st = gp_tangent(gp->dtype, gg, model->gbufs(), amp_ref, amp_alpha, v_ref, v_alpha, dnu,
                      sky_t, model->dsky, s);
  if (st) return st;
  st = linpatch_forward_metric(lp->dtype, g, model->lbufs(), model->dsky, model->pixel_area,
                               model->expo, model->counts, model->geo, model->resid, s);
  if (st) return st;
  st = linpatch_adjoint_gp(lp->dtype, g, model->lbufs(), model->resid, model->pixel_area, sky_a,
                           gp->g.has_alpha ? dnu : nullptr, model->u, s);
  if (st) return st;
  gg.uplanes = 1;
  gg.ucomplex = 1;
  return gp_adjoint_u(gp->dtype, gg, model->gbufs(), amp_ref, amp_alpha, dnu, model->u, out_ref,
                      out_alpha, s);
}

// out = K v + K^T v + K^T K v: the Gauss-Newton metric of the geoVI residual
// r(x) = (x - p) + J(p)^T D_p 2 (sqrt lam(x) - sqrt lam(p)) - m is
// (I + K)^T (I + K) = I + this.
int geo_metric(jubik_model* model, const void* amp_ref, const void* amp_alpha, const void* dnu,
               const void* v_ref, const void* v_alpha, void* out_ref, void* out_alpha,
               hipStream_t s) {
  const int dtype = model->lp->dtype;
  const int64_t n = (int64_t)model->gp->g.N * model->gp->g.N;
  const size_t es = elem_size(dtype);
  char* gv = static_cast<char*>(model->geo_vec);
  void* kv[2] = {gv, v_alpha ? gv + n * es : nullptr};              // K v
  void* kkv[2] = {gv + 2 * n * es, v_alpha ? gv + 3 * n * es : nullptr};  // K^T K v
  int st = model_cross(model, true, amp_ref, amp_alpha, dnu, v_ref, v_alpha, kv[0], kv[1], s);
  if (!st) st = model_cross(model, false, amp_ref, amp_alpha, dnu, v_ref, v_alpha, out_ref,
                            out_alpha, s);
  if (!st) st = model_cross(model, false, amp_ref, amp_alpha, dnu, kv[0], kv[1], kkv[0], kkv[1],
                            s);
  for (int f = 0; f < 2 && !st; ++f)
    if (f == 0 || v_alpha)
      st = vec_sum3(dtype, n, f ? out_alpha : out_ref, kv[f], kkv[f], f ? out_alpha : out_ref, s);
  return st;
}

int check_batch(const jubik_model* model, int64_t S) {
  if (S < 1 || S > model->max_samples)
    return fail(kInvalidArgument, "batch of " + std::to_string(S) +
                                      " samples outside [1, max_samples = " +
                                      std::to_string(model->max_samples) + "]");
  return kOk;
}

int check_samples_point(const jubik_model* model, int64_t S) {
  if (model->batch_S < 1)
    return fail(kNotReady, "no sample positions (call jubik_model_grad_batch first)");
  if (S != model->batch_S)
    return fail(kInvalidArgument, "the last batched gradient evaluated " +
                                      std::to_string(model->batch_S) + " samples, not " +
                                      std::to_string(S));
  return kOk;
}

// Conjugate gradients on (I + M) x = b, M the metric of model_metric: S
// systems at the base point (per_sample = false) or one system with M the
// mean over the S sample positions (per_sample = true).
// geo = true: one system (I + K + K^T + K^T K) x = b of geoVI (geo_metric).
int model_solve(jubik_model* model, int S, bool per_sample, const void* amp_ref,
                const void* amp_alpha, const void* dnu, const void* b_ref, const void* b_alpha,
                void* x_ref, void* x_alpha, int maxiter, double rtol, double absdelta,
                int check_every, double* info, hipStream_t s, bool geo = false) {
  const int dtype = model->lp->dtype;
  const int nsys = per_sample || geo ? 1 : S;
  const bool alpha = model->gp->g.has_alpha != 0;
  const int64_t n = (int64_t)model->gp->g.N * model->gp->g.N;
  const size_t es = elem_size(dtype);
  const size_t vec = (size_t)model->max_samples * 2 * n * es;  // one solver vector
  char* base = static_cast<char*>(model->cg_vec);
  auto vecs = [&](int k) {
    CgVec v;
    v.f[0] = base + k * vec;
    v.f[1] = alpha ? base + k * vec + (size_t)nsys * n * es : nullptr;
    return v;
  };
  CgVec r = vecs(0), p = vecs(1), Ap = vecs(2), Mv = vecs(3);
  CgVec b, x;
  b.f[0] = const_cast<void*>(b_ref);
  b.f[1] = const_cast<void*>(b_alpha);
  x.f[0] = x_ref;
  x.f[1] = x_alpha;
  CgDims d{nsys, alpha ? 2 : 1, n, rtol, absdelta};
  double* work = static_cast<double*>(model->cg_work);
  int st = cg_init(dtype, d, b, x, r, p, work, s);
  for (int it = 0; it < maxiter && !st; ++it) {
    st = geo ? geo_metric(model, amp_ref, amp_alpha, dnu, p.f[0], p.f[1], Mv.f[0], Mv.f[1], s)
             : model_metric(model, S, per_sample, amp_ref, amp_alpha, dnu, p.f[0], p.f[1],
                            Mv.f[0], Mv.f[1], s);
    if (!st) st = cg_matvec(dtype, d, p, Mv, per_sample && !geo ? 1.0 / S : 1.0, Ap, work, s);
    if (!st) st = cg_update(dtype, d, b, x, r, p, Ap, work, s);
    if (!st && check_every > 0 && (it + 1) % check_every == 0 && it + 1 < maxiter) {
      int active = 0;
      st = cg_active_count(d, work, s, &active);
      if (!st && active == 0) break;
    }
  }
  if (!st && info) st = cg_finish(d, work, info, s);
  return st;
}

int check_solve_args(const jubik_model* model, int maxiter, double rtol, int check_every) {
  if (!model->cg_vec) return fail(kNotReady, "no solver workspace (call jubik_model_reserve_solver)");
  if (maxiter < 0) return fail(kInvalidArgument, "maxiter must be non-negative");
  if (!(rtol >= 0.0)) return fail(kInvalidArgument, "rtol must be non-negative");
  if (check_every < 0) return fail(kInvalidArgument, "check_every must be non-negative");
  return kOk;
}

}  // namespace

int jubik_model_metric_batch(jubik_model* model, int64_t S, const void* amp_ref,
                             const void* amp_alpha, const void* dnu, const void* v_ref,
                             const void* v_alpha, void* out_ref, void* out_alpha, void* stream) {
  if (!model || !v_ref || !out_ref || !amp_ref) return fail(kInvalidArgument, "NULL argument");
  int st = check_batch(model, S);
  if (!st) st = check_alpha(model, amp_alpha, v_alpha, out_alpha, dnu);
  if (st) return st;
  if (!model->have_point)
    return fail(kNotReady, "no base point (call jubik_model_grad first)");
  DeviceGuard dg(model->lp->device);
  return model_metric(model, (int)S, false, amp_ref, amp_alpha, dnu, v_ref, v_alpha, out_ref,
                      out_alpha, as_stream(stream));
}

int jubik_model_geovi_point(jubik_model* model, const void* amp_ref, const void* amp_alpha,
                            const void* x_ref, const void* x_alpha, double zm, double mu,
                            const void* dnu, double* nll, void* stream) {
  if (!model || !amp_ref || !x_ref || !nll) return fail(kInvalidArgument, "NULL argument");
  int st = check_model(model);
  if (!st) st = check_alpha(model, amp_alpha, x_alpha, x_alpha, dnu);
  if (st) return st;
  if (!model->have_point) return fail(kNotReady, "no base point (call jubik_model_grad first)");
  if (model->max_samples < 1) return fail(kInvalidArgument, "model has no sample slot");
  DeviceGuard dg(model->lp->device);
  const size_t es = elem_size(model->lp->dtype);
  const int64_t npix = (int64_t)model->lp->g.J * model->lp->g.H * model->lp->g.W;
  const int64_t n = (int64_t)model->gp->g.N * model->gp->g.N;
  if (!model->geo) st = dev_alloc(&model->geo, 2 * npix * es, "geoVI scalings");
  if (!st && !model->geo_vec) st = dev_alloc(&model->geo_vec, 4 * n * es, "geoVI vectors");
  if (st) return st;
  model->geo_ready = false;
  // the evaluation at x into slot 1 (its gradient lands in geo_vec and is
  // not used); the base point stays in slot 0
  char* gv = static_cast<char*>(model->geo_vec);
  st = model_grad(model, 1, 1, amp_ref, amp_alpha, x_ref, x_alpha, zm, mu, dnu, nll, gv,
                  model->gp->g.has_alpha ? gv + n * es : nullptr, as_stream(stream));
  if (!st)
    st = geo_prep(model->lp->dtype, model->counts, model->lam_slot(0), model->lam_slot(1),
                  model->geo, static_cast<char*>(model->geo) + npix * es, npix,
                  as_stream(stream));
  if (!st) model->geo_ready = true;
  return st;
}

int jubik_model_geovi_apply(jubik_model* model, int mode, const void* amp_ref,
                            const void* amp_alpha, const void* dnu, const void* v_ref,
                            const void* v_alpha, void* out_ref, void* out_alpha, void* stream) {
  if (!model || !amp_ref || !out_ref || (mode != 2 && !v_ref))
    return fail(kInvalidArgument, "NULL argument");
  if (mode < 0 || mode > 3) return fail(kInvalidArgument, "geoVI mode must be 0 .. 3");
  int st = check_alpha(model, amp_alpha, mode == 2 ? out_alpha : v_alpha, out_alpha, dnu);
  if (st) return st;
  if (!model->geo_ready || !model->have_point)
    return fail(kNotReady, "no geoVI sample point (call jubik_model_geovi_point)");
  DeviceGuard dg(model->lp->device);
  hipStream_t s = as_stream(stream);
  if (mode == 0 || mode == 1)
    return model_cross(model, mode == 0, amp_ref, amp_alpha, dnu, v_ref, v_alpha, out_ref,
                       out_alpha, s);
  if (mode == 3)
    return geo_metric(model, amp_ref, amp_alpha, dnu, v_ref, v_alpha, out_ref, out_alpha, s);
  // mode 2: J(p)^T D_p eta
  jubik_linpatch* lp = model->lp;
  jubik_gp* gp = model->gp;
  LinGeom g = lp->g;
  g.S = 1;
  g.sky_shared = 1;
  GpGeom gg = gp->g;
  gg.S = 1;
  gg.sky_shared = 1;
  gg.uplanes = 1;
  gg.ucomplex = 1;
  const int64_t npix = (int64_t)g.J * g.H * g.W;
  const void* eta = static_cast<char*>(model->geo) + npix * elem_size(lp->dtype);
  st = sqrt_metric_cot(lp->dtype, model->expo, model->counts, model->lam_slot(0), eta,
                       model->resid, npix, 1, s);
  if (!st)
    st = linpatch_adjoint_gp(lp->dtype, g, model->lbufs(), model->resid, model->pixel_area,
                             model->sky_slot(0), gp->g.has_alpha ? dnu : nullptr, model->u, s);
  if (!st)
    st = gp_adjoint_u(gp->dtype, gg, model->gbufs(), amp_ref, amp_alpha, dnu, model->u, out_ref,
                      out_alpha, s);
  return st;
}

int jubik_model_geovi_solve(jubik_model* model, const void* amp_ref, const void* amp_alpha,
                            const void* dnu, const void* b_ref, const void* b_alpha, void* x_ref,
                            void* x_alpha, int maxiter, double rtol, double absdelta,
                            int check_every, double* info, void* stream) {
  if (!model || !amp_ref || !b_ref || !x_ref) return fail(kInvalidArgument, "NULL argument");
  int st = check_alpha(model, amp_alpha, b_alpha, x_alpha, dnu);
  if (!st) st = check_solve_args(model, maxiter, rtol, check_every);
  if (st) return st;
  if (!model->geo_ready || !model->have_point)
    return fail(kNotReady, "no geoVI sample point (call jubik_model_geovi_point)");
  DeviceGuard dg(model->lp->device);
  return model_solve(model, 1, false, amp_ref, amp_alpha, dnu, b_ref, b_alpha, x_ref, x_alpha,
                     maxiter, rtol, absdelta, check_every, info, as_stream(stream), true);
}

int jubik_model_metric_samples(jubik_model* model, int64_t S, const void* amp_ref,
                               const void* amp_alpha, const void* dnu, const void* v_ref,
                               const void* v_alpha, void* out_ref, void* out_alpha,
                               void* stream) {
  if (!model || !v_ref || !out_ref || !amp_ref) return fail(kInvalidArgument, "NULL argument");
  int st = check_batch(model, S);
  if (!st) st = check_alpha(model, amp_alpha, v_alpha, out_alpha, dnu);
  if (!st) st = check_samples_point(model, S);
  if (st) return st;
  DeviceGuard dg(model->lp->device);
  return model_metric(model, (int)S, true, amp_ref, amp_alpha, dnu, v_ref, v_alpha, out_ref,
                      out_alpha, as_stream(stream));
}

int jubik_model_reserve_solver(jubik_model* model) {
  if (!model) return fail(kInvalidArgument, "NULL argument");
  if (model->cg_vec) return kOk;
  DeviceGuard dg(model->lp->device);
  const int64_t n = (int64_t)model->gp->g.N * model->gp->g.N;
  const size_t vec = (size_t)model->max_samples * 2 * n * elem_size(model->lp->dtype);
  int st = dev_alloc(&model->cg_vec, 4 * vec, "solver vectors");
  if (!st)
    st = dev_alloc(&model->cg_work, cg_work_doubles(model->max_samples) * sizeof(double),
                   "solver scalars");
  if (st) {
    dev_free(model->cg_vec);
    dev_free(model->cg_work);
  }
  return st;
}

int jubik_model_metric_solve(jubik_model* model, int64_t S, const void* amp_ref,
                             const void* amp_alpha, const void* dnu, const void* b_ref,
                             const void* b_alpha, void* x_ref, void* x_alpha, int maxiter,
                             double rtol, double absdelta, int check_every, double* info,
                             void* stream) {
  if (!model || !b_ref || !x_ref || !amp_ref) return fail(kInvalidArgument, "NULL argument");
  int st = check_batch(model, S);
  if (!st) st = check_alpha(model, amp_alpha, b_alpha, x_alpha, dnu);
  if (!st) st = check_solve_args(model, maxiter, rtol, check_every);
  if (st) return st;
  if (!model->have_point)
    return fail(kNotReady, "no base point (call jubik_model_grad first)");
  DeviceGuard dg(model->lp->device);
  return model_solve(model, (int)S, false, amp_ref, amp_alpha, dnu, b_ref, b_alpha, x_ref,
                     x_alpha, maxiter, rtol, absdelta, check_every, info, as_stream(stream));
}

int jubik_model_kl_solve(jubik_model* model, int64_t S, const void* amp_ref,
                         const void* amp_alpha, const void* dnu, const void* b_ref,
                         const void* b_alpha, void* x_ref, void* x_alpha, int maxiter,
                         double rtol, double absdelta, int check_every, double* info,
                         void* stream) {
  if (!model || !b_ref || !x_ref || !amp_ref) return fail(kInvalidArgument, "NULL argument");
  int st = check_batch(model, S);
  if (!st) st = check_alpha(model, amp_alpha, b_alpha, x_alpha, dnu);
  if (!st) st = check_solve_args(model, maxiter, rtol, check_every);
  if (!st) st = check_samples_point(model, S);
  if (st) return st;
  DeviceGuard dg(model->lp->device);
  return model_solve(model, (int)S, true, amp_ref, amp_alpha, dnu, b_ref, b_alpha, x_ref,
                     x_alpha, maxiter, rtol, absdelta, check_every, info, as_stream(stream));
}

int jubik_model_sqrt_metric_batch(jubik_model* model, int64_t S, const void* amp_ref,
                                  const void* amp_alpha, const void* dnu, const void* noise,
                                  void* out_ref, void* out_alpha, void* stream) {
  if (!model || !noise || !out_ref || !amp_ref) return fail(kInvalidArgument, "NULL argument");
  if (S < 1 || S > model->max_samples)
    return fail(kInvalidArgument, "batch of " + std::to_string(S) +
                                      " samples outside [1, max_samples = " +
                                      std::to_string(model->max_samples) + "]");
  if (model->gp->g.has_alpha && (!amp_alpha || !out_alpha || !dnu))
    return fail(kInvalidArgument, "spectral-index inputs are NULL");
  if (!model->have_point)
    return fail(kNotReady, "no base point (call jubik_model_grad first)");
  jubik_linpatch* lp = model->lp;
  jubik_gp* gp = model->gp;
  DeviceGuard dg(lp->device);
  hipStream_t s = as_stream(stream);
  LinGeom g = lp->g;
  g.S = (int)S;
  g.sky_shared = 1;
  GpGeom gg = gp->g;
  gg.S = (int)S;
  gg.sky_shared = 1;
  gg.uplanes = 1;
  gg.ucomplex = 1;
  const int64_t n = (int64_t)g.J * g.H * g.W;
  int st = sqrt_metric_cot(lp->dtype, model->expo, model->counts, model->lam_slot(0), noise,
                           model->resid, n, (int)S, s);
  if (st) return st;
  st = linpatch_adjoint_gp(lp->dtype, g, model->lbufs(), model->resid, model->pixel_area,
                           model->sky_slot(0), gp->g.has_alpha ? dnu : nullptr, model->u, s);
  if (st) return st;
  return gp_adjoint_u(gp->dtype, gg, model->gbufs(), amp_ref, amp_alpha, dnu, model->u, out_ref,
                      out_alpha, s);
}

int jubik_model_metric(jubik_model* model, const void* amp_ref, const void* amp_alpha,
                       const void* dnu, const void* v_ref, const void* v_alpha, void* out_ref,
                       void* out_alpha, void* stream) {
  return jubik_model_metric_batch(model, 1, amp_ref, amp_alpha, dnu, v_ref, v_alpha, out_ref,
                                  out_alpha, stream);
}

int jubik_model_check_domain(jubik_model* model, int64_t* counts, void* stream) {
  if (!model) return fail(kInvalidArgument, "NULL argument");
  if (model->last_S < 1) return fail(kNotReady, "no gradient evaluation yet");
  DeviceGuard dg(model->lp->device);
  std::vector<double> nb(model->last_S);
  hipStream_t s = as_stream(stream);
  JUBIK_HIP(hipMemcpyAsync(nb.data(), model->nbad, nb.size() * sizeof(double),
                           hipMemcpyDeviceToHost, s));
  JUBIK_HIP(hipStreamSynchronize(s));
  std::string bad;
  for (size_t k = 0; k < nb.size(); ++k) {
    if (counts) counts[k] = (int64_t)nb[k];
    if (nb[k] > 0)
      bad += (bad.empty() ? "" : ", ") + std::to_string((int64_t)nb[k]) + " at sample " +
             std::to_string(k);
  }
  if (!bad.empty())
    return fail(kDomain, "lambda <= 0 at read-out pixels (" + bad +
                             "): the Poisson NLL is inf or NaN there; mask pixels with zero "
                             "exposure");
  return kOk;
}

int jubik_graph_capture_metric_batch(jubik_graph** out, const jubik_metric_desc* descs,
                                     int64_t count) {
  if (!out || !descs) return fail(kInvalidArgument, "NULL argument");
  *out = nullptr;
  if (count < 1 || count > 64) return fail(kInvalidArgument, "count must be in [1, 64]");
  if (timing_enabled())
    return fail(kInvalidArgument, "disable kernel timing before capturing a graph");
  for (int64_t i = 0; i < count; ++i) {
    const jubik_metric_desc& d = descs[i];
    const std::string at = "descriptor " + std::to_string(i) + ": ";
    if (!d.model || !d.v_ref || !d.out_ref || !d.amp_ref)
      return fail(kInvalidArgument, at + "NULL argument");
    int st = check_batch(d.model, d.samples);
    if (!st) st = check_alpha(d.model, d.amp_alpha, d.v_alpha, d.out_alpha, d.dnu);
    if (st) return st;
    if (!d.model->have_point)
      return fail(kNotReady, at + "no base point (call jubik_model_grad first)");
    if (d.model->lp->device != descs[0].model->lp->device)
      return fail(kInvalidArgument, at + "models on different devices");
    for (int64_t k = 0; k < i; ++k) {
      const jubik_model* o = descs[k].model;
      if (o == d.model || o->lp == d.model->lp || o->gp == d.model->gp)
        return fail(kInvalidArgument, at + "shares its model or a plan with descriptor " +
                                          std::to_string(k));
    }
  }
  DeviceGuard dg(descs[0].model->lp->device);
  const int n = (int)count;
  hipStream_t cs = nullptr;
  std::vector<hipStream_t> br(n, nullptr);
  std::vector<hipEvent_t> ev(n + 1, nullptr);
  int st = check_hip(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreate");
  for (int i = 0; i < n && !st; ++i)
    st = check_hip(hipStreamCreateWithFlags(&br[i], hipStreamNonBlocking), "hipStreamCreate");
  for (int i = 0; i <= n && !st; ++i)
    st = check_hip(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming), "hipEventCreate");
  auto* gr = new jubik_graph();
  gr->device = descs[0].model->lp->device;
  if (!st) st = check_hip(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal),
                          "hipStreamBeginCapture");
  if (!st) {
    int cst = check_hip(hipEventRecord(ev[0], cs), "hipEventRecord");
    for (int i = 0; i < n && !cst; ++i) {
      const jubik_metric_desc& d = descs[i];
      cst = check_hip(hipStreamWaitEvent(br[i], ev[0], 0), "hipStreamWaitEvent");
      if (cst) break;
      cst = model_metric(d.model, (int)d.samples, false, d.amp_ref, d.amp_alpha, d.dnu, d.v_ref,
                         d.v_alpha, d.out_ref, d.out_alpha, br[i]);
      const int jst = check_hip(hipEventRecord(ev[i + 1], br[i]), "hipEventRecord");
      const int wst = jst ? jst : check_hip(hipStreamWaitEvent(cs, ev[i + 1], 0),
                                            "hipStreamWaitEvent");
      if (!cst) cst = wst;
    }
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(cs, &g);
    gr->graph = g;
    st = cst;
    if (!st && e2 != hipSuccess) st = check_hip(e2, "hipStreamEndCapture");
    if (!st) st = check_hip(hipGraphInstantiate(&gr->exec, gr->graph, nullptr, nullptr, 0),
                            "hipGraphInstantiate");
  }
  for (hipEvent_t e : ev)
    if (e) (void)hipEventDestroy(e);
  for (hipStream_t b : br)
    if (b) (void)hipStreamDestroy(b);
  if (cs) (void)hipStreamDestroy(cs);
  if (st) {
    if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
    if (gr->graph) (void)hipGraphDestroy(gr->graph);
    delete gr;
    return st;
  }
  *out = gr;
  return kOk;
}

}  // extern "C"
