// rt_capi.hip -- the C-ABI of libemma_rt.so (declared in include/emma_rt.h):
// device management, parameters, the device-resident level tree, RadSolver
// for one level, and the coarse-step recursion captured as a HIP graph.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "emma_rt.h"
#include "rt_kernels.h"
#include "rt_sched.h"
#include "rt_stencil.h"

using namespace emma;

#define HIPCHK(x)                         \
  do {                                    \
    hipError_t e_ = (x);                  \
    if (e_ != hipSuccess) {               \
      last_hip_error() = e_;              \
      return EMMA_RT_ERR_HIP;             \
    }                                     \
  } while (0)

static hipError_t& last_hip_error() {
  static thread_local hipError_t e = hipSuccess;
  return e;
}

static_assert(emma::kItemPrio == emma::kSchedPrio && emma::kItemAcquireAgent == emma::kSchedAcquireAgent &&
                  emma::kItemReleaseAgent == emma::kSchedReleaseAgent,
              "item flags of the planner and the persistent kernel");
static_assert(emma::kNeiTransmissive == EMMA_RT_NEI_TRANSMISSIVE && emma::kNeiReflective == EMMA_RT_NEI_REFLECTIVE,
              "domain-face codes of the kernels and the C-ABI");
static_assert(emma::kListed == EMMA_RT_VARIANT_LISTED && emma::kPerDir == EMMA_RT_VARIANT_PERDIR &&
                  emma::kLatency == EMMA_RT_VARIANT_LATENCY,
              "halo variants of the kernels and the C-ABI");

namespace {

struct DevLevel {
  int level = 0, noct = 0, ncell = 0, cur = 0;
  int npair = 0;  // reflux pairs: (leaf, face) with a finer oct behind the face
  int nslot = 0;  // accumulators per flux component and group: npair x sub-slots
  int halo_per_dir = 0;  // some wave has more than 64 halo faces (no kListed variant)
  int variant = kListed;  // halo variant of the tiled kernel
  int chem_budget = 0;   // chemistry tries before a leaf goes to the straggler pass (0: none)
  int morton_m = 0;      // > 0: a complete periodic cube of 2^morton_m octs a side in Morton order
  int raster = 0;        // the tiled kernel walks the cube's tiles in raster order (morton_m > 0)
  float chem_keep = 0.f;  // long chains stay (see emma_rt_set_chem_keep)
  void* e[2] = {nullptr, nullptr};
  void* fx[2] = {nullptr, nullptr};
  void* fy[2] = {nullptr, nullptr};
  void* fz[2] = {nullptr, nullptr};
  void* src = nullptr;
  double* nhplus[2] = {nullptr, nullptr};
  double* eint[2] = {nullptr, nullptr};
  double* temp[2] = {nullptr, nullptr};
  double* nh = nullptr;
  int* child = nullptr;
  int* nei = nullptr;
  int* rpair = nullptr;  // [noct][6] reflux pair of each face in the coarser level, -1 if none
  void* racc = nullptr;
  double* t_first = nullptr;
  double* t_last = nullptr;
  std::vector<int> hchild;  // host copy with the reflux pairs encoded
  std::vector<int> hnei;    // host copy of nei_oct (the persistent schedule's planner)
  std::vector<int> hrpair;
  // hydro state (struct Wtype) for the WRADHYD coupling, and the octs' lower
  // corners for the test source; allocated on first use
  double* hyd[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};  // d u v w p dX E
  double* opos = nullptr;
  // star particles of the level's leaf cells (CSR over cells), set by
  // emma_rt_set_stars; star_first == nullptr: none
  int* star_first = nullptr;
  double* star_mass = nullptr;
  double* star_age = nullptr;
  int* star_rs = nullptr;
  int* star_is = nullptr;
  // device-side restart copy (emma_rt_save_state): one allocation, the
  // parts in the order of snap_parts(), and the buffer parity it was taken at
  char* snap = nullptr;
  size_t snap_bytes = 0;
  int snap_cur = -1;
};

// one launch of a cached coarse step: a persistent segment (consecutive level
// steps) or a COARSERAD transfer (putE2coarse / distribE) with its tree view
struct SchedLaunch {
  int kind = 0;  // 0 persistent segment, 1 putE2coarse, 2 distribE
  SchedArgs args;
  int blocks = 0;
  bool coupled = false;
  std::vector<unsigned char> tv;  // TreeView<R>
};

// a captured coarse step and what it needs while it exists: the buffer
// parity after it, and for the persistent schedule the device tables of its
// segments and their launches (replayed eagerly in EMMA_RT_TIMING_LAUNCH)
struct Cached {
  hipGraphExec_t ex = nullptr;
  std::vector<int> endpar;
  std::vector<void*> bufs;
  std::vector<SchedLaunch> launches;
  emma_rt_sched_info info;
};

// an operation of the recorded coarse step (persistent schedule)
struct SchedOp {
  int kind = 0;  // 0 level step (StepDesc<R>), 1 putE2coarse, 2 distribE (TreeView<R>)
  int k = 0;
  std::vector<unsigned char> blob;
};

struct GraphKey {
  int ncoarse;
  double aexp;
  std::vector<int> parity;
  int xion;
  int mode;
  int sched;
  int counters;
  bool operator<(const GraphKey& o) const {
    if (sched != o.sched) return sched < o.sched;
    if (counters != o.counters) return counters < o.counters;
    if (mode != o.mode) return mode < o.mode;
    if (ncoarse != o.ncoarse) return ncoarse < o.ncoarse;
    if (aexp != o.aexp) return aexp < o.aexp;
    if (xion != o.xion) return xion < o.xion;
    return parity < o.parity;
  }
};

}  // namespace

struct emma_rt_ctx {
  emma_rt_params p;
  int device = 0;
  hipStream_t stream = nullptr;
  int precision = EMMA_RT_FP64;
  int mode = EMMA_RT_RADSTEP;
  int rbytes = 8;
  int lmin = 0, lmax = -1;
  bool has_tree = false;
  std::vector<DevLevel> lev;  // index level - lmin
  int* d_status = nullptr;
  unsigned long long* d_counters = nullptr;
  // page-locked copies of the status word and the counters, read back by
  // end_call with asynchronous copies behind the call's work and one wait
  int* h_status = nullptr;
  unsigned long long* h_counters = nullptr;
  void* scratch = nullptr;  // this device's no-alloc staging area (emma_rt_advance_stencil_noalloc)
  double* d_photons = nullptr;
  // straggler list of the tiled mixed kernel (capacity: the largest level's
  // cells) and one list-length word per level-step launch of a call (slot
  // defer_seq, counted from 0 in every call; all cleared before the call)
  int* d_defer = nullptr;
  int* d_defer_ctl = nullptr;
  int defer_cap = 0;
  int defer_slots = 0;
  int defer_seq = 0;
  emma_rt_stats stats;
  int timing = 0;
  int counters = 1;  // the kernels' device counters (sub-steps, leaves, deferred cells)
  int xion = 0;
  double tphy = 0.;
  // FillRad / final-update coupling (WRADHYD) and the Stromgren test source
  int hydro = 0;
  int star = 0;
  double star_x = 0., star_y = 0., star_z = 0., star_mass = 0., star_lifetime = 0., star_srcint = 0.;
  // star-particle sources (stars_sources): srcint, tlife and the
  // DECREASE_EMMISIVITY_AFTER_TLIFE switch of the reference
  int stars = 0, stars_decrease = 0;
  double stars_srcint = 0., stars_tlife = 1.;
  std::vector<hipEvent_t> ev;  // pairs
  size_t ev_used = 0;
  std::vector<int> ev_level;   // level index of each pair, -1 for a whole graph replay
  double level_ms[kMaxLevels];  // device time per level of the last timed call
  int level_launches[kMaxLevels];
  std::map<GraphKey, Cached> graphs;
  // coarse-step schedule: the captured graph of level-step launches, or one
  // persistent kernel per run of level steps (EMMA_RT_SCHED_PERSISTENT)
  int sched = EMMA_RT_SCHED_GRAPH;
  int ncu = 0;  // compute units of the device
  int nxcc = 0;  // XCDs the work-groups run on (probed on first use), and their queue numbers
  int xcc_queue[16];
  emma_rt_sched_info sched_info;
  // recording of a coarse step (persistent schedule): the helpers below
  // record instead of launching while rec is set; FillRad's per-cell block
  // waits in pending_fill for its level's next step
  std::vector<SchedOp>* rec = nullptr;
  std::vector<unsigned char> pending_fill[kMaxLevels];
  // stencil path
  void* d_stencil = nullptr;
  int stride = 0;
  int stencil_rbytes = 0;
};

static void drop_graphs(emma_rt_ctx* c);
static void free_tree(emma_rt_ctx* c) {
  for (auto& L : c->lev) {
    for (int b = 0; b < 2; b++) {
      (void)hipFree(L.e[b]);
      (void)hipFree(L.fx[b]);
      (void)hipFree(L.fy[b]);
      (void)hipFree(L.fz[b]);
      (void)hipFree(L.nhplus[b]);
      (void)hipFree(L.eint[b]);
      (void)hipFree(L.temp[b]);
    }
    (void)hipFree(L.src);
    (void)hipFree(L.nh);
    (void)hipFree(L.child);
    (void)hipFree(L.nei);
    (void)hipFree(L.racc);
    (void)hipFree(L.rpair);
    (void)hipFree(L.t_first);
    (void)hipFree(L.t_last);
    for (int k = 0; k < 7; k++) (void)hipFree(L.hyd[k]);
    (void)hipFree(L.opos);
    (void)hipFree(L.snap);
    for (void* p : {(void*)L.star_first, (void*)L.star_mass, (void*)L.star_age, (void*)L.star_rs, (void*)L.star_is})
      (void)hipFree(p);
  }
  c->lev.clear();
  (void)hipFree(c->d_defer);
  (void)hipFree(c->d_defer_ctl);
  c->d_defer = nullptr;
  c->d_defer_ctl = nullptr;
  c->defer_cap = 0;
  c->defer_slots = 0;
  drop_graphs(c);
  c->has_tree = false;
}

static void drop_graphs(emma_rt_ctx* c) {
  if (c->graphs.empty()) return;
  (void)hipStreamSynchronize(c->stream);
  for (auto& kv : c->graphs) {
    if (kv.second.ex) (void)hipGraphExecDestroy(kv.second.ex);
    for (void* b : kv.second.bufs) (void)hipFree(b);
  }
  c->graphs.clear();
}

// chemrad's per-call constants (chem_utils.c:266-331), evaluated on the host
// the photon densities' floor per group (chem_utils.c:382, 772) and, for FP32
// storage, the smallest float at or above it
static double emin_floor(const emma_rt_params& p, int g) { return kEmin * p.factgrp[g] / p.unit_N; }
static float emin_floor_up(double f) {
  float up = (float)f;
  if ((double)up < f) up = nextafterf(up, INFINITY);
  return up;
}

static ChemConst make_chem(const emma_rt_params& p, double dtnew, double aexp) {
  ChemConst K;
  memset(&K, 0, sizeof(K));
  K.ngrp = p.ngrp;
  K.ncvgcool = p.ncvgcool;
  K.static_cosmo = (p.H0 == 0.) ? 1 : 0;
  K.fudgecool = p.fudgecool;
  K.unit_l = p.unit_l;
  K.unit_t = p.unit_t;
  K.unit_N = p.unit_N;
  K.unit_n = p.unit_n;
  K.unit_d = p.unit_d;
  K.unit_v = p.unit_v;
  K.aexp = aexp;
  K.a2 = aexp * aexp;
  K.a3 = aexp * aexp * aexp;
  K.pa2 = pow(aexp, 2);
  K.pa3 = pow(aexp, 3);
  K.pa4 = pow(aexp, 4);
  K.pa5 = pow(aexp, 5);
  K.pv2 = pow(p.unit_v, 2);
  K.pstar = p.unit_n * p.unit_d * pow(p.unit_v, 2);
  K.dt = dtnew * p.unit_t * pow(aexp, 2);
  K.emin = kPmin / (kGamma - 1.) / pow(aexp, 5) * p.unit_n * p.unit_d * pow(p.unit_v, 2);
  K.H0 = p.H0;
  K.om = p.om;
  K.ov = p.ov;
  K.k_e = p.unit_N / K.a3;
  K.k_f = p.unit_l / p.unit_t * p.unit_N / K.pa4;
  K.k_src = p.unit_N / p.unit_t / K.a2 / K.pa3;
  K.k_eint = p.unit_n * p.unit_d * K.pv2 / K.pa5;
  K.ki_e = K.pa3 / p.unit_N;
  K.ki_f = K.pa4 / p.unit_l * p.unit_t / p.unit_N;
  K.ki_eint = K.pa5 / (p.unit_n * p.unit_d * K.pv2);
  for (int g = 0; g < p.ngrp; g++) {
    K.heat[g] = p.alphae[g] * p.hnu[g] - p.alphai[g] * kHnu0;
    K.alphai[g] = p.alphai[g];
    K.factgrp[g] = p.factgrp[g];
    K.emin_grp[g] = emin_floor(p, g);
    K.emin_grp_up[g] = emin_floor_up(K.emin_grp[g]);
  }
  return K;
}

static double cloc_of(const emma_rt_params& p, double aexp) { return aexp * p.clight * 299792458. / p.unit_v; }

static double dtrad_of(const emma_rt_params& p, int level, double aexp) {
  return 0.85 / (aexp * p.clight * 299792458. / p.unit_v) / 3.0 / (double)(1 << level);
}

static int check_params(const emma_rt_params* p) {
  if (p == nullptr || p->ngrp < 1 || p->ngrp > EMMA_RT_MAX_GRP) return EMMA_RT_ERR_ARG;
  if (!(p->unit_t > 0) || !(p->unit_v > 0) || !(p->unit_N > 0) || !(p->clight > 0)) return EMMA_RT_ERR_ARG;
  // time groups of the atomic table (readAtomic, parameters.c:455-468):
  // ngrp = NGRP_SPACE x ngrp_time, bounds in years, non-decreasing
  if (p->ngrp_time < 1 || p->ngrp % p->ngrp_time != 0) return EMMA_RT_ERR_ARG;
  for (int i = 0; i < p->ngrp_time; i++)
    if (!(p->time_bound[i] <= p->time_bound[i + 1])) return EMMA_RT_ERR_ARG;
  if (p->lcoarse < 0) return EMMA_RT_ERR_ARG;
  return EMMA_RT_OK;
}

// ---------------------------------------------------------------- device mgmt
extern "C" int emma_rt_countdevices(int rank, int* count) {
  (void)rank;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (count) *count = (e == hipSuccess) ? n : 0;
  if (e != hipSuccess || n == 0) return EMMA_RT_ERR_NODEVICE;
  return EMMA_RT_OK;
}

extern "C" int emma_rt_initlocaldevice(int rank, int count) {
  if (count <= 0) return EMMA_RT_ERR_NODEVICE;
  HIPCHK(hipSetDevice(rank % count));
  return EMMA_RT_OK;
}

extern "C" int emma_rt_checkdevice(int rank, int* device) {
  (void)rank;
  int d = -1;
  HIPCHK(hipGetDevice(&d));
  if (device) *device = d;
  return EMMA_RT_OK;
}

extern "C" const char* emma_rt_strerror(int s) {
  switch (s) {
    case EMMA_RT_OK: return "ok";
    case EMMA_RT_ERR_ARG: return "bad argument";
    case EMMA_RT_ERR_HIP: return hipGetErrorString(last_hip_error());
    case EMMA_RT_ERR_NEG_PHOTON: return "negative or NaN photon density after transport";
    case EMMA_RT_ERR_EMIN: return "photon density below EMIN entering chemistry";
    case EMMA_RT_ERR_NOCONV: return "chemistry sub-stepping did not converge (fudgecool < 1e-20)";
    case EMMA_RT_ERR_NEIGHBOR: return "missing neighbour oct";
    case EMMA_RT_ERR_LAYOUT: return "struct layout mismatch";
    case EMMA_RT_ERR_NODEVICE: return "no GPU device";
    case EMMA_RT_ERR_STATE: return "call out of order";
    case EMMA_RT_ERR_SCHED: return "persistent coarse step: a dependency wait passed its bound";
    default: return "unknown status";
  }
}

// ---------------------------------------------------------------- context
// ---------------------------------------------------------------- no-alloc staging
// The staging area of emma_rt_advance_stencil_noalloc: a file-scope device
// array (each GPU the library runs on holds its own copy until the process
// ends) instead of the buffer emma_rt_create_radstencil allocates.  Host code
// never passes the array's name to a kernel or copy: it uses the device
// address hipGetSymbolAddress returns, looked up once per device with that
// device current (emma_rt_create) and kept per device.
static __device__ __align__(256) char scratch[EMMA_RT_NOALLOC_STENCIL_BYTES];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];

static int scratch_of_current_device(int dev, void** out) {
  if (dev < 0 || dev >= kMaxDevices) return EMMA_RT_ERR_ARG;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (ptr == nullptr) return EMMA_RT_ERR_HIP;
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return EMMA_RT_OK;
}

extern "C" int emma_rt_create(const emma_rt_params* p, emma_rt_ctx** out) {
  if (out == nullptr) return EMMA_RT_ERR_ARG;
  *out = nullptr;
  int st = check_params(p);
  if (st) return st;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return EMMA_RT_ERR_NODEVICE;
  emma_rt_ctx* c = new emma_rt_ctx();
  c->p = *p;
  memset(&c->stats, 0, sizeof(c->stats));
  memset(c->level_ms, 0, sizeof(c->level_ms));
  memset(c->level_launches, 0, sizeof(c->level_launches));
  if (hipGetDevice(&c->device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
      hipMalloc((void**)&c->d_status, 4 * sizeof(int)) != hipSuccess ||
      hipMalloc((void**)&c->d_counters, kCounterSlots * kCounterStride * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc((void**)&c->d_photons, sizeof(double)) != hipSuccess ||
      hipHostMalloc((void**)&c->h_status, 4 * sizeof(int)) != hipSuccess ||
      hipHostMalloc((void**)&c->h_counters, kCounterSlots * kCounterStride * sizeof(unsigned long long)) !=
          hipSuccess) {
    emma_rt_destroy(c);
    return EMMA_RT_ERR_HIP;
  }
  if (launch_fill((unsigned int*)c->d_status, 4, 0u, c->stream) != hipSuccess ||
      launch_fill((unsigned int*)c->d_counters, 2 * kCounterSlots * kCounterStride, 0u, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess) {
    emma_rt_destroy(c);
    return EMMA_RT_ERR_HIP;
  }
  st = scratch_of_current_device(c->device, &c->scratch);
  if (st) {
    emma_rt_destroy(c);
    return st;
  }
  if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || c->ncu < 1)
    c->ncu = 1;
  memset(&c->sched_info, 0, sizeof(c->sched_info));
  *out = c;
  return EMMA_RT_OK;
}

extern "C" int emma_rt_set_params(emma_rt_ctx* c, const emma_rt_params* p) {
  if (c == nullptr) return EMMA_RT_ERR_ARG;
  int st = check_params(p);
  if (st) return st;
  if (c->has_tree && p->ngrp != c->p.ngrp) return EMMA_RT_ERR_ARG;
  c->p = *p;
  drop_graphs(c);
  return EMMA_RT_OK;
}

extern "C" int emma_rt_destroy(emma_rt_ctx* c) {
  if (c == nullptr) return EMMA_RT_OK;
  (void)hipStreamSynchronize(c->stream);
  free_tree(c);
  stencil_free(c->d_stencil);
  for (auto e : c->ev) (void)hipEventDestroy(e);
  (void)hipFree(c->d_status);
  (void)hipFree(c->d_counters);
  (void)hipFree(c->d_photons);
  if (c->h_status) (void)hipHostFree(c->h_status);
  if (c->h_counters) (void)hipHostFree(c->h_counters);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return EMMA_RT_OK;
}

// ---------------------------------------------------------------- tree
template <typename T>
static hipError_t dalloc(T** p, size_t n) {
  return hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
}

// the halo variant of the tiled kernel a level gets unless the caller picks
// one (emma_rt_set_level_variant): kListed on complete levels; on sparse
// levels kLatency when the level has too few tiles to give every CU one,
// kPerDir otherwise
constexpr int kLatencyTiles = 256;
static int default_variant(const DevLevel& L) {
  if (!L.halo_per_dir) return kListed;
  const int tiles = (L.ncell + kTileCells - 1) / kTileCells;
  return tiles <= kLatencyTiles ? kLatency : kPerDir;
}

// log2 of the side when the level is a complete periodic cube of n^3 octs (n
// a power of two, 8 <= n <= 2^13) whose octs are in Morton order (x in the
// lowest bit) with the same-level neighbours nei_oct lists, else 0
static int morton_cube(int noct, const int* nei) {
  int m = 3;
  while (m <= 13 && (1LL << (3 * m)) < noct) m++;
  if (m > 13 || (1LL << (3 * m)) != noct) return 0;
  const unsigned n = 1u << m;
  auto spread = [](unsigned v) {  // bit i -> bit 3i
    unsigned long long r = 0;
    for (int i = 0; i < 13; i++) r |= (unsigned long long)((v >> i) & 1u) << (3 * i);
    return r;
  };
  auto key = [&](unsigned x, unsigned y, unsigned z) { return (long long)(spread(x) | spread(y) << 1 | spread(z) << 2); };
  for (int o = 0; o < noct; o++) {
    unsigned c[3] = {0, 0, 0};
    for (int i = 0; i < 3 * m; i++) c[i % 3] |= (unsigned)((o >> i) & 1) << (i / 3);
    for (int d = 0; d < 3; d++)
      for (int s = 0; s < 2; s++) {
        unsigned q[3] = {c[0], c[1], c[2]};
        q[d] = (q[d] + (s ? 1u : n - 1u)) & (n - 1u);
        if (nei[o * 6 + 2 * d + s] != key(q[0], q[1], q[2])) return 0;
      }
  }
  return m;
}

extern "C" int emma_rt_tree_create(emma_rt_ctx* c, int lmin, int lmax, const int* noct, const int* nei_oct,
                                   const int* child, int precision) {
  if (c == nullptr || noct == nullptr || nei_oct == nullptr || child == nullptr) return EMMA_RT_ERR_ARG;
  if (lmin < 1 || lmax < lmin || lmax > 30 || lmax - lmin + 1 > kMaxLevels) return EMMA_RT_ERR_ARG;
  if (precision != EMMA_RT_FP64 && precision != EMMA_RT_MIXED) return EMMA_RT_ERR_ARG;
  free_tree(c);
  const int nl = lmax - lmin + 1;
  const int G = c->p.ngrp;
  c->precision = precision;
  c->rbytes = precision == EMMA_RT_FP64 ? 8 : 4;
  c->lmin = lmin;
  c->lmax = lmax;
  c->lev.resize(nl);
  // host-side validation and reflux-slot assignment
  size_t on = 0, oc = 0;
  std::vector<size_t> off_n(nl), off_c(nl);
  for (int k = 0; k < nl; k++) {
    if (noct[k] < 0 || (k == 0 && noct[k] == 0)) return EMMA_RT_ERR_ARG;
    if ((long long)noct[k] * 8 > max_level_cells(G)) return EMMA_RT_ERR_ARG;
    off_n[k] = on;
    off_c[k] = oc;
    on += (size_t)noct[k] * 6;
    oc += (size_t)noct[k] * 8;
    DevLevel& L = c->lev[k];
    L.level = lmin + k;
    L.noct = noct[k];
    L.ncell = noct[k] * 8;
    L.hchild.assign(child + off_c[k], child + off_c[k] + L.ncell);
    L.hnei.assign(nei_oct + off_n[k], nei_oct + off_n[k] + (size_t)L.noct * 6);
  }
  for (int k = 0; k < nl; k++) {
    DevLevel& L = c->lev[k];
    const int nnext = (k + 1 < nl) ? c->lev[k + 1].noct : 0;
    for (int i = 0; i < L.ncell; i++) {
      int ch = L.hchild[i];
      if (ch < -1 || ch >= nnext) return EMMA_RT_ERR_ARG;
    }
    const int* nei = nei_oct + off_n[k];
    const int ncoarse = k > 0 ? c->lev[k - 1].ncell : 0;
    std::vector<unsigned char> rmask(ncoarse, 0);  // faces of the coarse leaves with an oct of level k behind
    for (int i = 0; i < L.noct * 6; i++) {
      int no = nei[i];
      if (no >= L.noct) return EMMA_RT_ERR_NEIGHBOR;
      if (no == EMMA_RT_NEI_TRANSMISSIVE || no == EMMA_RT_NEI_REFLECTIVE) continue;  // domain face
      if (no < 0) {
        if (k == 0) return EMMA_RT_ERR_NEIGHBOR;  // the coarsest level is complete
        int j = -1 - no;
        if (j >= ncoarse) return EMMA_RT_ERR_NEIGHBOR;
        if (c->lev[k - 1].hchild[j] >= 0) return EMMA_RT_ERR_NEIGHBOR;  // injected cells are leaves
        const int fc = (i % 6) ^ 1;  // the coarse leaf's face towards the oct
        if (rmask[j] & (1 << fc)) return EMMA_RT_ERR_NEIGHBOR;  // two octs behind one face
        rmask[j] |= (unsigned char)(1 << fc);
      }
    }
    if (k > 0) {
      // reflux pairs of level k-1, leaf by leaf, faces in order; child holds
      // -2 - (first pair << 3 | pairs) for those leaves
      DevLevel& C = c->lev[k - 1];
      std::vector<int> first(ncoarse, 0);
      long long np = 0;
      for (int j = 0; j < ncoarse; j++) {
        if (rmask[j] == 0) continue;
        first[j] = (int)np;
        const int m = __builtin_popcount(rmask[j]);
        C.hchild[j] = -2 - (int)((np << 3) | m);
        np += m;
        if (np >= (1LL << 28)) return EMMA_RT_ERR_ARG;
      }
      C.npair = (int)np;
      C.nslot = C.npair * (c->rbytes == 4 && (G == 1 || G == 3) ? kRefluxTiled : kRefluxCell);
      L.hrpair.assign((size_t)L.noct * 6, -1);
      for (int i = 0; i < L.noct * 6; i++) {
        if (nei[i] >= 0 || nei[i] <= EMMA_RT_NEI_REFLECTIVE) continue;
        const int j = -1 - nei[i], fc = (i % 6) ^ 1;
        L.hrpair[i] = first[j] + __builtin_popcount(rmask[j] & ((1u << fc) - 1u));
      }
    }
    // faces leaving a cell's oct whose neighbour lies outside the cell's tile,
    // counted per wave of 64 cells: complete levels have 64, sparse ones up
    // to 192 (every cell's three leaving faces)
    int maxh = 0;
    for (int w0 = 0; w0 < L.ncell; w0 += 64) {
      int cnt = 0;
      for (int cell = w0; cell < std::min(L.ncell, w0 + 64); cell++) {
        const int oct = cell >> 3, ic = cell & 7, oct0 = (cell / kTileCells) * kTileOcts;
        for (int d = 0; d < 3; d++) {
          const int no = nei[oct * 6 + 2 * d + ((ic >> d) & 1)];
          if (!(no >= oct0 && no < oct0 + kTileOcts)) cnt++;
        }
      }
      maxh = std::max(maxh, cnt);
    }
    L.halo_per_dir = maxh > 64 ? 1 : 0;
    L.variant = default_variant(L);
    L.morton_m = (k == 0) ? morton_cube(L.noct, nei) : 0;
    L.raster = 0;  // Morton order unless the caller (or the tuner) picks raster
  }
  // device allocation and upload
  for (int k = 0; k < nl; k++) {
    DevLevel& L = c->lev[k];
    const size_t gn = (size_t)G * L.ncell;
    for (int b = 0; b < 2; b++) {
      if (c->rbytes == 8) {
        HIPCHK(dalloc((double**)&L.e[b], gn));
        HIPCHK(dalloc((double**)&L.fx[b], gn));
        HIPCHK(dalloc((double**)&L.fy[b], gn));
        HIPCHK(dalloc((double**)&L.fz[b], gn));
      } else {
        HIPCHK(dalloc((float**)&L.e[b], gn));
        HIPCHK(dalloc((float**)&L.fx[b], gn));
        HIPCHK(dalloc((float**)&L.fy[b], gn));
        HIPCHK(dalloc((float**)&L.fz[b], gn));
      }
      HIPCHK(dalloc(&L.nhplus[b], L.ncell));
      HIPCHK(dalloc(&L.eint[b], L.ncell));
      HIPCHK(dalloc(&L.temp[b], L.ncell));
    }
    if (c->rbytes == 8)
      HIPCHK(dalloc((double**)&L.src, gn));
    else
      HIPCHK(dalloc((float**)&L.src, gn));
    HIPCHK(dalloc(&L.nh, L.ncell));
    HIPCHK(dalloc(&L.child, L.ncell));
    HIPCHK(dalloc(&L.nei, (size_t)L.noct * 6));
    const size_t accn = (size_t)L.nslot * 4 * (G == 1 ? 1 : (G == 3 ? 3 : kMaxGrp));
    HIPCHK(hipMalloc(&L.racc, std::max<size_t>(accn, 1) * c->rbytes));
    HIPCHK(hipMemset(L.racc, 0, std::max<size_t>(accn, 1) * c->rbytes));
    HIPCHK(hipMemcpy(L.child, L.hchild.data(), L.ncell * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(L.nei, nei_oct + off_n[k], (size_t)L.noct * 6 * sizeof(int), hipMemcpyHostToDevice));
    if (k > 0) {
      HIPCHK(dalloc(&L.rpair, (size_t)L.noct * 6));
      HIPCHK(hipMemcpy(L.rpair, L.hrpair.data(), (size_t)L.noct * 6 * sizeof(int), hipMemcpyHostToDevice));
    }
  }
  if (c->rbytes == 4 && (G == 1 || G == 3)) {  // the tiled kernel's straggler list
    int cap = 0;
    for (auto& L : c->lev) cap = std::max(cap, L.ncell);
    HIPCHK(dalloc(&c->d_defer, (size_t)cap));
    c->defer_cap = cap;
  }
  c->has_tree = true;
  if (c->xion) {
    c->xion = 0;
    return emma_rt_enable_xion_times(c, 1);
  }
  return EMMA_RT_OK;
}

static DevLevel* level_of(emma_rt_ctx* c, int level) {
  if (c == nullptr || !c->has_tree || level < c->lmin || level > c->lmax) return nullptr;
  return &c->lev[level - c->lmin];
}

// floor (photon densities, FP32 storage): a value at or above the group's
// floor must not round below it (ncell values per group)
template <typename R>
static int upload_r(void* dst, const double* src, size_t n, hipStream_t s, const emma_rt_params* floor = nullptr,
                    size_t ncell = 0) {
  if (sizeof(R) == sizeof(double)) {
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    return EMMA_RT_OK;
  }
  std::vector<R> tmp(n);
  for (size_t i = 0; i < n; i++) tmp[i] = (R)src[i];
  if (floor != nullptr) {
    for (size_t i = 0; i < n; i++) {
      const int g = (int)(i / ncell);
      const double f = emin_floor(*floor, g);
      if (src[i] >= f && (double)tmp[i] < f) tmp[i] = (R)emin_floor_up(f);
    }
  }
  HIPCHK(hipMemcpyAsync(dst, tmp.data(), n * sizeof(R), hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  return EMMA_RT_OK;
}

template <typename R>
static int download_r(double* dst, const void* src, size_t n, hipStream_t s) {
  if (sizeof(R) == sizeof(double)) {
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return EMMA_RT_OK;
  }
  std::vector<R> tmp(n);
  HIPCHK(hipMemcpyAsync(tmp.data(), src, n * sizeof(R), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (size_t i = 0; i < n; i++) dst[i] = (double)tmp[i];
  return EMMA_RT_OK;
}

static int upload_rf(emma_rt_ctx* c, void* dst, const double* src, size_t n, bool photon_floor = false,
                     size_t ncell = 0) {
  const emma_rt_params* fl = photon_floor ? &c->p : nullptr;
  return c->rbytes == 8 ? upload_r<double>(dst, src, n, c->stream) : upload_r<float>(dst, src, n, c->stream, fl, ncell);
}

static int download_rf(emma_rt_ctx* c, double* dst, const void* src, size_t n) {
  return c->rbytes == 8 ? download_r<double>(dst, src, n, c->stream) : download_r<float>(dst, src, n, c->stream);
}

extern "C" int emma_rt_set_level(emma_rt_ctx* c, int level, const double* e, const double* fx, const double* fy,
                                 const double* fz, const double* src, const double* nhplus, const double* eint,
                                 const double* nh, const double* temp) {
  DevLevel* L = level_of(c, level);
  if (L == nullptr) return EMMA_RT_ERR_STATE;
  if (!e || !fx || !fy || !fz || !src || !nhplus || !eint || !nh || !temp) return EMMA_RT_ERR_ARG;
  const size_t gn = (size_t)c->p.ngrp * L->ncell;
  const int b = L->cur;
  int st;
  if ((st = upload_rf(c, L->e[b], e, gn, true, L->ncell)) || (st = upload_rf(c, L->fx[b], fx, gn)) ||
      (st = upload_rf(c, L->fy[b], fy, gn)) || (st = upload_rf(c, L->fz[b], fz, gn)) ||
      (st = upload_rf(c, L->src, src, gn)))
    return st;
  HIPCHK(hipMemcpy(L->nhplus[b], nhplus, L->ncell * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(L->eint[b], eint, L->ncell * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(L->temp[b], temp, L->ncell * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(L->nh, nh, L->ncell * sizeof(double), hipMemcpyHostToDevice));
  return EMMA_RT_OK;
}

extern "C" int emma_rt_get_level(emma_rt_ctx* c, int level, double* e, double* fx, double* fy, double* fz,
                                 double* src, double* nhplus, double* eint, double* nh, double* temp) {
  DevLevel* L = level_of(c, level);
  if (L == nullptr) return EMMA_RT_ERR_STATE;
  HIPCHK(hipStreamSynchronize(c->stream));
  const size_t gn = (size_t)c->p.ngrp * L->ncell;
  const int b = L->cur;
  int st;
  if (e && (st = download_rf(c, e, L->e[b], gn))) return st;
  if (fx && (st = download_rf(c, fx, L->fx[b], gn))) return st;
  if (fy && (st = download_rf(c, fy, L->fy[b], gn))) return st;
  if (fz && (st = download_rf(c, fz, L->fz[b], gn))) return st;
  if (src && (st = download_rf(c, src, L->src, gn))) return st;
  if (nhplus) HIPCHK(hipMemcpy(nhplus, L->nhplus[b], L->ncell * sizeof(double), hipMemcpyDeviceToHost));
  if (eint) HIPCHK(hipMemcpy(eint, L->eint[b], L->ncell * sizeof(double), hipMemcpyDeviceToHost));
  if (nh) HIPCHK(hipMemcpy(nh, L->nh, L->ncell * sizeof(double), hipMemcpyDeviceToHost));
  if (temp) HIPCHK(hipMemcpy(temp, L->temp[b], L->ncell * sizeof(double), hipMemcpyDeviceToHost));
  return EMMA_RT_OK;
}

extern "C" int emma_rt_enable_xion_times(emma_rt_ctx* c, int enable) {
  if (c == nullptr) return EMMA_RT_ERR_ARG;
  if (!c->has_tree) {
    c->xion = enable ? 1 : 0;
    return EMMA_RT_OK;
  }
  if (enable && !c->xion) {
    for (auto& L : c->lev) {
      HIPCHK(dalloc(&L.t_first, L.ncell));
      HIPCHK(dalloc(&L.t_last, L.ncell));
      std::vector<double> m1(L.ncell, -1.);
      HIPCHK(hipMemcpy(L.t_first, m1.data(), L.ncell * sizeof(double), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(L.t_last, m1.data(), L.ncell * sizeof(double), hipMemcpyHostToDevice));
    }
  }
  if (!enable && c->xion) {
    for (auto& L : c->lev) {
      (void)hipFree(L.t_first);
      (void)hipFree(L.t_last);
      L.t_first = L.t_last = nullptr;
    }
  }
  c->xion = enable ? 1 : 0;
  drop_graphs(c);
  return EMMA_RT_OK;
}

extern "C" int emma_rt_get_xion_times(emma_rt_ctx* c, int level, double* t_first, double* t_last) {
  DevLevel* L = level_of(c, level);
  if (L == nullptr || !c->xion) return EMMA_RT_ERR_STATE;
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(t_first, L->t_first, L->ncell * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(t_last, L->t_last, L->ncell * sizeof(double), hipMemcpyDeviceToHost));
  return EMMA_RT_OK;
}

extern "C" int emma_rt_set_tphy(emma_rt_ctx* c, double tphy) {
  if (c == nullptr) return EMMA_RT_ERR_ARG;
  c->tphy = tphy;
  drop_graphs(c);
  return EMMA_RT_OK;
}

// ---------------------------------------------------------------- level step
template <typename R>
static LevelArgs<R> level_args(emma_rt_ctx* c, int k, double dtnew, double aexp, int chemonly, bool counters) {
  DevLevel& L = c->lev[k];
  const int a = L.cur, b = 1 - L.cur;
  LevelArgs<R> A;
  memset(&A, 0, sizeof(A));
  A.ncell = L.ncell;
  A.ng = c->p.ngrp;
  A.level = L.level;
  A.chemonly = chemonly;
  if (c->mode == EMMA_RT_COARSERAD) {
    A.transport = (k == 0) && !chemonly;
    A.keep_split_flux = 1;
    A.update_split = A.transport;
  } else {
    A.transport = 1;
    A.keep_split_flux = 0;
    A.update_split = 0;
  }
  A.variant = L.variant;
  A.raster_m = (L.raster && L.morton_m > 0) ? L.morton_m : 0;
  if (c->d_defer != nullptr && L.chem_budget > 0 && c->defer_seq < c->defer_slots) {
    A.chem_budget = L.chem_budget;
    A.chem_keep = L.chem_keep;
    A.defer_cap = c->defer_cap;
    A.defer_list = c->d_defer;
    A.defer_ctl = c->d_defer_ctl + c->defer_seq++;
  }
  A.e = (const R*)L.e[a];
  A.fx = (const R*)L.fx[a];
  A.fy = (const R*)L.fy[a];
  A.fz = (const R*)L.fz[a];
  A.nhplus = L.nhplus[a];
  A.eint = L.eint[a];
  A.temp = L.temp[a];
  A.src = (R*)L.src;
  A.nh = L.nh;
  A.eo = (R*)L.e[b];
  A.fxo = (R*)L.fx[b];
  A.fyo = (R*)L.fy[b];
  A.fzo = (R*)L.fz[b];
  A.nhpluso = L.nhplus[b];
  A.einto = L.eint[b];
  A.tempo = L.temp[b];
  A.child = L.child;
  A.nei = L.nei;
  A.rpair = L.rpair;
  A.racc = (R*)L.racc;
  A.nslot = L.nslot;
  if (k > 0) {
    DevLevel& C = c->lev[k - 1];
    A.ncoarse = C.ncell;
    A.ce = (const R*)C.e[C.cur];
    A.cfx = (const R*)C.fx[C.cur];
    A.cfy = (const R*)C.fy[C.cur];
    A.cfz = (const R*)C.fz[C.cur];
    A.cracc = (R*)C.racc;
    A.cnslot = C.nslot;
  }
  if (k + 1 < (int)c->lev.size()) {
    DevLevel& F = c->lev[k + 1];
    A.nfine = F.ncell;
    A.fe = (const R*)F.e[F.cur];
    A.ffx = (const R*)F.fx[F.cur];
    A.ffy = (const R*)F.fy[F.cur];
    A.ffz = (const R*)F.fz[F.cur];
    A.fsrc = (const R*)F.src;
    A.fnhplus = F.nhplus[F.cur];
    A.feint = F.eint[F.cur];
    A.ftemp = F.temp[F.cur];
    A.fnh = F.nh;
    A.ft_first = F.t_first;
    A.ft_last = F.t_last;
  }
  if (c->xion) {
    A.t_first = L.t_first;
    A.t_last = L.t_last;
  }
  A.tphy = c->tphy;
  const double cl = cloc_of(c->p, aexp);
  const double dx = pow(0.5, L.level);
  A.c = (R)cl;
  A.hc = (R)(0.5 * cl);
  A.dtsurdx = (R)(dtnew / dx);
  A.status = c->d_status;
  A.counters = counters ? c->d_counters : nullptr;
  return A;
}

static hipError_t launch_level(emma_rt_ctx* c, int k, double dtnew, double aexp, int chemonly, bool counters) {
  DevLevel& L = c->lev[k];
  const ChemConst K = make_chem(c->p, dtnew, aexp);
  hipError_t e;
  if (c->rbytes == 8)
    e = launch_level_step<double>(level_args<double>(c, k, dtnew, aexp, chemonly, counters), K, c->stream);
  else
    e = launch_level_step<float>(level_args<float>(c, k, dtnew, aexp, chemonly, counters), K, c->stream);
  L.cur = 1 - L.cur;
  return e;
}

static hipEvent_t next_event(emma_rt_ctx* c) {
  if (c->ev_used == c->ev.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    c->ev.push_back(e);
  }
  return c->ev[c->ev_used++];
}

template <typename R>
static void record_step(emma_rt_ctx* c, int k, double dtnew, double aexp, int chemonly) {
  StepDesc<R> D;
  memset(&D, 0, sizeof(D));
  D.a = level_args<R>(c, k, dtnew, aexp, chemonly, c->counters != 0);
  // no straggler pass inside the persistent kernel: every leaf finishes its
  // chemistry in its tile
  D.a.chem_budget = 0;
  D.a.defer_cap = 0;
  D.a.defer_list = nullptr;
  D.a.defer_ctl = nullptr;
  D.K = make_chem(c->p, dtnew, aexp);
  // the kernel holds the listed and per-direction halo variants; the latency
  // variant (more registers than the kernel's occupancy allows) runs as
  // per-direction
  D.variant = c->lev[k].variant == kListed ? kListed : kPerDir;
  if (!c->pending_fill[k].empty()) {
    memcpy(&D.f, c->pending_fill[k].data(), sizeof(D.f));
    D.fill = 1;
    c->pending_fill[k].clear();
  }
  SchedOp op;
  op.kind = 0;
  op.k = k;
  op.blob.resize(sizeof(D));
  memcpy(op.blob.data(), &D, sizeof(D));
  c->rec->push_back(std::move(op));
  c->lev[k].cur = 1 - c->lev[k].cur;
}

static int timed_launch(emma_rt_ctx* c, int k, double dtnew, double aexp, int chemonly) {
  if (c->rec) {
    if (c->rbytes == 8)
      record_step<double>(c, k, dtnew, aexp, chemonly);
    else
      record_step<float>(c, k, dtnew, aexp, chemonly);
    return EMMA_RT_OK;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (c->timing) {
    e0 = next_event(c);
    e1 = next_event(c);
    if (!e0 || !e1) return EMMA_RT_ERR_HIP;
    c->ev_level.push_back(k);
    HIPCHK(hipEventRecord(e0, c->stream));
  }
  HIPCHK(launch_level(c, k, dtnew, aexp, chemonly, c->counters != 0));
  if (c->timing) HIPCHK(hipEventRecord(e1, c->stream));
  c->stats.level_steps++;
  return EMMA_RT_OK;
}

// list-length words for `n` level-step launches in one call (allocated
// outside any stream capture: the calls below run before BeginCapture)
static int ensure_defer_slots(emma_rt_ctx* c, long long n) {
  if (c->d_defer == nullptr || n <= c->defer_slots) return EMMA_RT_OK;
  if (n > (1LL << 24)) return EMMA_RT_ERR_ARG;
  HIPCHK(hipStreamSynchronize(c->stream));
  // captured coarse steps hold the old list-length words
  drop_graphs(c);
  (void)hipFree(c->d_defer_ctl);
  c->d_defer_ctl = nullptr;
  c->defer_slots = 0;
  const int m = (int)((n + 255) & ~255LL);
  HIPCHK(dalloc(&c->d_defer_ctl, (size_t)m));
  c->defer_slots = m;
  return EMMA_RT_OK;
}

// level-step launches of `ncoarse` coarse steps
static long long steps_per_call(emma_rt_ctx* c, int ncoarse) {
  long long steps = 0;
  for (int k = 0; k < (int)c->lev.size(); k++)
    if (c->lev[k].noct > 0) steps += (c->mode == EMMA_RT_COARSERAD) ? (long long)ncoarse : ((long long)ncoarse << k);
  return steps;
}

// the device words a call starts from: status, straggler list lengths,
// counters (fill kernels, so that a coarse step can capture them at the
// head of its graph)
static int clear_call_words(emma_rt_ctx* c) {
  HIPCHK(launch_fill((unsigned int*)c->d_status, 4, 0u, c->stream));
  if (c->defer_slots > 0) HIPCHK(launch_fill((unsigned int*)c->d_defer_ctl, c->defer_slots, 0u, c->stream));
  if (c->counters) HIPCHK(launch_fill((unsigned int*)c->d_counters, 2 * kCounterSlots * kCounterStride, 0u, c->stream));
  return EMMA_RT_OK;
}

// clear_device = false: the caller runs clear_call_words inside its graph
static int begin_call(emma_rt_ctx* c, bool clear_device = true) {
  if (clear_device) {
    int st = clear_call_words(c);
    if (st) return st;
  }
  c->defer_seq = 0;
  memset(&c->stats, 0, sizeof(c->stats));
  c->ev_used = 0;
  c->ev_level.clear();
  return EMMA_RT_OK;
}

static int end_call(emma_rt_ctx* c) {
  // the status word (and the counters) copied behind the call's work, one wait
  const bool counted = c->counters != 0;
  HIPCHK(hipMemcpyAsync(c->h_status, c->d_status, 4 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  if (counted)
    HIPCHK(hipMemcpyAsync(c->h_counters, c->d_counters, kCounterSlots * kCounterStride * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const int* status = c->h_status;
  const unsigned long long* slots = c->h_counters;
  unsigned long long cnt[3] = {0ull, 0ull, 0ull};
  if (counted) {
    for (int i = 0; i < kCounterSlots; i++) {
      cnt[0] += slots[(size_t)i * kCounterStride + 0];
      cnt[1] += slots[(size_t)i * kCounterStride + 1];
      cnt[2] += slots[(size_t)i * kCounterStride + 2];
    }
  }
  c->stats.status = status[0];
  c->stats.bad_level = status[1];
  c->stats.bad_cell = status[2];
  c->stats.chem_substeps = counted ? (long long)cnt[0] : -1;
  c->stats.cell_updates = counted ? (long long)cnt[1] : -1;
  c->stats.deferred_cells = counted ? (long long)cnt[2] : -1;
  double ms = 0.;
  for (int k = 0; k < kMaxLevels; k++) {
    c->level_ms[k] = 0.;
    c->level_launches[k] = 0;
  }
  for (size_t i = 0; i + 1 < c->ev_used; i += 2) {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, c->ev[i], c->ev[i + 1]));
    ms += t;
    const int k = c->ev_level[i / 2];
    if (k >= 0) {
      c->level_ms[k] += t;
      c->level_launches[k]++;
    }
  }
  c->stats.kernel_ms = ms;
  c->stats.timed_spans = (int)(c->ev_used / 2);
  return status[0];
}

template <typename R>
static TreeView<R> tree_view(emma_rt_ctx* c) {
  TreeView<R> t;
  memset(&t, 0, sizeof(t));
  t.nlev = (int)c->lev.size();
  t.ng = c->p.ngrp;
  t.lcoarse = c->lmin;
  for (int k = 0; k < t.nlev; k++) {
    DevLevel& L = c->lev[k];
    t.ncell[k] = L.ncell;
    t.child[k] = L.child;
    t.e[k] = (R*)L.e[L.cur];
    t.fx[k] = (R*)L.fx[L.cur];
    t.fy[k] = (R*)L.fy[L.cur];
    t.fz[k] = (R*)L.fz[L.cur];
  }
  return t;
}

// FillRad's COARSERAD block (src_utils.c:717-733): split cells of the coarse
// level gather the volume-weighted E, F of their leaves
template <typename R>
static void record_transfer(emma_rt_ctx* c, int kind) {
  const TreeView<R> t = tree_view<R>(c);
  SchedOp op;
  op.kind = kind;
  op.blob.resize(sizeof(t));
  memcpy(op.blob.data(), &t, sizeof(t));
  c->rec->push_back(std::move(op));
}

static hipError_t launch_put(emma_rt_ctx* c) {
  if (c->rec) {
    if (c->rbytes == 8)
      record_transfer<double>(c, 1);
    else
      record_transfer<float>(c, 1);
    return hipSuccess;
  }
  if (c->rbytes == 8) return launch_put_e2coarse<double>(tree_view<double>(c), c->stream);
  return launch_put_e2coarse<float>(tree_view<float>(c), c->stream);
}

// RadSolver's COARSERAD push of the coarse values to the leaves (rad_utils.c:2009-2013)
static hipError_t launch_distrib(emma_rt_ctx* c) {
  if (c->rec) {
    if (c->rbytes == 8)
      record_transfer<double>(c, 2);
    else
      record_transfer<float>(c, 2);
    return hipSuccess;
  }
  if (c->rbytes == 8) return launch_distrib_e<double>(tree_view<double>(c), c->stream);
  return launch_distrib_e<float>(tree_view<float>(c), c->stream);
}

// ---------------------------------------------------------------- FillRad / hydro coupling
static bool hydro_ready(emma_rt_ctx* c) {
  for (auto& L : c->lev)
    if (L.noct > 0 && L.hyd[0] == nullptr) return false;
  return true;
}

static bool opos_ready(emma_rt_ctx* c) {
  for (auto& L : c->lev)
    if (L.noct > 0 && L.opos == nullptr) return false;
  return true;
}

// FillRad's per-cell block on level index k (src_utils.c:738-838); a no-op
// unless the hydro coupling or the test source is on
template <typename R>
static hipError_t launch_fill_cells(emma_rt_ctx* c, int k, double aexp) {
  DevLevel& L = c->lev[k];
  if (!(c->hydro || c->star || c->stars) || L.noct == 0) return hipSuccess;
  FillArgs<R> f;
  memset(&f, 0, sizeof(f));
  f.ncell = L.ncell;
  f.ng = c->p.ngrp;
  f.hydro = c->hydro;
  f.star = c->star;
  f.child = L.child;
  f.nh = L.nh;
  f.eint = L.eint[L.cur];
  f.nhplus = L.nhplus[L.cur];
  f.temp = L.temp[L.cur];
  f.src = (R*)L.src;
  f.d = L.hyd[0];
  f.p = L.hyd[4];
  f.dX = L.hyd[5];
  f.opos = L.opos;
  if (c->star) {
    // stromgren_source (src_utils.c:339-412): the uniform parts on the host,
    // in the reference's operation order (glibc pow and fmod)
    const double MYR = 3.1536e13, SOLAR_MASS = 1.989e30;  // param.h:31, 37
    const int nsub = 1;
    f.X0 = 1. / (double)(1ULL << c->lmin);
    f.dxcur = pow(0.5, L.level);
    f.xs = c->star_x;
    f.ys = c->star_y;
    f.zs = c->star_z;
    const double tcur = c->tphy;
    const double tcur_in_yrs = tcur * c->p.unit_t / MYR * 1e6;
    f.star_g0 = f.star_g1 = 0;
    if (fmod(tcur_in_yrs, c->star_lifetime) < c->star_lifetime / nsub) {
      const int ns = c->p.ngrp / c->p.ngrp_time;  // NGRP_SPACE
      for (int it = 0; it < c->p.ngrp_time; it++)
        if (tcur_in_yrs >= c->p.time_bound[it] && tcur_in_yrs < c->p.time_bound[it + 1]) {
          f.star_g0 = it * ns;
          f.star_g1 = it * ns + ns;
        }
    }
    const double srcint = c->star_mass * SOLAR_MASS * c->star_srcint * (tcur > 0 ? 1. : 0) * nsub / 8;
    f.star_src = srcint / pow(f.X0 * c->p.unit_l, 3) * c->p.unit_t / c->p.unit_N * pow(aexp, 2);
  }
  if (c->stars) {
    // stars_sources (src_utils.c:209-335): the uniform factors on the host;
    // a level without particles gets its sources cleared
    f.stars = 1;
    f.stars_decrease = c->stars_decrease;
    f.nspace = c->p.ngrp / c->p.ngrp_time;
    f.star_first = L.star_first;
    f.star_mass = L.star_mass;
    f.star_age = L.star_age;
    f.star_rs = L.star_rs;
    f.star_is = L.star_is;
    f.stars_unit_d = c->p.unit_d;
    f.stars_srcint = c->stars_srcint;
    f.stars_dx3 = pow(pow(0.5, L.level), 3);
    f.stars_unit_t = c->p.unit_t;
    f.stars_unit_N = c->p.unit_N;
    f.stars_pa2 = pow(aexp, 2);
    f.stars_tphy = c->tphy;
    f.stars_tlife = c->stars_tlife;
  }
  f.pa3 = pow(aexp, 3);
  f.pa5 = pow(aexp, 5);
  f.pstar = c->p.unit_n * c->p.unit_d * pow(c->p.unit_v, 2);
  f.unit_N = c->p.unit_N;
  if (c->rec) {  // fused into the level's next step
    c->pending_fill[k].resize(sizeof(f));
    memcpy(c->pending_fill[k].data(), &f, sizeof(f));
    return hipSuccess;
  }
  return launch_fill_level<R>(f, c->stream);
}

static hipError_t fill_cells(emma_rt_ctx* c, int k, double aexp) {
  if (c->rbytes == 8) return launch_fill_cells<double>(c, k, aexp);
  return launch_fill_cells<float>(c, k, aexp);
}

// RadSolver's injection into the hydro state after the level's final update
static hipError_t to_hydro(emma_rt_ctx* c, int k) {
  DevLevel& L = c->lev[k];
  if (!c->hydro || L.noct == 0) return hipSuccess;
  HydroArgs h;
  h.ncell = L.ncell;
  h.eint = L.eint[L.cur];
  h.nhplus = L.nhplus[L.cur];
  h.nh = L.nh;
  h.d = L.hyd[0];
  h.u = L.hyd[1];
  h.v = L.hyd[2];
  h.w = L.hyd[3];
  h.p = L.hyd[4];
  h.dX = L.hyd[5];
  h.E = L.hyd[6];
  if (c->rec) {  // fused into the level's last recorded step
    for (auto it = c->rec->rbegin(); it != c->rec->rend(); ++it) {
      if (it->kind != 0 || it->k != k) continue;
      if (c->rbytes == 8) {
        StepDesc<double>* D = reinterpret_cast<StepDesc<double>*>(it->blob.data());
        D->h = h;
        D->hydro = 1;
      } else {
        StepDesc<float>* D = reinterpret_cast<StepDesc<float>*>(it->blob.data());
        D->h = h;
        D->hydro = 1;
      }
      return hipSuccess;
    }
    return hipErrorInvalidValue;
  }
  return launch_rad_to_hydro(h, c->stream);
}

extern "C" int emma_rt_set_hydro(emma_rt_ctx* c, int level, const double* d, const double* u, const double* v,
                                 const double* w, const double* p, const double* dX) {
  DevLevel* L = level_of(c, level);
  if (L == nullptr) return EMMA_RT_ERR_STATE;
  if (L->ncell == 0) return EMMA_RT_OK;
  if (!d || !u || !v || !w || !p || !dX) return EMMA_RT_ERR_ARG;
  const size_t nb = (size_t)L->ncell * sizeof(double);
  for (int k = 0; k < 7; k++)
    if (L->hyd[k] == nullptr) HIPCHK(hipMalloc((void**)&L->hyd[k], nb));
  const double* in[6] = {d, u, v, w, p, dX};
  for (int k = 0; k < 6; k++) HIPCHK(hipMemcpy(L->hyd[k], in[k], nb, hipMemcpyHostToDevice));
  HIPCHK(launch_fill((unsigned int*)L->hyd[6], 2 * L->ncell, 0u, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return EMMA_RT_OK;
}

extern "C" int emma_rt_get_hydro(emma_rt_ctx* c, int level, double* p, double* dX, double* E) {
  DevLevel* L = level_of(c, level);
  if (L == nullptr) return EMMA_RT_ERR_STATE;
  if (L->ncell == 0) return EMMA_RT_OK;
  if (L->hyd[0] == nullptr) return EMMA_RT_ERR_STATE;
  const size_t nb = (size_t)L->ncell * sizeof(double);
  HIPCHK(hipStreamSynchronize(c->stream));
  if (p) HIPCHK(hipMemcpy(p, L->hyd[4], nb, hipMemcpyDeviceToHost));
  if (dX) HIPCHK(hipMemcpy(dX, L->hyd[5], nb, hipMemcpyDeviceToHost));
  if (E) HIPCHK(hipMemcpy(E, L->hyd[6], nb, hipMemcpyDeviceToHost));
  return EMMA_RT_OK;
}

extern "C" int emma_rt_set_oct_pos(emma_rt_ctx* c, int level, const double* xyz) {
  DevLevel* L = level_of(c, level);
  if (L == nullptr) return EMMA_RT_ERR_STATE;
  if (L->noct == 0) return EMMA_RT_OK;
  if (!xyz) return EMMA_RT_ERR_ARG;
  const size_t nb = (size_t)L->noct * 3 * sizeof(double);
  if (L->opos == nullptr) HIPCHK(hipMalloc((void**)&L->opos, nb));
  HIPCHK(hipMemcpy(L->opos, xyz, nb, hipMemcpyHostToDevice));
  return EMMA_RT_OK;
}

extern "C" int emma_rt_set_stromgren_source(emma_rt_ctx* c, int enable, double x, double y, double z, double mass,
                                            double lifetime, double srcint) {
  if (c == nullptr) return EMMA_RT_ERR_ARG;
  if (enable && !(lifetime > 0)) return EMMA_RT_ERR_ARG;
  c->star = enable ? 1 : 0;
  if (enable) c->stars = 0;  // one source kind per build (WRADTEST or STARS)
  c->star_x = x;
  c->star_y = y;
  c->star_z = z;
  c->star_mass = mass;
  c->star_lifetime = lifetime;
  c->star_srcint = srcint;
  drop_graphs(c);
  return EMMA_RT_OK;
}

extern "C" int emma_rt_set_stars(emma_rt_ctx* c, int level, int nstar, const int* cell, const double* mass,
                                 const double* age, const int* radiative_state, const int* is_star) {
  DevLevel* L = level_of(c, level);
  if (L == nullptr) return EMMA_RT_ERR_STATE;
  if (nstar < 0 || (nstar > 0 && (!cell || !mass || !age || !radiative_state || !is_star))) return EMMA_RT_ERR_ARG;
  const int ntime = c->p.ngrp_time;
  std::vector<int> first((size_t)L->ncell + 1, 0);
  for (int k = 0; k < nstar; k++) {
    const int j = cell[k];
    if (j < 0 || j >= L->ncell || (k > 0 && j < cell[k - 1])) return EMMA_RT_ERR_ARG;  // cells ascending
    if (L->hchild[j] >= 0) return EMMA_RT_ERR_ARG;  // particles sit in leaf cells
    if (radiative_state[k] < 0 || radiative_state[k] >= ntime) return EMMA_RT_ERR_ARG;
    first[j + 1]++;
  }
  for (int j = 0; j < L->ncell; j++) first[j + 1] += first[j];
  for (void* p : {(void*)L->star_first, (void*)L->star_mass, (void*)L->star_age, (void*)L->star_rs, (void*)L->star_is})
    (void)hipFree(p);
  L->star_first = nullptr;
  L->star_mass = L->star_age = nullptr;
  L->star_rs = L->star_is = nullptr;
  drop_graphs(c);
  if (nstar == 0) return EMMA_RT_OK;
  const size_t n = (size_t)nstar;
  HIPCHK(dalloc(&L->star_first, (size_t)L->ncell + 1));
  HIPCHK(dalloc(&L->star_mass, n));
  HIPCHK(dalloc(&L->star_age, n));
  HIPCHK(dalloc(&L->star_rs, n));
  HIPCHK(dalloc(&L->star_is, n));
  HIPCHK(hipMemcpy(L->star_first, first.data(), first.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(L->star_mass, mass, n * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(L->star_age, age, n * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(L->star_rs, radiative_state, n * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(L->star_is, is_star, n * sizeof(int), hipMemcpyHostToDevice));
  return EMMA_RT_OK;
}

extern "C" int emma_rt_set_star_sources(emma_rt_ctx* c, int enable, double srcint, double tlife,
                                        int decrease_after_tlife) {
  if (c == nullptr) return EMMA_RT_ERR_ARG;
  if (enable && !(tlife > 0)) return EMMA_RT_ERR_ARG;
  c->stars = enable ? 1 : 0;
  if (enable) c->star = 0;  // one source kind per build (WRADTEST or STARS, src_utils.c:458-500)
  c->stars_srcint = srcint;
  c->stars_tlife = tlife;
  c->stars_decrease = decrease_after_tlife ? 1 : 0;
  drop_graphs(c);
  return EMMA_RT_OK;
}

extern "C" int emma_rt_set_coupling(emma_rt_ctx* c, int hydro) {
  if (c == nullptr) return EMMA_RT_ERR_ARG;
  c->hydro = hydro ? 1 : 0;
  drop_graphs(c);
  return EMMA_RT_OK;
}

// every level set up for what the coupling and the source read
static int coupling_ready(emma_rt_ctx* c) {
  if (c->hydro && !hydro_ready(c)) return EMMA_RT_ERR_STATE;
  if (c->star && !opos_ready(c)) return EMMA_RT_ERR_STATE;
  return EMMA_RT_OK;
}

extern "C" int emma_rt_set_mode(emma_rt_ctx* c, int mode) {
  if (c == nullptr) return EMMA_RT_ERR_ARG;
  if (mode != EMMA_RT_RADSTEP && mode != EMMA_RT_COARSERAD) return EMMA_RT_ERR_ARG;
  c->mode = mode;
  drop_graphs(c);
  return EMMA_RT_OK;
}

extern "C" int emma_rt_fillrad(emma_rt_ctx* c, int level, double aexp) {
  DevLevel* L = level_of(c, level);
  if (L == nullptr) return EMMA_RT_ERR_STATE;
  int st = coupling_ready(c);
  if (st) return st;
  if (c->mode == EMMA_RT_COARSERAD && level == c->lmin && c->lev.size() > 1) HIPCHK(launch_put(c));
  HIPCHK(fill_cells(c, level - c->lmin, aexp));
  HIPCHK(hipStreamSynchronize(c->stream));
  return EMMA_RT_OK;
}

extern "C" int emma_rt_radsolver(emma_rt_ctx* c, int level, double dtnew, double aexp, int chemonly) {
  DevLevel* L = level_of(c, level);
  if (L == nullptr) return EMMA_RT_ERR_STATE;
  int st = coupling_ready(c);
  if (st) return st;
  st = ensure_defer_slots(c, 1);
  if (st) return st;
  st = begin_call(c);
  if (st) return st;
  st = timed_launch(c, level - c->lmin, dtnew, aexp, chemonly);
  if (st) return st;
  if (c->mode == EMMA_RT_COARSERAD && level == c->lmin && c->lev.size() > 1) HIPCHK(launch_distrib(c));
  HIPCHK(to_hydro(c, level - c->lmin));
  return end_call(c);
}

// COARSERAD recursion of Advance_level_RAD (advanceamr.c:1309-1537): every
// level takes one step of the coarse level's dt, finer levels first; the
// coarse level gathers its leaves before its step and pushes back after it.
static int recurse_coarserad(emma_rt_ctx* c, int k, double aexp) {
  const int nl = (int)c->lev.size();
  if (k + 1 < nl && c->lev[k + 1].noct > 0) {
    int st = recurse_coarserad(c, k + 1, aexp);
    if (st) return st;
  }
  if (k == 0 && nl > 1) HIPCHK(launch_put(c));
  HIPCHK(fill_cells(c, k, aexp));
  int st = timed_launch(c, k, dtrad_of(c->p, c->lmin, aexp), aexp, 0);
  if (st) return st;
  if (k == 0 && nl > 1) HIPCHK(launch_distrib(c));
  HIPCHK(to_hydro(c, k));
  return EMMA_RT_OK;
}

// RADSTEP recursion of Advance_level (advanceamr.c:243-1300): per step of
// level l, the finer level takes two steps first, then level l steps.
static int recurse(emma_rt_ctx* c, int k, double aexp) {
  if (c->mode == EMMA_RT_COARSERAD) return recurse_coarserad(c, k, aexp);
  const int nl = (int)c->lev.size();
  const DevLevel& L = c->lev[k];
  const double dt = dtrad_of(c->p, L.level, aexp);
  const int nsub = (k == 0) ? 1 : 2;
  for (int s = 0; s < nsub; s++) {
    if (k + 1 < nl && c->lev[k + 1].noct > 0) {
      int st = recurse(c, k + 1, aexp);
      if (st) return st;
    }
    HIPCHK(fill_cells(c, k, aexp));  // FillRad before RadSolver (advanceamr.c:1035)
    int st = timed_launch(c, k, dt, aexp, 0);
    if (st) return st;
    HIPCHK(to_hydro(c, k));
  }
  return EMMA_RT_OK;
}

// ---------------------------------------------------------------- persistent coarse step
// The coarse step is recorded (the recursion above with the launch helpers
// recording instead of launching), each run of consecutive level steps is
// planned (rt_sched.h) and its tables go to the device once; the captured
// graph then holds one persistent launch per run (plus the COARSERAD
// transfers between runs).
constexpr long long kSchedWaitTicks = 200000000LL;  // 2 s of s_memrealtime (100 MHz)

// The XCDs the device's work-groups run on (the XCC_ID register, read by a
// probe grid of 8 work-groups per CU): one queue of the persistent kernel
// each, numbered in order of their ids.
static int probe_xcds(emma_rt_ctx* c) {
  if (c->nxcc > 0) return EMMA_RT_OK;
  const int n = 8 * c->ncu;
  int* d = nullptr;
  HIPCHK(hipMalloc((void**)&d, (size_t)n * sizeof(int)));
  std::vector<int> h(n, -1);
  hipError_t e = launch_xcc_probe(d, n, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(h.data(), d, (size_t)n * sizeof(int), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  HIPCHK(e);
  bool seen[16] = {false};
  for (int x : h)
    if (x >= 0 && x < 16) seen[x] = true;
  int nq = 0;
  for (int i = 0; i < 16; i++) c->xcc_queue[i] = seen[i] ? nq++ : -1;
  if (nq == 0) return EMMA_RT_ERR_HIP;
  c->nxcc = nq;
  return EMMA_RT_OK;
}

static std::vector<SchedLevelTopo> sched_topo(emma_rt_ctx* c) {
  std::vector<SchedLevelTopo> t(c->lev.size());
  for (size_t k = 0; k < c->lev.size(); k++) {
    t[k].noct = c->lev[k].noct;
    t[k].nei = c->lev[k].hnei.data();
    t[k].child = c->lev[k].hchild.data();
  }
  return t;
}

static size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

template <typename R>
static int build_persistent(emma_rt_ctx* c, const std::vector<SchedOp>& ops, Cached& out) {
  const int ng = c->p.ngrp;
  const std::vector<SchedLevelTopo> topo = sched_topo(c);
  const std::vector<double> cost(c->lev.size(), 1.);
  memset(&out.info, 0, sizeof(out.info));
  size_t i = 0;
  while (i < ops.size()) {
    if (ops[i].kind != 0) {
      SchedLaunch l;
      l.kind = ops[i].kind;
      l.tv = ops[i].blob;
      out.launches.push_back(std::move(l));
      i++;
      continue;
    }
    size_t j = i;
    std::vector<int> steps;
    bool coupled = false;
    while (j < ops.size() && ops[j].kind == 0) {
      const StepDesc<R>* D = reinterpret_cast<const StepDesc<R>*>(ops[j].blob.data());
      coupled = coupled || D->fill || D->hydro;
      steps.push_back(ops[j].k);
      j++;
    }
    const int bpc = persistent_blocks_per_cu<R>(ng, coupled);
    if (bpc < 1) return EMMA_RT_ERR_HIP;
    const int blocks = bpc * c->ncu;
    int st = probe_xcds(c);
    if (st) return st;
    SchedPlan plan;
    const int nq = c->nxcc;
    if (plan_schedule(topo, steps, cost, blocks, plan, nullptr, nq) != 0) return EMMA_RT_ERR_ARG;
    const size_t n = (size_t)plan.nitems, m = plan.deps.size();
    const size_t o_desc = 0, o_step = align256(o_desc + steps.size() * sizeof(StepDesc<R>));
    const size_t o_tile = align256(o_step + n * sizeof(int)), o_ntile = align256(o_tile + n * sizeof(int));
    const size_t o_prio = align256(o_ntile + n * sizeof(int));
    const size_t o_doff = align256(o_prio + n), o_deps = align256(o_doff + (n + 1) * sizeof(int));
    const size_t o_done = align256(o_deps + std::max<size_t>(m, 1) * sizeof(int));
    const size_t o_qoff = align256(o_done + n * sizeof(int)), o_ctl = align256(o_qoff + (nq + 1) * sizeof(int));
    const size_t total = o_ctl + (nq + 2) * sizeof(int);
    char* buf = nullptr;
    HIPCHK(hipMalloc((void**)&buf, total));
    out.bufs.push_back(buf);
    std::vector<unsigned char> descs(steps.size() * sizeof(StepDesc<R>));
    for (size_t q = i; q < j; q++) memcpy(descs.data() + (q - i) * sizeof(StepDesc<R>), ops[q].blob.data(), sizeof(StepDesc<R>));
    std::vector<int> ctl0(nq + 2, 0);  // queues' next tickets, work-groups that left, epoch
    ctl0[nq + 1] = 1;
    HIPCHK(hipMemcpy(buf + o_desc, descs.data(), descs.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(buf + o_step, plan.item_step.data(), n * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(buf + o_tile, plan.item_tile.data(), n * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(buf + o_ntile, plan.item_ntile.data(), n * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(buf + o_prio, plan.item_prio.data(), n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(buf + o_doff, plan.dep_off.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice));
    if (m > 0) HIPCHK(hipMemcpy(buf + o_deps, plan.deps.data(), m * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(buf + o_done, 0, n * sizeof(int)));
    HIPCHK(hipMemcpy(buf + o_qoff, plan.qoff.data(), (nq + 1) * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(buf + o_ctl, ctl0.data(), ctl0.size() * sizeof(int), hipMemcpyHostToDevice));
    SchedLaunch l;
    l.kind = 0;
    l.blocks = blocks;
    l.coupled = coupled;
    l.args.steps = buf + o_desc;
    l.args.item_step = (const int*)(buf + o_step);
    l.args.item_tile = (const int*)(buf + o_tile);
    l.args.item_ntile = (const int*)(buf + o_ntile);
    l.args.item_flags = (const unsigned char*)(buf + o_prio);
    l.args.qoff = (const int*)(buf + o_qoff);
    l.args.dep_off = (const int*)(buf + o_doff);
    l.args.deps = (const int*)(buf + o_deps);
    l.args.done = (int*)(buf + o_done);
    l.args.ctl = (int*)(buf + o_ctl);
    l.args.nitems = (int)n;
    l.args.nq = nq;
    for (int i = 0; i < 16; i++) l.args.xcc_queue[i] = c->xcc_queue[i];
    l.args.status = c->d_status;
    l.args.wait_ticks = kSchedWaitTicks;
    out.launches.push_back(std::move(l));
    out.info.items += (long long)n;
    out.info.deps += (long long)m;
    out.info.blocks = std::max(out.info.blocks, blocks);
    out.info.launches++;
    out.info.est_span += plan.est_span;
    out.info.est_chain += plan.est_chain;
    out.info.est_work += plan.est_work;
    i = j;
  }
  return EMMA_RT_OK;
}

template <typename R>
static int run_launches(emma_rt_ctx* c, const std::vector<SchedLaunch>& launches, bool timed) {
  for (const SchedLaunch& l : launches) {
    if (l.kind == 0) {
      hipEvent_t e0 = nullptr, e1 = nullptr;
      if (timed) {
        e0 = next_event(c);
        e1 = next_event(c);
        if (!e0 || !e1) return EMMA_RT_ERR_HIP;
        c->ev_level.push_back(-1);
        HIPCHK(hipEventRecord(e0, c->stream));
      }
      HIPCHK(launch_persistent<R>(l.args, c->p.ngrp, l.coupled, l.blocks, c->stream));
      if (timed) HIPCHK(hipEventRecord(e1, c->stream));
    } else if (l.kind == 1) {
      HIPCHK(launch_put_e2coarse<R>(*reinterpret_cast<const TreeView<R>*>(l.tv.data()), c->stream));
    } else {
      HIPCHK(launch_distrib_e<R>(*reinterpret_cast<const TreeView<R>*>(l.tv.data()), c->stream));
    }
  }
  return EMMA_RT_OK;
}

static int run_cached(emma_rt_ctx* c, const Cached& e, bool timed) {
  return c->rbytes == 8 ? run_launches<double>(c, e.launches, timed) : run_launches<float>(c, e.launches, timed);
}

// record ncoarse coarse steps from the current buffer parity, plan and
// upload them; the parity is left as it was (endpar: after the steps)
static int record_persistent(emma_rt_ctx* c, int ncoarse, double aexp, Cached& e) {
  std::vector<SchedOp> ops;
  std::vector<int> par;
  for (auto& L : c->lev) par.push_back(L.cur);
  c->rec = &ops;
  int rs = EMMA_RT_OK;
  for (int i = 0; i < ncoarse && rs == EMMA_RT_OK; i++) rs = recurse(c, 0, aexp);
  c->rec = nullptr;
  for (auto& L : c->lev) e.endpar.push_back(L.cur);
  for (size_t k = 0; k < c->lev.size(); k++) {
    c->lev[k].cur = par[k];
    c->pending_fill[k].clear();
  }
  if (rs) return rs;
  rs = c->rbytes == 8 ? build_persistent<double>(c, ops, e) : build_persistent<float>(c, ops, e);
  return rs;
}

static void free_cached(Cached& e) {
  if (e.ex) (void)hipGraphExecDestroy(e.ex);
  e.ex = nullptr;
  for (void* b : e.bufs) (void)hipFree(b);
  e.bufs.clear();
}

extern "C" int emma_rt_coarse_step(emma_rt_ctx* c, double aexp, int ncoarse) {
  if (c == nullptr || !c->has_tree) return EMMA_RT_ERR_STATE;
  if (ncoarse < 1) return EMMA_RT_ERR_ARG;
  int st = coupling_ready(c);
  if (st) return st;
  st = ensure_defer_slots(c, steps_per_call(c, ncoarse));
  if (st) return st;
  // replays of the captured coarse step clear the call's device words at
  // the head of the graph; eager launches (TIMING_LAUNCH) clear them here
  const bool eager = c->timing == 1;
  st = begin_call(c, eager);
  if (st) return st;
  if (eager && c->sched == EMMA_RT_SCHED_GRAPH) {  // eager launches with per-kernel events
    for (int i = 0; i < ncoarse; i++) {
      st = recurse(c, 0, aexp);
      if (st) return st;
    }
    return end_call(c);
  }
  GraphKey key;
  key.ncoarse = ncoarse;
  key.aexp = aexp;
  key.xion = c->xion;
  key.mode = c->mode;
  key.sched = c->sched;
  key.counters = c->counters;
  for (auto& L : c->lev) key.parity.push_back(L.cur);
  auto it = c->graphs.find(key);
  if (it == c->graphs.end()) {
    Cached e;
    hipGraph_t g;
    if (c->sched == EMMA_RT_SCHED_PERSISTENT) {
      int rs = record_persistent(c, ncoarse, aexp, e);
      if (rs) {
        free_cached(e);
        return rs;
      }
      hipError_t be = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
      if (be != hipSuccess) {
        free_cached(e);
        HIPCHK(be);
      }
      rs = clear_call_words(c);
      if (rs == EMMA_RT_OK) rs = run_cached(c, e, false);
      hipError_t ce = hipStreamEndCapture(c->stream, &g);
      if (rs || ce != hipSuccess) {
        if (ce == hipSuccess) (void)hipGraphDestroy(g);
        free_cached(e);
        if (rs) return rs;
        HIPCHK(ce);
      }
    } else {
      HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int rs = clear_call_words(c);
      const int timing = c->timing;
      c->timing = 0;  // no event records inside the graph
      for (int i = 0; i < ncoarse && rs == EMMA_RT_OK; i++) rs = recurse(c, 0, aexp);
      c->timing = timing;
      hipError_t ce = hipStreamEndCapture(c->stream, &g);
      for (auto& L : c->lev) e.endpar.push_back(L.cur);
      // nothing ran during the capture: restore the buffer parity the replay
      // starts from (the launch below advances it), also on the error paths
      for (size_t k = 0; k < c->lev.size(); k++) c->lev[k].cur = key.parity[k];
      if (rs || ce != hipSuccess) {
        if (ce == hipSuccess) (void)hipGraphDestroy(g);
        if (rs) return rs;
        HIPCHK(ce);
      }
      memset(&e.info, 0, sizeof(e.info));
    }
    hipError_t ie = hipGraphInstantiate(&e.ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
      e.ex = nullptr;
      free_cached(e);
      HIPCHK(ie);
    }
    it = c->graphs.emplace(key, std::move(e)).first;
    c->stats.level_steps = 0;
  }
  c->sched_info = it->second.info;
  if (c->timing == 1) {  // persistent schedule, eager: one event pair per persistent launch
    st = run_cached(c, it->second, true);
    if (st) return st;
  } else if (c->timing == 2) {  // one event pair around the replay
    hipEvent_t e0 = next_event(c), e1 = next_event(c);
    if (!e0 || !e1) return EMMA_RT_ERR_HIP;
    c->ev_level.push_back(-1);
    HIPCHK(hipEventRecord(e0, c->stream));
    HIPCHK(hipGraphLaunch(it->second.ex, c->stream));
    HIPCHK(hipEventRecord(e1, c->stream));
  } else {
    HIPCHK(hipGraphLaunch(it->second.ex, c->stream));
  }
  for (size_t k = 0; k < c->lev.size(); k++) c->lev[k].cur = it->second.endpar[k];
  c->stats.level_steps = steps_per_call(c, ncoarse);
  return end_call(c);
}

extern "C" int emma_rt_set_schedule(emma_rt_ctx* c, int sched) {
  if (c == nullptr) return EMMA_RT_ERR_STATE;
  if (sched != EMMA_RT_SCHED_GRAPH && sched != EMMA_RT_SCHED_PERSISTENT) return EMMA_RT_ERR_ARG;
  if (sched != c->sched) drop_graphs(c);
  c->sched = sched;
  return EMMA_RT_OK;
}

extern "C" int emma_rt_get_schedule(emma_rt_ctx* c, int* sched, emma_rt_sched_info* info) {
  if (c == nullptr) return EMMA_RT_ERR_STATE;
  if (sched) *sched = c->sched;
  if (info) *info = c->sched_info;
  return EMMA_RT_OK;
}

extern "C" int emma_rt_plan_schedule(int nlev, const int* noct, const int* nei_oct, const int* child, int nsteps,
                                     const int* step_level, const double* tile_cost, const int* chunk, int slots,
                                     int nqueues, long long cap_items, long long cap_deps, long long* nitems,
                                     long long* ndeps, int* qoff, int* item_step, int* item_tile, int* item_ntile,
                                     unsigned char* item_flags, int* item_rank, int* dep_off, int* deps,
                                     double* est) {
  if (nlev < 1 || nlev > kMaxLevels || noct == nullptr || nei_oct == nullptr || child == nullptr || nsteps < 0 ||
      (nsteps > 0 && step_level == nullptr) || nitems == nullptr || ndeps == nullptr)
    return EMMA_RT_ERR_ARG;
  std::vector<SchedLevelTopo> topo(nlev);
  size_t on = 0, oc = 0;
  for (int k = 0; k < nlev; k++) {
    if (noct[k] < 0) return EMMA_RT_ERR_ARG;
    topo[k].noct = noct[k];
    topo[k].nei = nei_oct + on;
    topo[k].child = child + oc;
    on += (size_t)noct[k] * 6;
    oc += (size_t)noct[k] * 8;
  }
  std::vector<int> steps(step_level, step_level + nsteps);
  std::vector<double> cost(nlev, 1.);
  if (tile_cost)
    for (int k = 0; k < nlev; k++) cost[k] = tile_cost[k];
  std::vector<int> cs;
  if (chunk) cs.assign(chunk, chunk + nlev);
  SchedPlan plan;
  if (nqueues < 1 || nqueues > 64) return EMMA_RT_ERR_ARG;
  if (plan_schedule(topo, steps, cost, slots, plan, chunk ? &cs : nullptr, nqueues) != 0) return EMMA_RT_ERR_ARG;
  *nitems = plan.nitems;
  *ndeps = (long long)plan.deps.size();
  if (est) {
    est[0] = plan.est_span;
    est[1] = plan.est_chain;
    est[2] = plan.est_work;
  }
  if (cap_items < plan.nitems || cap_deps < (long long)plan.deps.size()) return cap_items == 0 ? EMMA_RT_OK : EMMA_RT_ERR_ARG;
  if (item_step) memcpy(item_step, plan.item_step.data(), plan.item_step.size() * sizeof(int));
  if (item_tile) memcpy(item_tile, plan.item_tile.data(), plan.item_tile.size() * sizeof(int));
  if (item_ntile) memcpy(item_ntile, plan.item_ntile.data(), plan.item_ntile.size() * sizeof(int));
  if (qoff) memcpy(qoff, plan.qoff.data(), plan.qoff.size() * sizeof(int));
  if (item_flags) memcpy(item_flags, plan.item_prio.data(), plan.item_prio.size());
  if (item_rank) memcpy(item_rank, plan.item_rank.data(), plan.item_rank.size() * sizeof(int));
  if (dep_off) memcpy(dep_off, plan.dep_off.data(), plan.dep_off.size() * sizeof(int));
  if (deps && !plan.deps.empty()) memcpy(deps, plan.deps.data(), plan.deps.size() * sizeof(int));
  return EMMA_RT_OK;
}

extern "C" int emma_rt_set_chem_budget(emma_rt_ctx* c, int level, int budget) {
  if (c == nullptr || !c->has_tree || budget < 0) return EMMA_RT_ERR_ARG;
  if (level != -1 && level_of(c, level) == nullptr) return EMMA_RT_ERR_ARG;
  if (c->d_defer == nullptr && budget > 0) return EMMA_RT_ERR_ARG;  // FP64 or generic-group path: no tiled kernel
  for (auto& L : c->lev)
    if (level == -1 || L.level == level) L.chem_budget = budget;
  drop_graphs(c);
  return EMMA_RT_OK;
}

extern "C" int emma_rt_set_level_variant(emma_rt_ctx* c, int level, int variant) {
  if (c == nullptr || !c->has_tree) return EMMA_RT_ERR_ARG;
  if (variant < EMMA_RT_VARIANT_AUTO || variant > EMMA_RT_VARIANT_LATENCY) return EMMA_RT_ERR_ARG;
  if (level != -1 && level_of(c, level) == nullptr) return EMMA_RT_ERR_ARG;
  for (auto& L : c->lev) {
    if (level != -1 && L.level != level) continue;
    if (variant == EMMA_RT_VARIANT_LISTED && L.halo_per_dir) return EMMA_RT_ERR_ARG;  // more halo faces than slots
  }
  for (auto& L : c->lev)
    if (level == -1 || L.level == level) L.variant = variant == EMMA_RT_VARIANT_AUTO ? default_variant(L) : variant;
  drop_graphs(c);
  return EMMA_RT_OK;
}

extern "C" int emma_rt_set_level_order(emma_rt_ctx* c, int level, int order) {
  if (c == nullptr || !c->has_tree) return EMMA_RT_ERR_STATE;
  if (order != EMMA_RT_ORDER_MORTON && order != EMMA_RT_ORDER_RASTER) return EMMA_RT_ERR_ARG;
  for (int k = 0; k < (int)c->lev.size(); k++) {
    if (level >= 0 && c->lmin + k != level) continue;
    DevLevel& L = c->lev[k];
    if (order == EMMA_RT_ORDER_RASTER && L.morton_m == 0) {
      if (level >= 0) return EMMA_RT_ERR_ARG;
      continue;
    }
    L.raster = order == EMMA_RT_ORDER_RASTER;
  }
  if (level >= 0 && level_of(c, level) == nullptr) return EMMA_RT_ERR_ARG;
  drop_graphs(c);
  return EMMA_RT_OK;
}

extern "C" int emma_rt_get_level_order(emma_rt_ctx* c, int level, int* order, int* cube_side_log2) {
  DevLevel* L = level_of(c, level);
  if (L == nullptr) return EMMA_RT_ERR_ARG;
  if (order) *order = L->raster ? EMMA_RT_ORDER_RASTER : EMMA_RT_ORDER_MORTON;
  if (cube_side_log2) *cube_side_log2 = L->morton_m;
  return EMMA_RT_OK;
}

extern "C" int emma_rt_get_level_tuning(emma_rt_ctx* c, int level, int* variant, int* budget, int* halo_per_dir,
                                        double* keep) {
  DevLevel* L = level_of(c, level);
  if (L == nullptr) return EMMA_RT_ERR_ARG;
  if (variant) *variant = L->variant;
  if (budget) *budget = L->chem_budget;
  if (keep) *keep = L->chem_keep;
  if (halo_per_dir) *halo_per_dir = L->halo_per_dir;
  return EMMA_RT_OK;
}

extern "C" int emma_rt_set_chem_keep(emma_rt_ctx* c, int level, double keep) {
  if (c == nullptr || !c->has_tree || !(keep >= 0.) || keep > 1e30) return EMMA_RT_ERR_ARG;
  if (level != -1 && level_of(c, level) == nullptr) return EMMA_RT_ERR_ARG;
  if (c->d_defer == nullptr && keep > 0.) return EMMA_RT_ERR_ARG;
  for (auto& L : c->lev)
    if (level == -1 || L.level == level) L.chem_keep = (float)keep;
  drop_graphs(c);
  return EMMA_RT_OK;
}

extern "C" int emma_rt_get_chem_budget(emma_rt_ctx* c, int level, int* budget) {
  DevLevel* L = level_of(c, level);
  if (L == nullptr || budget == nullptr) return EMMA_RT_ERR_ARG;
  *budget = L->chem_budget;
  return EMMA_RT_OK;
}

extern "C" int emma_rt_enable_timing(emma_rt_ctx* c, int enable) {
  if (c == nullptr || enable < 0 || enable > 2) return EMMA_RT_ERR_ARG;
  c->timing = enable;
  return EMMA_RT_OK;
}

extern "C" int emma_rt_enable_counters(emma_rt_ctx* c, int enable) {
  if (c == nullptr || enable < 0 || enable > 1) return EMMA_RT_ERR_ARG;
  c->counters = enable;
  return EMMA_RT_OK;
}

// ---------------------------------------------------------------- restart copy
// The arrays a level step changes, for the current buffer parity: the state
// (double-buffered), the source and n_H (FillRad writes them), the reflux
// accumulators, the reionisation times and the hydro fields the injection
// writes.  (pointer, bytes) pairs; absent arrays are skipped.
static std::vector<std::pair<void*, size_t>> snap_parts(emma_rt_ctx* c, DevLevel& L) {
  std::vector<std::pair<void*, size_t>> v;
  const size_t gn = (size_t)c->p.ngrp * L.ncell * c->rbytes, cn = (size_t)L.ncell * sizeof(double);
  const int b = L.cur;
  for (void* p : {L.e[b], L.fx[b], L.fy[b], L.fz[b], L.src}) v.push_back({p, gn});
  for (void* p : {(void*)L.nhplus[b], (void*)L.eint[b], (void*)L.temp[b], (void*)L.nh}) v.push_back({p, cn});
  const size_t accn = (size_t)L.nslot * 4 * (c->p.ngrp == 1 ? 1 : (c->p.ngrp == 3 ? 3 : kMaxGrp));
  if (accn) v.push_back({L.racc, accn * c->rbytes});
  if (L.t_first) v.push_back({(void*)L.t_first, cn});
  if (L.t_last) v.push_back({(void*)L.t_last, cn});
  for (int k = 4; k < 7; k++)
    if (L.hyd[k]) v.push_back({(void*)L.hyd[k], cn});
  return v;
}

extern "C" int emma_rt_save_state(emma_rt_ctx* c) {
  if (c == nullptr || !c->has_tree) return EMMA_RT_ERR_STATE;
  for (auto& L : c->lev) {
    if (L.noct == 0) continue;
    const auto parts = snap_parts(c, L);
    size_t need = 0;
    for (auto& p : parts) need += (p.second + 255) & ~(size_t)255;
    if (L.snap_bytes < need) {
      (void)hipFree(L.snap);
      L.snap = nullptr;
      L.snap_bytes = 0;
      HIPCHK(hipMalloc((void**)&L.snap, need));
      L.snap_bytes = need;
    }
    size_t off = 0;
    for (auto& p : parts) {
      HIPCHK(hipMemcpyAsync(L.snap + off, p.first, p.second, hipMemcpyDeviceToDevice, c->stream));
      off += (p.second + 255) & ~(size_t)255;
    }
    L.snap_cur = L.cur;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return EMMA_RT_OK;
}

extern "C" int emma_rt_restore_state(emma_rt_ctx* c) {
  if (c == nullptr || !c->has_tree) return EMMA_RT_ERR_STATE;
  for (auto& L : c->lev)
    if (L.noct > 0 && L.snap_cur < 0) return EMMA_RT_ERR_STATE;
  for (auto& L : c->lev) {
    if (L.noct == 0) continue;
    L.cur = L.snap_cur;
    const auto parts = snap_parts(c, L);
    size_t off = 0;
    for (auto& p : parts) {
      if (off + p.second > L.snap_bytes) return EMMA_RT_ERR_STATE;  // arrays added since the save
      HIPCHK(hipMemcpyAsync(p.first, L.snap + off, p.second, hipMemcpyDeviceToDevice, c->stream));
      off += (p.second + 255) & ~(size_t)255;
    }
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return EMMA_RT_OK;
}

extern "C" int emma_rt_get_level_times(emma_rt_ctx* c, int nlev, double* ms, int* launches) {
  if (c == nullptr || nlev < 0 || nlev > kMaxLevels) return EMMA_RT_ERR_ARG;
  for (int k = 0; k < nlev; k++) {
    if (ms) ms[k] = c->level_ms[k];
    if (launches) launches[k] = c->level_launches[k];
  }
  return EMMA_RT_OK;
}

extern "C" int emma_rt_get_stats(emma_rt_ctx* c, emma_rt_stats* out) {
  if (c == nullptr || out == nullptr) return EMMA_RT_ERR_ARG;
  *out = c->stats;
  if (c->has_tree) {
    HIPCHK(launch_fill((unsigned int*)c->d_photons, 2, 0u, c->stream));
    for (auto& L : c->lev) {
      const double vol = pow(0.5, 3 * L.level);
      if (c->rbytes == 8)
        HIPCHK(launch_photon_sum<double>((const double*)L.e[L.cur], L.child, L.ncell, c->p.ngrp, vol, c->d_photons,
                                         c->stream));
      else
        HIPCHK(launch_photon_sum<float>((const float*)L.e[L.cur], L.child, L.ncell, c->p.ngrp, vol, c->d_photons,
                                        c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(&out->photons, c->d_photons, sizeof(double), hipMemcpyDeviceToHost));
  }
  return EMMA_RT_OK;
}

// ---------------------------------------------------------------- stencil path
extern "C" int emma_rt_rgrid_layout(int ngrp, int real_bytes, long* out) {
  if (out == nullptr) return EMMA_RT_ERR_ARG;
  RgridLayout L;
  if (!rgrid_layout(ngrp, real_bytes, &L)) return EMMA_RT_ERR_ARG;
  memcpy(out, L.v, sizeof(L.v));
  return EMMA_RT_OK;
}

extern "C" int emma_rt_check_layout(int ngrp, int real_bytes, const long* host) {
  if (host == nullptr) return EMMA_RT_ERR_ARG;
  RgridLayout L;
  if (!rgrid_layout(ngrp, real_bytes, &L)) return EMMA_RT_ERR_ARG;
  for (int i = 0; i < 22; i++)
    if (L.v[i] != host[i]) return EMMA_RT_ERR_LAYOUT;
  return EMMA_RT_OK;
}

extern "C" int emma_rt_create_radstencil(emma_rt_ctx* c, int stride, int real_bytes) {
  if (c == nullptr || stride <= 0 || (real_bytes != 4 && real_bytes != 8)) return EMMA_RT_ERR_ARG;
  RgridLayout L;
  if (!rgrid_layout(c->p.ngrp, real_bytes, &L)) return EMMA_RT_ERR_ARG;
  stencil_free(c->d_stencil);
  c->d_stencil = nullptr;
  HIPCHK(hipMalloc(&c->d_stencil, (size_t)stride * L.v[19]));
  c->stride = stride;
  c->stencil_rbytes = real_bytes;
  return EMMA_RT_OK;
}

extern "C" int emma_rt_destroy_radstencil(emma_rt_ctx* c) {
  if (c == nullptr) return EMMA_RT_ERR_ARG;
  stencil_free(c->d_stencil);
  c->d_stencil = nullptr;
  c->stride = 0;
  return EMMA_RT_OK;
}

extern "C" int emma_rt_create_pinned_stencil(emma_rt_ctx* c, int stride, int real_bytes, void** out) {
  if (c == nullptr || out == nullptr || stride <= 0 || (real_bytes != 4 && real_bytes != 8)) return EMMA_RT_ERR_ARG;
  RgridLayout L;
  if (!rgrid_layout(c->p.ngrp, real_bytes, &L)) return EMMA_RT_ERR_ARG;
  *out = nullptr;
  HIPCHK(hipHostMalloc(out, (size_t)stride * L.v[19], hipHostMallocDefault));
  memset(*out, 0, (size_t)stride * L.v[19]);
  return EMMA_RT_OK;
}

extern "C" int emma_rt_alloc_pinned(size_t bytes, void** out) {
  if (out == nullptr || bytes == 0) return EMMA_RT_ERR_ARG;
  *out = nullptr;
  HIPCHK(hipHostMalloc(out, bytes, hipHostMallocDefault));
  memset(*out, 0, bytes);
  return EMMA_RT_OK;
}

extern "C" int emma_rt_destroy_pinned_stencil(void* stencil) {
  if (stencil != nullptr) HIPCHK(hipHostFree(stencil));
  return EMMA_RT_OK;
}

// the stencil path on a device staging area `dbuf` holding at least nread
// records of REAL = real_bytes
static int advance_stencil_on(emma_rt_ctx* c, void* dbuf, int real_bytes, void* stencil, int nread, int level,
                              double dxcur, double dtnew, double aexp, int chemonly) {
  if (c->mode == EMMA_RT_COARSERAD && c->p.lcoarse < 1) return EMMA_RT_ERR_STATE;
  if (nread == 0) return EMMA_RT_OK;
  // condadvec of advancerad / advanceradGPU (rad_utils.c:1808-1812,
  // rad_utils_gpu.cu:1058-1062)
  const int coarserad = c->mode == EMMA_RT_COARSERAD;
  const int transport = coarserad ? (level == c->p.lcoarse && !chemonly) : 1;
  RgridLayout L;
  rgrid_layout(c->p.ngrp, real_bytes, &L);
  int st = begin_call(c);
  if (st) return st;
  // The kernel reads the 7 oct slots and each centre cell's rfieldnew and
  // writes rfieldnew and rflux (every entry); deltaR, rfield and the fluxes
  // of the host's buffer are neither read nor changed.  So only those parts
  // cross PCIe: the oct slots (contiguous at the start of each record), then
  // per centre cell its rfieldnew, and back per centre cell rfieldnew + rflux
  // (contiguous at the end of CELLFLUX_R).
  const size_t rsize = (size_t)L.v[19], ooff = (size_t)L.v[20], osize = 7 * (size_t)L.v[17];
  const size_t noff = (size_t)L.v[21], fsize = (size_t)L.v[12], rfn = (size_t)L.v[15], rtype = (size_t)L.v[0];
  const size_t back = fsize - rfn;  // rfieldnew + rflux
  char* h = (char*)stencil;
  char* d = (char*)dbuf;
  HIPCHK(hipMemcpy2DAsync(d + ooff, rsize, h + ooff, rsize, osize, (size_t)nread, hipMemcpyHostToDevice, c->stream));
  for (int ic = 0; ic < 8; ic++) {
    const size_t o = noff + ic * fsize + rfn;
    HIPCHK(hipMemcpy2DAsync(d + o, rsize, h + o, rsize, rtype, (size_t)nread, hipMemcpyHostToDevice, c->stream));
  }
  const ChemConst K = make_chem(c->p, dtnew, aexp);
  const double cl = cloc_of(c->p, aexp);
  HIPCHK(launch_stencil(dbuf, L, c->p.ngrp, real_bytes, nread, cl, dtnew / dxcur, chemonly, coarserad, transport, K,
                        c->d_status, c->counters ? c->d_counters : nullptr, c->stream));
  // without transport the kernel writes no fluxes: rfieldnew alone comes back
  const size_t nback = transport ? back : rtype;
  for (int ic = 0; ic < 8; ic++) {
    const size_t o = noff + ic * fsize + rfn;
    HIPCHK(hipMemcpy2DAsync(h + o, rsize, d + o, rsize, nback, (size_t)nread, hipMemcpyDeviceToHost, c->stream));
  }
  return end_call(c);
}

extern "C" int emma_rt_advance_stencil(emma_rt_ctx* c, void* stencil, int nread, int level, double dxcur,
                                       double dtnew, double aexp, int chemonly) {
  if (c == nullptr || stencil == nullptr || nread < 0) return EMMA_RT_ERR_ARG;
  if (c->d_stencil == nullptr) return EMMA_RT_ERR_STATE;
  if (nread > c->stride) return EMMA_RT_ERR_ARG;
  return advance_stencil_on(c, c->d_stencil, c->stencil_rbytes, stencil, nread, level, dxcur, dtnew, aexp, chemonly);
}

extern "C" int emma_rt_advance_stencil_noalloc(emma_rt_ctx* c, void* stencil, int nread, int level, double dxcur,
                                               double dtnew, double aexp, int chemonly, int real_bytes) {
  if (c == nullptr || stencil == nullptr || nread < 0 || (real_bytes != 4 && real_bytes != 8)) return EMMA_RT_ERR_ARG;
  RgridLayout L;
  if (!rgrid_layout(c->p.ngrp, real_bytes, &L)) return EMMA_RT_ERR_ARG;
  if ((size_t)nread * (size_t)L.v[19] > sizeof(scratch)) return EMMA_RT_ERR_ARG;  // beyond the staging area
  if (c->scratch == nullptr) return EMMA_RT_ERR_STATE;
  return advance_stencil_on(c, c->scratch, real_bytes, stencil, nread, level, dxcur, dtnew, aexp, chemonly);
}

// ---------------------------------------------------------------- grid output
// The serial STDIO grid output of dumpalloct_serial (io.c:1233-1363) for the
// radiation, chemistry and hydro fields the library holds: per field id one
// file <folder>grid_<name>/<name>.<ndump %05d>.p<rank %05d> holding int n,
// float tsim, float xmin, xmax, ymin, ymax, zmin, zmax, then n floats, one per
// leaf cell, levels lcoarse..lmax, octs in level order, cells 0..7.  Field
// values are rounded to float as assign_grid_field (io.c:32-200) returns them.
namespace {
// the field names of readOutputParam_grid (parameters.c:132-177), by id
const char* const kGridFieldName[] = {
    "x", "y", "z", "l", "cpu", "gdata_d", "density", "gdata.p", "res", "f0", "f1", "f2", "marked",
    "field_d", "field_u", "field_v", "field_w", "field_p", "rfield_e0", "rfield_fx0", "rfield_fy0",
    "rfield_fz0", "rfield_e1", "rfield_fx1", "rfield_fy1", "rfield_fz1", "rfield_snfb", "rfield_e2",
    "rfield_fx2", "rfield_fy2", "rfield_fz2", "rfield_src", "xion", "field_dX", "field_dXHE", "field_dXXHE",
    "field_xHE", "field_xxHE", "rfield_temp", "t_first_xion", "t_last_xion", "rho_first_xion"};
constexpr int kGridFields = 42;

struct HostLevel {
  std::vector<double> e, fx, fy, fz, src, nhplus, nh, temp, tf, tl, hyd[6], opos;
};
}  // namespace

// whether the library holds field `id` for this context (1), or not (0)
static int grid_field_ok(emma_rt_ctx* c, int id) {
  const int G = c->p.ngrp;
  bool hydro = true, xion = c->xion != 0;
  for (auto& L : c->lev)
    if (L.noct > 0 && L.hyd[0] == nullptr) hydro = false;
  if (id >= 0 && id <= 4) return 1;                       // position, level, cpu
  if (id >= 13 && id <= 17) return hydro;                 // d u v w p
  if (id >= 18 && id <= 25) return ((id - 18) / 4) < G;  // group 0, 1
  if (id >= 27 && id <= 30) return G > 2;                 // group 2
  if (id == 31 || id == 32 || id == 38) return 1;         // source sum, x, T
  if (id == 33) return hydro;                             // dX
  if (id == 39 || id == 40) return xion;                  // reionisation times
  return 0;  // gravity, marked, snfb, helium, rho_first_xion: not held here
}

// the levels' host copies of what the grid fields read
static int grid_snapshot(emma_rt_ctx* c, std::vector<HostLevel>& H) {
  const int G = c->p.ngrp, nl = (int)c->lev.size();
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int k = 0; k < nl; k++) {
    DevLevel& L = c->lev[k];
    if (L.noct == 0) continue;
    HostLevel& h = H[k];
    const size_t n = L.ncell, gn = (size_t)G * n;
    for (auto* v : {&h.e, &h.fx, &h.fy, &h.fz, &h.src}) v->resize(gn);
    for (auto* v : {&h.nhplus, &h.nh, &h.temp}) v->resize(n);
    int st = emma_rt_get_level(c, L.level, h.e.data(), h.fx.data(), h.fy.data(), h.fz.data(), h.src.data(),
                               h.nhplus.data(), nullptr, h.nh.data(), h.temp.data());
    if (st) return st;
    h.opos.resize((size_t)L.noct * 3);
    HIPCHK(hipMemcpy(h.opos.data(), L.opos, h.opos.size() * sizeof(double), hipMemcpyDeviceToHost));
    if (c->xion && L.t_first) {
      h.tf.resize(n);
      h.tl.resize(n);
      HIPCHK(hipMemcpy(h.tf.data(), L.t_first, n * sizeof(double), hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(h.tl.data(), L.t_last, n * sizeof(double), hipMemcpyDeviceToHost));
    }
    for (int j = 0; j < 6; j++) {
      if (L.hyd[j] == nullptr) continue;
      h.hyd[j].resize(n);
      HIPCHK(hipMemcpy(h.hyd[j].data(), L.hyd[j], n * sizeof(double), hipMemcpyDeviceToHost));
    }
  }
  return EMMA_RT_OK;
}

// assign_grid_field (io.c:32-200): field `id` of cell i of level k, as float
static float grid_value(emma_rt_ctx* c, const std::vector<HostLevel>& H, int id, int k, size_t i, int rank) {
  const int G = c->p.ngrp;
  const HostLevel& h = H[k];
  const DevLevel& L = c->lev[k];
  const size_t n = L.ncell, o = i >> 3, ic = i & 7;
  const double dxc = pow(0.5, L.level);
  {
    switch (id) {
      case 0: return (float)(h.opos[o * 3 + 0] + (double)(ic & 1) * dxc);
      case 1: return (float)(h.opos[o * 3 + 1] + (double)((ic >> 1) & 1) * dxc);
      case 2: return (float)(h.opos[o * 3 + 2] + (double)(ic >> 2) * dxc);
      case 3: return (float)L.level;
      case 4: return (float)rank;
      case 13: case 14: case 15: case 16: case 17: return (float)h.hyd[id - 13][i];
      case 31: {  // the reference sums into a float
        float r = 0.f;
        for (int g = 0; g < G; g++) r = (float)((double)r + h.src[g * n + i]);
        return r;
      }
      case 32: return (float)(h.nhplus[i] / h.nh[i]);
      case 33: return (float)h.hyd[5][i];
      case 38: return (float)h.temp[i];
      case 39: return (float)h.tf[i];
      case 40: return (float)h.tl[i];
      default: break;
    }
    const int g = id < 26 ? (id - 18) / 4 : 2, q = id < 26 ? (id - 18) % 4 : id - 27;
    const std::vector<double>& a = q == 0 ? h.e : (q == 1 ? h.fx : (q == 2 ? h.fy : h.fz));
    return (float)a[g * n + i];
  }
}

// the dumped cells: leaves, and every cell of level lmax (io.c:1322)
static bool grid_dumped(emma_rt_ctx* c, const DevLevel& L, size_t i) { return L.hchild[i] < 0 || L.level == c->lmax; }

static int check_grid_ids(emma_rt_ctx* c, int nfield, const int* field_ids, bool* want) {
  for (int f = 0; f < nfield; f++) {
    if (field_ids[f] < 0 || field_ids[f] >= kGridFields || !grid_field_ok(c, field_ids[f])) return EMMA_RT_ERR_ARG;
    if (want) want[field_ids[f]] = true;
  }
  if (!opos_ready(c)) return EMMA_RT_ERR_STATE;  // positions and file bounds need the octs' corners
  return EMMA_RT_OK;
}

extern "C" int emma_rt_grid_field(emma_rt_ctx* c, int field_id, int rank, float* out, long long capacity,
                                  long long* n_out) {
  if (c == nullptr || n_out == nullptr || capacity < 0 || rank < 0) return EMMA_RT_ERR_ARG;
  if (!c->has_tree) return EMMA_RT_ERR_STATE;
  int st = check_grid_ids(c, 1, &field_id, nullptr);
  if (st) return st;
  long long n = 0;
  for (auto& L : c->lev)
    for (size_t i = 0; i < (size_t)L.ncell; i++) n += grid_dumped(c, L, i);
  *n_out = n;
  if (out == nullptr) return EMMA_RT_OK;  // size query
  if (capacity < n) return EMMA_RT_ERR_ARG;
  std::vector<HostLevel> H(c->lev.size());
  st = grid_snapshot(c, H);
  if (st) return st;
  long long j = 0;
  for (int k = 0; k < (int)c->lev.size(); k++)
    for (size_t i = 0; i < (size_t)c->lev[k].ncell; i++)
      if (grid_dumped(c, c->lev[k], i)) out[j++] = grid_value(c, H, field_id, k, i, rank);
  return EMMA_RT_OK;
}

extern "C" int emma_rt_dump_grid(emma_rt_ctx* c, const char* folder, int ndump, int rank, double tsim, int nfield,
                                 const int* field_ids) {
  if (c == nullptr || folder == nullptr || nfield < 1 || field_ids == nullptr || ndump < 0 || rank < 0)
    return EMMA_RT_ERR_ARG;
  if (!c->has_tree) return EMMA_RT_ERR_STATE;
  bool want[kGridFields] = {false};
  int st = check_grid_ids(c, nfield, field_ids, want);
  if (st) return st;
  const int nl = (int)c->lev.size();
  std::vector<HostLevel> H(nl);
  st = grid_snapshot(c, H);
  if (st) return st;
  // file bounds over the dumped cells' octs (oct corner, oct corner + oct size)
  float xmin = 2.f, xmax = -1.f, ymin = 2.f, ymax = -1.f, zmin = 2.f, zmax = -1.f;
  int ncell_out = 0;
  for (int k = 0; k < nl; k++) {
    const DevLevel& L = c->lev[k];
    const double dx = pow(0.5, L.level - 1);
    for (int o = 0; o < L.noct; o++) {
      const double* p = &H[k].opos[(size_t)o * 3];
      for (int ic = 0; ic < 8; ic++) {
        if (!grid_dumped(c, L, (size_t)o * 8 + ic)) continue;
        ncell_out++;
        // the reference compares in float against its float bounds
        if (p[0] < xmin) xmin = (float)p[0];
        if (p[1] < ymin) ymin = (float)p[1];
        if (p[2] < zmin) zmin = (float)p[2];
        if (p[0] + dx > xmax) xmax = (float)(p[0] + dx);
        if (p[1] + dx > ymax) ymax = (float)(p[1] + dx);
        if (p[2] + dx > zmax) zmax = (float)(p[2] + dx);
      }
    }
  }
  std::vector<float> buf;
  for (int id = 0; id < kGridFields; id++) {
    if (!want[id]) continue;
    std::string dir = std::string(folder) + "grid_" + kGridFieldName[id] + "/";
    (void)mkdir(dir.c_str(), 0755);
    char name[64];
    snprintf(name, sizeof(name), "%s.%05d.p%05d", kGridFieldName[id], ndump, rank);
    FILE* fp = fopen((dir + name).c_str(), "wb");
    if (fp == nullptr) return EMMA_RT_ERR_ARG;
    // the header's time: the reference writes sizeof(float) bytes of its
    // REAL tsim (io.c:1282), i.e. the first four bytes of the double
    unsigned char tbytes[sizeof(float)];
    memcpy(tbytes, &tsim, sizeof(tbytes));
    const float bounds[6] = {xmin, xmax, ymin, ymax, zmin, zmax};
    bool ok = fwrite(&ncell_out, sizeof(int), 1, fp) == 1 && fwrite(tbytes, sizeof(tbytes), 1, fp) == 1 &&
              fwrite(bounds, sizeof(float), 6, fp) == 6;
    for (int k = 0; k < nl && ok; k++) {
      const DevLevel& L = c->lev[k];
      buf.clear();
      for (size_t i = 0; i < (size_t)L.ncell; i++)
        if (grid_dumped(c, L, i)) buf.push_back(grid_value(c, H, id, k, i, rank));
      if (!buf.empty()) ok = fwrite(buf.data(), sizeof(float), buf.size(), fp) == buf.size();
    }
    if (fclose(fp) != 0) ok = false;
    if (!ok) return EMMA_RT_ERR_ARG;
  }
  return EMMA_RT_OK;
}
