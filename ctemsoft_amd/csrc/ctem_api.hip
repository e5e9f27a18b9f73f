// C-ABI of the EBSD master-pattern path (include/ctem_ebsd.h): contexts, crystal upload,
// incident directions, batched plans and the CalcLgh3 drop-in. Host orchestration only;
// the arithmetic runs in the kernels of ctem_kernels.hip and ctem_mixed.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ctem_ebsd.h"
#include "../../include/ctem_ecp.h"
#include "ctem_internal.h"
#include "ctem_laue.h"
#include "ctem_sg.h"

using namespace ctem;

namespace {

thread_local std::string g_err;

int32_t fail(int32_t code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return fail(CTEM_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

const double kPi = 3.141592653589793238;

// Device allocation owned by a context or plan.
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void release() { if (p) (void)hipFree(p); p = nullptr; bytes = 0; }
  int32_t ensure(size_t b) {
    if (b <= bytes && p) return CTEM_OK;
    release();
    if (b == 0) return CTEM_OK;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) { p = nullptr; return fail(CTEM_E_HIP, "hipMalloc(%zu): %s", b, hipGetErrorString(e)); }
    bytes = b;
    return CTEM_OK;
  }
  template <class T> T* as() const { return (T*)p; }
};

template <class T>
int32_t upload(DBuf& d, const T* h, size_t n, hipStream_t s) {
  int32_t rc = d.ensure(n * sizeof(T));
  if (rc) return rc;
  if (n) HIPCHK(hipMemcpyAsync(d.p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  return CTEM_OK;
}

// Taylor coefficients 1/j!
double inv_fact(int j) { double f = 1.0; for (int i = 2; i <= j; i++) f *= i; return 1.0 / f; }

// bucket sizes: multiples of 32 (or of a coarser multiple of 32)
int npad_of(int n, int multiple = 32) { return std::max(multiple, (n + multiple - 1) / multiple * multiple); }

}  // namespace

// The per-site structure-matrix tables S_p(g-h) and their B-derivatives over the box of all
// differences of list entries. They depend on the sites, the metric and the box only, not on the
// energy, so the crystals of an energy series share one (the key holds those inputs' bytes).
struct SiteTable {
  int device = 0;
  DBuf slut, dslut;
  std::vector<unsigned char> key;
  ~SiteTable() {
    (void)hipSetDevice(device);
    slut.release();
    dslut.release();
  }
};

// One uploaded crystal (one energy bin). A plan keeps the crystal that was current when it
// was created, so that several bins can be resident at once under one context.
struct Crystal {
  int device = 0;
  DBuf g, hkl, cut1, cut2, dbd, lut, rmt;
  std::shared_ptr<SiteTable> st;
  int32_t nref = 0, nsite = 0, ncell = 0;
  int32_t lut_lo[3] = {0, 0, 0}, lut_dim[3] = {0, 0, 0};
  int32_t slut_lo[3] = {0, 0, 0}, slut_dim[3] = {0, 0, 0};
  int64_t sncell = 0;
  double mlambda = 0.0;
  float dynupz = 0.0f, sgcutoff = 0.0f;
  float nat = 0.0f;  // total orbit size, float(sum(nat))
  float lutmax = 0.0f;  // max |re|, |im| over the float32 LUT
  // the LUT is inversion-symmetric (U(-g) = U(g), a centrosymmetric structure with the origin
  // on an inversion centre) to within 2^-20 of its max: DynMat is then complex symmetric to
  // far below the mixed form's storage resolution
  bool centro = false;
  std::vector<int32_t> h_hkl;
  // host copies of the sites for the structure-matrix tables of later plans
  std::vector<int32_t> h_site_z, h_norbit;
  std::vector<float> h_site_b;
  std::vector<double> h_orbit;
  double h_rmt[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};

  ~Crystal() {
    (void)hipSetDevice(device);
    for (DBuf* b : {&g, &hkl, &cut1, &cut2, &dbd, &lut, &rmt}) b->release();
  }
  DevRefl refl() const {
    DevRefl R;
    R.nref = nref; R.g = g.as<float>(); R.hkl = hkl.as<int32_t>(); R.cut1 = cut1.as<float>();
    R.cut2 = cut2.as<float>(); R.dbd = dbd.as<uint8_t>(); R.sgcutoff = sgcutoff;
    return R;
  }
  DevLut lutv(int32_t* err) const {
    DevLut L;
    L.lut = lut.as<float2>(); L.slut = st->slut.as<double2>();
    for (int i = 0; i < 3; i++) {
      L.lo[i] = lut_lo[i]; L.dim[i] = lut_dim[i];
      L.slo[i] = slut_lo[i]; L.sdim[i] = slut_dim[i];
    }
    L.ncell = ncell; L.sncell = sncell; L.nsite = nsite; L.err = err;
    return L;
  }
};

// Work space of ctem_plan_run and ctem_plan_adjoint, shared by the plans of a context (calls
// on one context are sequential); it grows to the largest chunk run and is kept.
struct Work {
  DBuf A, W, Psi, sq, dirs, partial, wmax, colsum, order, amod;
  DBuf g_dots, g_lut, g_dots2, aM, aQ, aPhi, aHm, aEbar, aBlk, aW2, aDbar, aSig, aMax;
  void release() {
    for (DBuf* b : {&A, &W, &Psi, &sq, &dirs, &partial, &wmax, &colsum, &order, &amod, &g_dots, &g_lut, &g_dots2, &aM, &aQ, &aPhi,
                    &aHm, &aEbar, &aBlk, &aW2, &aDbar, &aSig, &aMax})
      b->release();
  }
};

struct ctem_ctx {
  int device = 0;
  int cu_count = 256;
  hipStream_t stream = nullptr;
  std::shared_ptr<Crystal> cr;  // the crystal that plans created next will use
  std::weak_ptr<SiteTable> last_st;  // the latest S table, reused while a crystal holds it
  Work w;
  // scratch for the CalcLgh3 drop-in
  DBuf s_A, s_W, s_Psi, s_sq, s_err, s_lam, s_lgh, s_dirs, s_nns, s_zero, s_off, s_part, s_dots;
  hipEvent_t ev[2] = {nullptr, nullptr};
  std::vector<ctem_plan*> plans;  // live plans, destroyed with the context
  unsigned char* scratch = nullptr;  // this device's static work space (ctem_calclgh3_static)
  // bumped whenever a work-space buffer is reallocated: graphs captured before hold the old
  // addresses and are captured again
  uint64_t wgen = 0;
  uint64_t plan_serial = 0;  // plans are numbered for graph keys
  // one captured graph over a list of plans (ctem_plans_run)
  hipGraphExec_t mgraph = nullptr;
  std::vector<uint64_t> mgraph_key;
};

// squaring counts of one chunk of matrices and their maximum
struct ChunkSched {
  bool known = false;
  std::vector<int32_t> sq;
  int smax = 0;
};

struct ctem_plan {
  ctem_ctx* ctx = nullptr;
  int32_t numk = 0, nz = 0;
  double dz = 0, thick = 0;
  DBuf k, kn, lam, ij;
  // selection
  bool selected = false;
  DBuf nns, nnw, soff, woff, sidx, ssg, widx, wsg, err;
  std::vector<int32_t> h_nns, h_nnw;
  std::vector<int64_t> h_soff, h_woff;
  // results
  DBuf dots;
  bool ran = false;
  int32_t keep_dir = -1, keep_n = 0;
  DBuf keep_lgh;
  std::shared_ptr<Crystal> cr;  // the crystal of this plan (current at creation)
  std::vector<int32_t> h_sq;
  // timing of the last run
  double t_total = 0, t_gemm = 0, f_gemm = 0;
  int32_t n_gemm = 0;
  // algorithmic flops of the last FP64 run by the reference's Pade rule (math.f90:208-219):
  // total F = 8n^2w + 8n^3(q+j) + (32/3)n^3 + 16n^2Z + 4n^2P, its product part 8n^3(q+j) and
  // its solve part (32/3)n^3, summed over the directions
  double fa_total = -1, fa_prod = -1, fa_solve = -1;
  int32_t terms_taylor = CTEM_SPLIT_FP16X3, terms_square = CTEM_SPLIT_FP16X3;  // mixed-form split schemes
  int32_t symmetry = 1;  // split mixed form: symmetric products when the crystal's LUT allows
  std::vector<hipEvent_t> evpool;
  size_t evused = 0;
  // bucket layout from the last selection: (npad, directions) per bucket, and the bucket order
  // of all directions on the device
  std::vector<std::pair<int, std::vector<int32_t>>> buckets;
  DBuf dirs;
  // squaring counts of each chunk from an earlier run; the same directions and crystal give
  // the same matrices, so later runs need no read-back. Key: (form, npad, chunk start).
  std::map<std::tuple<int, int, int>, ChunkSched> sched;
  uint64_t serial = 0;
  // tuning: padded sizes are multiples of bucket_multiple (32 or 64); a chunk of one bucket
  // holds as many directions as fit in chunk_bytes of work space
  int32_t bucket_multiple = 32;
  size_t chunk_bytes = (size_t)24 << 30;
  uint64_t bufsig = 0;  // addresses of the beam-set buffers of the last selection
  // graph mode: the whole run captured once and replayed (ctem_plan_set_graph)
  int32_t graph_mode = 0;
  hipGraphExec_t graph = nullptr;
  std::vector<uint64_t> graph_key;
  // electron channeling pattern plan (ctem_ecp_plan_create): mode, thicknesses, the foil normal
  // of the selection, the channeling structure matrix table, ScatMat's store map and the results
  int32_t ecp_mode = -1;  // -1: a master-pattern plan
  int32_t ecp_nt = 0, ecp_numt = 0;
  float ecp_t0 = 0.0f, ecp_dt = 0.0f, ecp_zint = 0.0f;
  double ecp_nat = 1.0, ecp_s0 = 1.0;
  DBuf fn, ecp_slut, ecp_store, ecp_thick, ecp_sr, ecp_part;
};

// ------------------------------------------------------------------------------------------
// ------------------------------------------------------------------------------------------
// Work space of the allocation-free CalcLgh3 entry (ctem_calclgh3_static): one array in device
// memory per GPU, for matrices of up to kStaticMaxN strong beams and kStaticMaxZ depth steps.
// Each GPU the library runs on holds it until the process ends.
constexpr int kStaticMaxN = 256;
constexpr int kStaticMaxZ = 256;
constexpr size_t kStaticMat = (size_t)kStaticMaxN * kStaticMaxN * sizeof(double2);  // 1 MiB
constexpr size_t kStaticBytes = 8 * kStaticMat + 64 * 1024;                         // 8 MiB + 64 KiB
static __device__ __align__(256) unsigned char scratch[kStaticBytes];

namespace {
// device address of `scratch` on each device, looked up once per device with that device
// current (the host-side name is only a stand-in for the symbol, not a device address)
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;

int32_t scratch_address(int device, void** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(CTEM_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
  }
  *out = g_scratch_ptr[device];
  return CTEM_OK;
}
}  // namespace

extern "C" const char* ctem_last_error(void) { return g_err.c_str(); }

extern "C" int32_t ctem_ctx_create(int32_t device, ctem_ctx** out) {
  if (!out) return fail(CTEM_E_ARG, "ctem_ctx_create: out is NULL");
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0)
    return fail(CTEM_E_HIP, "no HIP device available (%s)", e == hipSuccess ? "count 0" : hipGetErrorString(e));
  if (device < 0 || device >= ndev) return fail(CTEM_E_ARG, "device %d out of range (%d devices)", device, ndev);
  HIPCHK(hipSetDevice(device));
  ctem_ctx* c = new ctem_ctx();
  c->device = device;
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
      c->cu_count = prop.multiProcessorCount;
  }
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete c; return fail(CTEM_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e)); }
  (void)hipEventCreate(&c->ev[0]);
  (void)hipEventCreate(&c->ev[1]);
  void* sp = nullptr;
  int32_t rc = scratch_address(device, &sp);
  if (rc) { ctem_ctx_destroy(c); return rc; }
  c->scratch = (unsigned char*)sp;
  *out = c;
  return CTEM_OK;
}

extern "C" int32_t ctem_ctx_destroy(ctem_ctx* c) {
  if (!c) return CTEM_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  while (!c->plans.empty()) ctem_plan_destroy(c->plans.back());
  if (c->mgraph) (void)hipGraphExecDestroy(c->mgraph);
  c->cr.reset();
  c->w.release();
  for (DBuf* b : {&c->s_A, &c->s_W, &c->s_Psi, &c->s_sq, &c->s_err, &c->s_lam, &c->s_lgh, &c->s_dirs,
                  &c->s_nns, &c->s_zero, &c->s_off, &c->s_part, &c->s_dots})
    b->release();
  for (auto& e : c->ev) if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return CTEM_OK;
}

extern "C" int32_t ctem_device_info(ctem_ctx* c, char* name, int32_t name_len, int32_t* cu_count) {
  if (!c) return fail(CTEM_E_ARG, "null context");
  hipDeviceProp_t p;
  HIPCHK(hipGetDeviceProperties(&p, c->device));
  if (name && name_len > 0) { snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName); }
  if (cu_count) *cu_count = p.multiProcessorCount;
  return CTEM_OK;
}

// ------------------------------------------------------------------------------------------
// per-site structure-matrix table: S_p(d) = sum over the orbit of
// cmplx(Z^2 DBWF) * cmplx(cos, sin) in float32, summed in double, for every d = g-h of the
// LUT box (CTEMEBSDmaster.f90:418-449; DBWF = exp(-B |d|^2/4) with |d| from the float32
// CalcLength, no occupancy factor).
// S_p(d) for every d in the box lo..lo+dim-1, site after site, column-major (h fastest)
// dout receives dS_p/dB_p = sum over the orbit of -|d|^2/4 Z^2 DBWF (cos, sin), in double
// (the derivative of the smooth function the float32 terms round).
// Cells are independent: the (site, l) planes are shared out over host threads, each cell
// computed exactly as in a serial loop.
static void build_slut(const ctem_crystal* c, const int32_t lo[3], const int32_t dim[3],
                       std::vector<double>& out, std::vector<double>& dout) {
#pragma clang fp contract(off)
  const size_t ncell = (size_t)dim[0] * dim[1] * dim[2];
  out.assign((size_t)2 * ncell * c->nsite, 0.0);
  dout.assign((size_t)2 * ncell * c->nsite, 0.0);
  const double tpi = 2.0 * kPi;
  std::vector<size_t> obase(c->nsite + 1, 0);
  for (int p = 0; p < c->nsite; p++) obase[p + 1] = obase[p] + c->site_norbit[p];
  auto plane = [&](int p, int l) {
#pragma clang fp contract(off)
    const float zf = (float)c->site_z[p];
    const double znsq = (double)(zf * zf);
    const int no = c->site_norbit[p];
    const double* orb = c->orbit + 3 * obase[p];
    for (int kk = 0; kk < dim[1]; kk++)
      for (int h = 0; h < dim[0]; h++) {
        const int d[3] = {h + lo[0], kk + lo[1], l + lo[2]};
        const float df[3] = {(float)d[0], (float)d[1], (float)d[2]};
        const float len = std::sqrt(dot_r(df, df, c->rmt));
        const float len2 = len * len;
        const double kkl = (double)(0.25f * len2);
        double sr = 0.0, si = 0.0, dr = 0.0, di = 0.0;
        for (int a = 0; a < no; a++) {
          double s = 0.0;
          for (int t = 0; t < 3; t++) s += (double)d[t] * orb[3 * a + t];
          const double arg = tpi * s;
          const double dbwf = std::exp(-((double)c->site_b[p] * kkl));
          const float pre = (float)(znsq * dbwf);
          double sd, cd;
          ::sincos(arg, &sd, &cd);  // one sincos call, as the reference's compiled code makes
          const float cs = (float)cd, sn = (float)sd;
          const float re = pre * cs, im = pre * sn;
          sr += (double)re;
          si += (double)im;
          dr += -kkl * znsq * dbwf * cd;
          di += -kkl * znsq * dbwf * sd;
        }
        const size_t cell = (size_t)h + (size_t)dim[0] * ((size_t)kk + (size_t)dim[1] * l);
        out[2 * ((size_t)p * ncell + cell)] = sr;
        out[2 * ((size_t)p * ncell + cell) + 1] = si;
        dout[2 * ((size_t)p * ncell + cell)] = dr;
        dout[2 * ((size_t)p * ncell + cell) + 1] = di;
      }
  };
  const int nplane = c->nsite * dim[2];
  const int nthr = std::max(1, std::min<int>(std::min(16u, std::max(1u, std::thread::hardware_concurrency())), nplane));
  std::atomic<int> next{0};
  auto work = [&]() {
    for (int q; (q = next.fetch_add(1)) < nplane;) plane(q / dim[2], q % dim[2]);
  };
  std::vector<std::thread> pool;
  try {
    for (int t = 1; t < nthr; t++) pool.emplace_back(work);
  } catch (...) {
    // no more threads: the calling thread and those started share the planes
  }
  work();
  for (auto& t : pool) t.join();
}

extern "C" int32_t ctem_upload_crystal(ctem_ctx* c, const ctem_crystal* x) {
#pragma clang fp contract(off)
  if (!c || !x) return fail(CTEM_E_ARG, "ctem_upload_crystal: null argument");
  if (x->nref < 1 || !x->hkl || !x->ucg || !x->dbdiff || !x->lut)
    return fail(CTEM_E_ARG, "ctem_upload_crystal: empty reflection list or LUT");
  if (x->hkl[0] != 0 || x->hkl[1] != 0 || x->hkl[2] != 0)
    return fail(CTEM_E_ARG, "ctem_upload_crystal: the first list entry must be 000");
  if (x->nsite < 0 || x->nsite > 4) return fail(CTEM_E_RANGE, "ctem_upload_crystal: 0..4 sites supported (got %d)", x->nsite);
  if (x->nsite > 0 && (!x->site_z || !x->site_b || !x->site_norbit || !x->orbit))
    return fail(CTEM_E_ARG, "ctem_upload_crystal: null site arrays");
  for (int p = 0; p < x->nsite; p++)
    if (x->site_norbit[p] < 1) return fail(CTEM_E_ARG, "ctem_upload_crystal: empty orbit");
  for (int i = 0; i < 3; i++)
    if (x->lut_dim[i] <= 0) return fail(CTEM_E_ARG, "ctem_upload_crystal: bad LUT extent");
  HIPCHK(hipSetDevice(c->device));
  const int n = x->nref;
  const int64_t ncell64 = (int64_t)x->lut_dim[0] * x->lut_dim[1] * x->lut_dim[2];
  if (ncell64 > (1ll << 30)) return fail(CTEM_E_RANGE, "LUT too large");
  const int ncell = (int)ncell64;
  // The S_p(g-h) table covers every difference of two list entries. The LUT box is the
  // caller's: the reference's list can hold entries beyond half of it (hexagonal axes), and
  // the assembly kernel reads the LUT at the reference's column-major offsets, checked
  // against the allocation (CTEM_E_RANGE when a lookup leaves it).
  int hmax[3] = {0, 0, 0};
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) hmax[d] = std::max(hmax[d], std::abs(x->hkl[3 * i + d]));
  int32_t slo[3], sdim[3];
  for (int d = 0; d < 3; d++) { slo[d] = -2 * hmax[d]; sdim[d] = 4 * hmax[d] + 1; }
  const int64_t sncell = (int64_t)sdim[0] * sdim[1] * sdim[2];
  if (sncell * std::max(1, x->nsite) > (1ll << 31)) return fail(CTEM_E_RANGE, "reflection list spans too large a box");
  std::vector<float> g(3 * n), cut1(n), cut2(n);
  std::vector<uint8_t> dbd(n);
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) g[3 * i + d] = (float)x->hkl[3 * i + d];
    // lUg = abs(Ucg)*mLambda stored in float32; cut1/cut2 in float32 (gvectors.f90:1029-1032)
    const float lug = (float)(std::abs(std::complex<double>(x->ucg[2 * i], x->ucg[2 * i + 1])) * x->mlambda);
    cut1[i] = x->cutoff * lug;
    cut2[i] = x->weakcutoff * lug;
    dbd[i] = x->dbdiff[i] ? 1 : 0;
  }
  std::vector<float> lutf(2 * (size_t)ncell);
  float lutmax = 0.0f;
  for (size_t i = 0; i < 2 * (size_t)ncell; i++) {
    lutf[i] = (float)x->lut[i];
    lutmax = std::max(lutmax, std::fabs(lutf[i]));
  }
  // inversion symmetry of the LUT: a box symmetric about the origin and |U(-g) - U(g)| small
  bool centro = true;
  for (int d = 0; d < 3; d++) centro = centro && (x->lut_lo[d] == -(x->lut_lo[d] + x->lut_dim[d] - 1));
  if (centro) {
    double lmax = 0.0, dmax = 0.0;
    for (int64_t i = 0; i < ncell; i++) {
      const int64_t j = ncell - 1 - i;  // the cell of (-h, -k, -l) in a box symmetric about 0
      lmax = std::max(lmax, std::max(std::fabs(x->lut[2 * i]), std::fabs(x->lut[2 * i + 1])));
      dmax = std::max(dmax, std::max(std::fabs(x->lut[2 * i] - x->lut[2 * j]), std::fabs(x->lut[2 * i + 1] - x->lut[2 * j + 1])));
    }
    centro = !(dmax > std::ldexp(lmax, -20));  // NaN entries: not symmetric
  }
  // the S table's key: every input of build_slut
  std::vector<unsigned char> key;
  auto put = [&key](const void* v, size_t b) {
    key.insert(key.end(), (const unsigned char*)v, (const unsigned char*)v + b);
  };
  put(slo, sizeof(slo));
  put(sdim, sizeof(sdim));
  put(x->rmt, 9 * sizeof(double));
  put(&x->nsite, sizeof(x->nsite));
  size_t norb = 0;
  if (x->nsite > 0) {
    put(x->site_z, x->nsite * sizeof(int32_t));
    put(x->site_b, x->nsite * sizeof(float));
    put(x->site_norbit, x->nsite * sizeof(int32_t));
    for (int p = 0; p < x->nsite; p++) norb += (size_t)x->site_norbit[p];
    put(x->orbit, 3 * norb * sizeof(double));
  }
  std::shared_ptr<SiteTable> st = c->last_st.lock();
  if (!st || st->device != c->device || st->key != key) {
    std::vector<double> slut, dslut;
    build_slut(x, slo, sdim, slut, dslut);
    st = std::make_shared<SiteTable>();
    st->device = c->device;
    st->key = std::move(key);
    int32_t rs;
    if ((rs = upload(st->slut, slut.data(), slut.size(), c->stream))) return rs;
    if ((rs = upload(st->dslut, dslut.data(), dslut.size(), c->stream))) return rs;
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  int nat = 0;
  for (int p = 0; p < x->nsite; p++) nat += x->site_norbit[p];
  int32_t rc;
  auto cr = std::make_shared<Crystal>();
  cr->device = c->device;
  if ((rc = upload(cr->g, g.data(), g.size(), c->stream))) return rc;
  if ((rc = upload(cr->hkl, x->hkl, (size_t)3 * n, c->stream))) return rc;
  if ((rc = upload(cr->cut1, cut1.data(), n, c->stream))) return rc;
  if ((rc = upload(cr->cut2, cut2.data(), n, c->stream))) return rc;
  if ((rc = upload(cr->dbd, dbd.data(), n, c->stream))) return rc;
  if ((rc = upload(cr->lut, lutf.data(), lutf.size(), c->stream))) return rc;
  cr->st = st;
  if ((rc = upload(cr->rmt, x->rmt, 9, c->stream))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  cr->nref = n;
  cr->nsite = x->nsite;
  cr->ncell = ncell;
  for (int i = 0; i < 3; i++) {
    cr->lut_lo[i] = x->lut_lo[i]; cr->lut_dim[i] = x->lut_dim[i];
    cr->slut_lo[i] = slo[i]; cr->slut_dim[i] = sdim[i];
  }
  cr->sncell = sncell;
  c->last_st = st;
  cr->lutmax = lutmax;
  cr->centro = centro;
  cr->mlambda = x->mlambda;
  cr->dynupz = x->dynupz;
  cr->sgcutoff = x->sgcutoff;
  cr->nat = (float)nat;
  cr->h_hkl.assign(x->hkl, x->hkl + 3 * n);
  if (x->nsite > 0) {
    cr->h_site_z.assign(x->site_z, x->site_z + x->nsite);
    cr->h_site_b.assign(x->site_b, x->site_b + x->nsite);
    cr->h_norbit.assign(x->site_norbit, x->site_norbit + x->nsite);
    cr->h_orbit.assign(x->orbit, x->orbit + 3 * (size_t)nat);
  }
  for (int i = 0; i < 9; i++) cr->h_rmt[i] = x->rmt[i];
  c->cr = cr;
  return CTEM_OK;
}

// ------------------------------------------------------------------------------------------
// incident directions (kvectors.f90:421-542, 781-947)
namespace {
struct KList {
  int32_t cap, n;
  int32_t* ij; double* k; double* kn;
  const double* dsm; double mlambda; int hemi;
};

void push_k(KList& L, int i, int j, const double ks_in[3]) {
  if (L.n < L.cap && L.ij) {
    double ks[3] = {ks_in[0], ks_in[1], L.hemi < 0 ? -ks_in[2] : ks_in[2]};
    const double x = std::sqrt(ks[0] * ks[0] + ks[1] * ks[1] + ks[2] * ks[2]);
    if (x != 0.0) for (int t = 0; t < 3; t++) ks[t] = ks[t] / x;
    for (int t = 0; t < 3; t++) ks[t] = ks[t] / L.mlambda;
    for (int a = 0; a < 3; a++) {
      double s = 0.0;
      for (int b = 0; b < 3; b++) s += L.dsm[b * 3 + a] * ks[b];
      L.k[3 * L.n + a] = s;
    }
    L.ij[2 * L.n] = i;
    L.ij[2 * L.n + 1] = j;
    L.kn[L.n] = 1.0 / L.mlambda;
  }
  L.n++;
}

// The reference's compiled code evaluates sin and cos of one argument through a single
// sincos call, which differs from separate sin and cos calls in the last bit for some
// arguments; the same call is used here.
inline void sincos_ref(double a, double* s, double* c) { ::sincos(a, s, c); }

int sextant(double x, double y) {
  const double srt = (double)1.732050808f;
  const double xx = std::fabs(x * srt);
  if (y >= 0) { if (y >= xx) return 0; return (x > 0.0) ? 1 : 5; }
  if (std::fabs(y) >= xx) return 3;
  return (x > 0.0) ? 2 : 4;
}

void add_k(KList& L, double delta, int i, int j, int usehex) {
#pragma clang fp contract(off)
  const double srt = 0.86602540, rtt = 1.7320508076, prea = 0.525037568, preb = 1.050075136,
               prec = 0.90689968, pred = 2.09439510, ipi = 1.0 / kPi;
  if (i == 0 && j == 0) return;
  double x, y, ks[3];
  if (usehex) {
    const float jh = (float)j * 0.5f;
    x = (double)((float)i - jh) * delta;
    y = (double)j * delta * srt;
    double XX, YY, xp, yp, sn, cs;
    switch (sextant(x, y)) {
      case 0: case 3: sincos_ref(x * prec / y, &sn, &cs); XX = preb * y * cs; YY = preb * y * sn; break;
      case 1: case 4: xp = y + rtt * x; yp = y * pred / xp; sincos_ref(yp, &sn, &cs); XX = prea * xp * sn; YY = prea * xp * cs; break;
      default: xp = y - rtt * x; yp = y * pred / xp; sincos_ref(yp, &sn, &cs); XX = prea * xp * sn; YY = -prea * xp * cs; break;
    }
    const double q = XX * XX + YY * YY;
    ks[0] = 0.5 * XX * std::sqrt(4.0 - q);
    ks[1] = 0.5 * YY * std::sqrt(4.0 - q);
    ks[2] = 1.0 - 0.5 * q;
    push_k(L, i - j / 2 + (j % 2) / 2, j, ks);
  } else {
    x = i * delta;
    y = j * delta;
    double sn, cs;
    if (std::fabs(x) <= std::fabs(y)) {
      const double q = 2.0 * y * ipi * std::sqrt(kPi - y * y);
      sincos_ref(x * kPi * 0.25 / y, &sn, &cs);
      ks[0] = q * sn; ks[1] = q * cs; ks[2] = 1.0 - 2.0 * y * y * ipi;
    } else {
      const double q = 2.0 * x * ipi * std::sqrt(kPi - x * x);
      sincos_ref(y * kPi * 0.25 / x, &sn, &cs);
      ks[0] = q * cs; ks[1] = q * sn; ks[2] = 1.0 - 2.0 * x * x * ipi;
    }
    push_k(L, i, j, ks);
  }
}
}  // namespace

extern "C" int32_t ctem_kvectors(int32_t npx, int32_t npy, int32_t isym, int32_t usehex,
                                 int32_t hemisphere, const double dsm[9], double mlambda,
                                 int32_t cap, int32_t* ij, double* k, double* kn, int32_t* numk) {
  if (npx < 1 || npy < 1 || !dsm || !numk || !(mlambda > 0) || (hemisphere != 1 && hemisphere != -1))
    return fail(CTEM_E_ARG, "ctem_kvectors: bad arguments");
  if ((ij || k || kn) && !(ij && k && kn)) return fail(CTEM_E_ARG, "ctem_kvectors: ij, k, kn must all be given");
  KList L{ij ? cap : 0, 0, ij, k, kn, dsm, mlambda, hemisphere};
  const double delta = usehex ? 2.0 * std::sqrt(kPi) / std::pow(3.0, 0.75) / (double)npx
                              : std::sqrt(kPi * 0.5) / (double)npx;
  const double head[3] = {0.0, 0.0, 1.0};
  push_k(L, 0, 0, head);
  switch (isym) {
    case 1: for (int j = -npy; j <= npy; j++) for (int i = -npx; i <= npx; i++) add_k(L, delta, i, j, usehex); break;
    case 2: for (int j = 0; j <= npy; j++) for (int i = -npx; i <= npx; i++) add_k(L, delta, i, j, usehex); break;
    case 3: case 4: case 10: for (int j = 0; j <= npy; j++) for (int i = 0; i <= npx; i++) add_k(L, delta, i, j, usehex); break;
    case 5: case 11: for (int i = 0; i <= npx; i++) for (int j = 0; j <= i; j++) add_k(L, delta, i, j, usehex); break;
    case 6: case 7: for (int j = 0; j <= npy; j++) for (int i = 0; i <= npy; i++) add_k(L, delta, i, j, usehex); break;
    case 8: for (int j = 0; j <= npy; j++) for (int i = j; i <= npy; i++) add_k(L, delta, i, j, usehex); break;
    case 9: for (int j = 0; j <= npx; j++) for (int i = 2 * j; i <= npx; i++) add_k(L, delta, i, j, usehex); break;
    case 12:
      for (int j = 0; j <= npy / 2; j++)
        for (int i = 2 * j; i <= npy; i++) { add_k(L, delta, i, j, usehex); add_k(L, delta, i - j, -j, usehex); }
      break;
    default: return fail(CTEM_E_ARG, "ctem_kvectors: Laue group %d not in 1..12", isym);
  }
  *numk = L.n;
  if (ij && L.n > cap) return fail(CTEM_E_ARG, "ctem_kvectors: capacity %d < %d directions", cap, L.n);
  return CTEM_OK;
}

extern "C" int32_t ctem_laue_equiv(int32_t ipx, int32_t ipy, int32_t isym, int32_t* iequiv, int32_t* nequiv) {
  if (!iequiv || !nequiv) return fail(CTEM_E_ARG, "ctem_laue_equiv: null output");
  int32_t q[24] = {0};
  const int ne = laue_equiv(ipx, ipy, isym, q);
  memcpy(iequiv, q, sizeof q);
  *nequiv = ne;
  return CTEM_OK;
}

// ------------------------------------------------------------------------------------------
// plans
extern "C" int32_t ctem_plan_create(ctem_ctx* c, int32_t numk, const double* k, const double* kn,
                                    int32_t nz, const float* lam, double dz, double thick,
                                    ctem_plan** out) {
  if (!c || !out) return fail(CTEM_E_ARG, "ctem_plan_create: null argument");
  *out = nullptr;
  if (numk < 0 || (numk > 0 && (!k || !kn))) return fail(CTEM_E_ARG, "ctem_plan_create: bad direction list");
  if (nz < 1 || nz > 4096 || !lam) return fail(CTEM_E_ARG, "ctem_plan_create: need 1..4096 depth weights");
  if (!(dz > 0) || !(thick > 0)) return fail(CTEM_E_ARG, "ctem_plan_create: dz and thick must be positive");
  for (int64_t i = 0; i < 3ll * numk; i++)
    if (!std::isfinite(k[i])) return fail(CTEM_E_ARG, "ctem_plan_create: non-finite wave vector");
  HIPCHK(hipSetDevice(c->device));
  ctem_plan* p = new ctem_plan();
  p->ctx = c;
  p->cr = c->cr;
  p->serial = ++c->plan_serial;
  p->numk = numk; p->nz = nz; p->dz = dz; p->thick = thick;
  std::vector<double> lamd(lam, lam + nz);
  int32_t rc;
  if ((rc = upload(p->k, k, (size_t)3 * numk, c->stream)) || (rc = upload(p->kn, kn, (size_t)numk, c->stream)) ||
      (rc = upload(p->lam, lamd.data(), (size_t)nz, c->stream))) {
    delete p;
    return rc;
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) { delete p; return fail(CTEM_E_HIP, "plan upload failed"); }
  c->plans.push_back(p);
  *out = p;
  return CTEM_OK;
}

extern "C" int32_t ctem_plan_destroy(ctem_plan* p) {
  if (!p) return CTEM_OK;
  (void)hipSetDevice(p->ctx->device);
  (void)hipStreamSynchronize(p->ctx->stream);
  if (p->graph) (void)hipGraphExecDestroy(p->graph);
  if (p->ctx->mgraph) {  // a graph over a plan list may hold this plan's buffers
    (void)hipGraphExecDestroy(p->ctx->mgraph);
    p->ctx->mgraph = nullptr;
    p->ctx->mgraph_key.clear();
  }
  for (DBuf* b : {&p->k, &p->kn, &p->lam, &p->ij, &p->nns, &p->nnw, &p->soff, &p->woff, &p->sidx,
                  &p->ssg, &p->widx, &p->wsg, &p->err, &p->dots, &p->keep_lgh, &p->dirs, &p->fn, &p->ecp_slut,
                  &p->ecp_store, &p->ecp_thick, &p->ecp_sr, &p->ecp_part})
    b->release();
  p->cr.reset();
  for (auto e : p->evpool) (void)hipEventDestroy(e);
  auto& v = p->ctx->plans;
  v.erase(std::remove(v.begin(), v.end(), p), v.end());
  delete p;
  return CTEM_OK;
}

static int32_t check_err_flag(DBuf& err, hipStream_t s, const char* what) {
  int32_t h = 0;
  HIPCHK(hipMemcpyAsync(&h, err.p, sizeof h, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (h & 32) return fail(CTEM_E_STATE, "%s: a matrix needs more squarings than the cached run schedule", what);
  if (h & 1) return fail(CTEM_E_RANGE, "%s: an excitation-error angle left the reproduced cosf range (|g| too large for k)", what);
  if (h & 2) return fail(CTEM_E_NUMERIC, "%s: non-finite dynamical matrix", what);
  if (h & 4) return fail(CTEM_E_NUMERIC, "%s: matrix norm needs more squarings than allowed", what);
  if (h & 16)
    return fail(CTEM_E_NUMERIC, "%s: a matrix entry left the split-FP16 range (|x| >= 16); "
                "use CTEM_SPLIT_BF16X6 for this crystal", what);
  if (h & 8)
    return fail(CTEM_E_RANGE, "%s: a beam pair's g-h lies outside the LUT allocation (the reference's "
                "Compute_DynMat reads outside its LUT array there); pass a larger LUT box", what);
  return CTEM_OK;
}

extern "C" int32_t ctem_plan_select(ctem_plan* p, int32_t* nns_out, int32_t* nnw_out) {
  if (!p) return fail(CTEM_E_ARG, "null plan");
  ctem_ctx* c = p->ctx;
  if (!p->cr) p->cr = c->cr;
  if (!p->cr) return fail(CTEM_E_STATE, "ctem_plan_select: no crystal uploaded");
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const int nk = p->numk;
  int32_t rc;
  if ((rc = p->nns.ensure(4 * (size_t)std::max(nk, 1))) || (rc = p->nnw.ensure(4 * (size_t)std::max(nk, 1))) ||
      (rc = p->err.ensure(4)) || (rc = p->soff.ensure(8 * (size_t)std::max(nk, 1))) ||
      (rc = p->woff.ensure(8 * (size_t)std::max(nk, 1))))
    return rc;
  HIPCHK(hipMemsetAsync(p->err.p, 0, 4, s));
  SelectOut o{};
  o.nns = p->nns.as<int32_t>(); o.nnw = p->nnw.as<int32_t>(); o.err = p->err.as<int32_t>();
  DevRefl R = p->cr->refl();
  const float* fnp = p->fn.p ? p->fn.as<float>() : nullptr;
  launch_select(s, p->k.as<double>(), nk, R, p->cr->rmt.as<double>(), fnp, o, 0);
  HIPCHK(hipGetLastError());
  if ((rc = check_err_flag(p->err, s, "beam selection"))) return rc;
  p->h_nns.resize(nk); p->h_nnw.resize(nk);
  if (nk) {
    HIPCHK(hipMemcpyAsync(p->h_nns.data(), p->nns.p, 4 * (size_t)nk, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(p->h_nnw.data(), p->nnw.p, 4 * (size_t)nk, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  p->h_soff.resize(nk); p->h_woff.resize(nk);
  int64_t ts = 0, tw = 0;
  for (int i = 0; i < nk; i++) { p->h_soff[i] = ts; p->h_woff[i] = tw; ts += p->h_nns[i]; tw += p->h_nnw[i]; }
  if ((rc = upload(p->soff, p->h_soff.data(), nk, s)) || (rc = upload(p->woff, p->h_woff.data(), nk, s)) ||
      (rc = p->sidx.ensure(4 * (size_t)std::max<int64_t>(ts, 1))) || (rc = p->ssg.ensure(4 * (size_t)std::max<int64_t>(ts, 1))) ||
      (rc = p->widx.ensure(4 * (size_t)std::max<int64_t>(tw, 1))) || (rc = p->wsg.ensure(4 * (size_t)std::max<int64_t>(tw, 1))))
    return rc;
  o.soff = p->soff.as<int64_t>(); o.woff = p->woff.as<int64_t>();
  o.sidx = p->sidx.as<int32_t>(); o.ssg = p->ssg.as<float>();
  o.widx = p->widx.as<int32_t>(); o.wsg = p->wsg.as<float>();
  launch_select(s, p->k.as<double>(), nk, R, p->cr->rmt.as<double>(), fnp, o, 1);
  HIPCHK(hipGetLastError());
  // bucket layout: directions grouped by padded size, in list order within a bucket
  {
    std::map<int, std::vector<int32_t>> bk;
    for (int i = 0; i < nk; i++) bk[npad_of(p->h_nns[i], p->bucket_multiple)].push_back(i);
    p->buckets.assign(bk.begin(), bk.end());
    std::vector<int32_t> order;
    order.reserve(nk);
    for (auto& kv : p->buckets) order.insert(order.end(), kv.second.begin(), kv.second.end());
    if ((rc = upload(p->dirs, order.data(), order.size(), s))) return rc;
  }
  HIPCHK(hipStreamSynchronize(s));
  const void* bufs[] = {p->nns.p, p->nnw.p, p->soff.p, p->woff.p, p->sidx.p, p->ssg.p, p->widx.p, p->wsg.p,
                        p->dirs.p, p->err.p};
  uint64_t h = 1469598103934665603ull;
  for (const void* q : bufs) h = (h ^ (uint64_t)(uintptr_t)q) * 1099511628211ull;
  p->bufsig = h;
  p->selected = true;
  if (nns_out) memcpy(nns_out, p->h_nns.data(), 4 * (size_t)nk);
  if (nnw_out) memcpy(nnw_out, p->h_nnw.data(), 4 * (size_t)nk);
  return CTEM_OK;
}

extern "C" int32_t ctem_plan_get_beams(ctem_plan* p, int32_t ik, int32_t* strong, float* strongsg,
                                       int32_t* weak, float* weaksg) {
  if (!p || !p->selected) return fail(CTEM_E_STATE, "ctem_plan_get_beams: run ctem_plan_select first");
  if (ik < 0 || ik >= p->numk) return fail(CTEM_E_ARG, "direction %d out of range", ik);
  hipStream_t s = p->ctx->stream;
  const size_t ns = p->h_nns[ik], nw = p->h_nnw[ik];
  if (strong) HIPCHK(hipMemcpyAsync(strong, p->sidx.as<int32_t>() + p->h_soff[ik], 4 * ns, hipMemcpyDeviceToHost, s));
  if (strongsg) HIPCHK(hipMemcpyAsync(strongsg, p->ssg.as<float>() + p->h_soff[ik], 4 * ns, hipMemcpyDeviceToHost, s));
  if (weak && nw) HIPCHK(hipMemcpyAsync(weak, p->widx.as<int32_t>() + p->h_woff[ik], 4 * nw, hipMemcpyDeviceToHost, s));
  if (weaksg && nw) HIPCHK(hipMemcpyAsync(weaksg, p->wsg.as<float>() + p->h_woff[ik], 4 * nw, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return CTEM_OK;
}

// ------------------------------------------------------------------------------------------
// the matrix exponential and depth stage, FP64
namespace {

struct GemmTimer {
  ctem_plan* p;  // may be null (no timing)
  hipStream_t s;
  hipEvent_t a = nullptr, b = nullptr;
  void begin() {
    if (!p) return;
    if (p->evused + 2 > p->evpool.size()) {
      hipEvent_t e1, e2;
      (void)hipEventCreate(&e1); (void)hipEventCreate(&e2);
      p->evpool.push_back(e1); p->evpool.push_back(e2);
    }
    a = p->evpool[p->evused++]; b = p->evpool[p->evused++];
    (void)hipEventRecord(a, s);
  }
  void end() { if (p) (void)hipEventRecord(b, s); }
};

constexpr double kTheta64 = 0.9;   // ||A/2^s||_1 bound for the degree-16 Taylor polynomial
constexpr int kSmax = 48;

// A (nb matrices, npad^2 each) holds 2 pi i dz M' on entry and is scaled by 2^-s in place.
// W holds 5*nb*npad^2 work space; on exit E_b = expm(A_b) is in the fifth slice of W (the
// returned pointer). h_n[b] are the true sizes (for the algorithmic flop count), h_sq
// receives the squaring counts.
// cs (may be null): with cs->known the squaring counts of this chunk are taken from it and not
// read back (no host synchronisation); otherwise they are read back and stored in it.
double2* expm_fp64(ctem_plan* tp, hipStream_t s, int npad, int nb, double2* A, double2* W,
                   int32_t* d_sq, int32_t* d_err, const std::vector<int>& h_n,
                   std::vector<int32_t>& h_sq, int32_t& rc_out, double* d_amod = nullptr,
                   std::vector<double>* h_amod = nullptr, ChunkSched* cs = nullptr) {
  rc_out = CTEM_OK;
  const int64_t per = (int64_t)npad * npad;
  double2* A2 = W;
  double2* A3 = W + per * nb;
  double2* A4 = W + 2 * per * nb;
  double2* H = W + 3 * per * nb;
  double2* T = W + 4 * per * nb;
#define EXPM_CHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { rc_out = fail(CTEM_E_HIP, "%s: %s", #x, hipGetErrorString(e_)); return nullptr; } } while (0)
  launch_norm_scale(s, npad, nb, A, kTheta64, kSmax, d_sq, d_err, d_amod);
  EXPM_CHK(hipGetLastError());
  if (cs && cs->known) {
    h_sq = cs->sq;
  } else {
    h_sq.resize(nb);
    EXPM_CHK(hipMemcpyAsync(h_sq.data(), d_sq, 4 * (size_t)nb, hipMemcpyDeviceToHost, s));
    if (d_amod && h_amod) {
      h_amod->resize(nb);
      EXPM_CHK(hipMemcpyAsync(h_amod->data(), d_amod, 8 * (size_t)nb, hipMemcpyDeviceToHost, s));
    }
    EXPM_CHK(hipStreamSynchronize(s));
  }
  int smax = 0;
  for (int b = 0; b < nb; b++) smax = std::max(smax, h_sq[b]);
  if (cs && !cs->known) { cs->sq = h_sq; cs->smax = smax; cs->known = true; }
  if (cs) launch_sq_check(s, nb, d_sq, smax, d_err);
  double n3 = 0.0;
  for (int b = 0; b < nb; b++) n3 += (double)h_n[b] * h_n[b] * h_n[b];

  auto sq_gemm = [&](double2* C, const double2* X, const double2* Y, double betaI,
                     std::initializer_list<std::pair<const double2*, double>> ms,
                     const int32_t* active, int step, double flops) -> int32_t {
    ZGemm g{};
    g.m = npad; g.n = npad; g.k = npad;
    g.ldx = g.ldy = g.ldc = g.ldm = npad;
    g.X = X; g.sx = per; g.Y = Y; g.sy = per; g.C = C; g.sc = per; g.sm = per;
    g.alpha = 1.0; g.betaI = betaI;
    int q = 0;
    for (auto& m : ms) { g.M[q] = m.first; g.beta[q] = m.second; q++; }
    g.active = active; g.step = step; g.nb = nb;
    GemmTimer t{tp, s};
    t.begin();
    launch_zgemm(s, g);
    t.end();
    HIPCHK(hipGetLastError());
    if (tp) { tp->f_gemm += flops; tp->n_gemm++; }
    return CTEM_OK;
  };
  int32_t rc;
  const double f3 = 8.0 * n3;
  if ((rc = sq_gemm(A2, A, A, 0.0, {}, nullptr, 0, f3))) { rc_out = rc; return nullptr; }
  if ((rc = sq_gemm(A3, A, A2, 0.0, {}, nullptr, 0, f3))) { rc_out = rc; return nullptr; }
  if ((rc = sq_gemm(A4, A2, A2, 0.0, {}, nullptr, 0, f3))) { rc_out = rc; return nullptr; }
  {
    const double2* M[4] = {A, A2, A3, A4};
    const double beta[4] = {inv_fact(13), inv_fact(14), inv_fact(15), inv_fact(16)};
    launch_combo(s, npad, nb, H, inv_fact(12), M, beta);
    EXPM_CHK(hipGetLastError());
  }
  if ((rc = sq_gemm(T, A4, H, inv_fact(8), {{A, inv_fact(9)}, {A2, inv_fact(10)}, {A3, inv_fact(11)}}, nullptr, 0, f3))) { rc_out = rc; return nullptr; }
  if ((rc = sq_gemm(H, A4, T, inv_fact(4), {{A, inv_fact(5)}, {A2, inv_fact(6)}, {A3, inv_fact(7)}}, nullptr, 0, f3))) { rc_out = rc; return nullptr; }
  if ((rc = sq_gemm(T, A4, H, 1.0, {{A, 1.0}, {A2, inv_fact(2)}, {A3, inv_fact(3)}}, nullptr, 0, f3))) { rc_out = rc; return nullptr; }
  // squarings with per-matrix counts; the result ends in T after the parity copy
  double2* buf[2] = {T, H};
  for (int i = 0; i < smax; i++) {
    double fl = 0.0;
    for (int b = 0; b < nb; b++) if (i < h_sq[b]) fl += 8.0 * h_n[b] * h_n[b] * h_n[b];
    if ((rc = sq_gemm(buf[(i + 1) & 1], buf[i & 1], buf[i & 1], 0.0, {}, d_sq, i, fl))) { rc_out = rc; return nullptr; }
  }
  launch_parity_copy(s, npad, nb, d_sq, H, T);
  EXPM_CHK(hipGetLastError());
#undef EXPM_CHK
  return T;
}

// expm of the block matrix [[X, Y], [0, X]] held as the pair (X, Y) (npad x npad each): the same
// degree-16 Taylor evaluation and squarings as expm_fp64, every product of two pairs
// (X1 X2, X1 Y2 + Y1 X2) taken as three n-size GEMMs (the dense 2n-size product is eight). X and
// Y are scaled in place; W holds 10 nb npad^2 entries. Returns the pair (E, L) in W: E at the
// returned pointer, L at E + 5 nb npad^2 (L = the Frechet derivative of expm at X in the direction
// Y when X, Y are unscaled on entry).
double2* expm_pair_fp64(hipStream_t s, int npad, int nb, double2* X, double2* Y, double2* W, int32_t* d_sq,
                        int32_t* d_err, std::vector<int32_t>& h_sq, int32_t& rc_out) {
  rc_out = CTEM_OK;
  const int64_t per = (int64_t)npad * npad;
  const int64_t half = 5 * per * nb;  // X parts in W[0, half), Y parts in W[half, 2 half)
  double2* X2 = W;
  double2* X3 = W + per * nb;
  double2* X4 = W + 2 * per * nb;
  double2* XH = W + 3 * per * nb;
  double2* XT = W + 4 * per * nb;
  double2* Y2 = X2 + half;
  double2* Y3 = X3 + half;
  double2* Y4 = X4 + half;
  double2* YH = XH + half;
  double2* YT = XT + half;
#define PAIR_CHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { rc_out = fail(CTEM_E_HIP, "%s: %s", #x, hipGetErrorString(e_)); return nullptr; } } while (0)
  launch_norm_scale_pair(s, npad, nb, X, Y, kTheta64, kSmax, d_sq, d_err);
  PAIR_CHK(hipGetLastError());
  h_sq.resize(nb);
  PAIR_CHK(hipMemcpyAsync(h_sq.data(), d_sq, 4 * (size_t)nb, hipMemcpyDeviceToHost, s));
  PAIR_CHK(hipStreamSynchronize(s));
  int smax = 0;
  for (int b = 0; b < nb; b++) smax = std::max(smax, h_sq[b]);
  // C = X*Y + betaI I + sum beta_j M_j over the nb matrices (M may be C itself: the epilogue reads
  // each addend element before it writes that element of C)
  auto gemm = [&](double2* C, const double2* A_, const double2* B_, double betaI,
                  std::initializer_list<std::pair<const double2*, double>> ms, const int32_t* active,
                  int step) -> bool {
    ZGemm g{};
    g.m = g.n = g.k = npad;
    g.ldx = g.ldy = g.ldc = g.ldm = npad;
    g.X = A_; g.sx = per; g.Y = B_; g.sy = per; g.C = C; g.sc = per; g.sm = per;
    g.alpha = 1.0; g.betaI = betaI;
    int q = 0;
    for (auto& m : ms) { g.M[q] = m.first; g.beta[q] = m.second; q++; }
    g.active = active; g.step = step; g.nb = nb;
    launch_zgemm(s, g);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { rc_out = fail(CTEM_E_HIP, "pair product: %s", hipGetErrorString(e)); return false; }
    return true;
  };
  // (CX, CY) = (AX, AY)(BX, BY) + betaI (I, 0) + sum beta_j (MX_j, MY_j)
  auto pmul = [&](double2* CX, double2* CY, const double2* AX, const double2* AY, const double2* BX,
                  const double2* BY, double betaI, std::initializer_list<std::pair<const double2*, double>> mx,
                  std::initializer_list<std::pair<const double2*, double>> my, const int32_t* active,
                  int step) -> bool {
    if (!gemm(CX, AX, BX, betaI, mx, active, step)) return false;
    if (!gemm(CY, AX, BY, 0.0, my, active, step)) return false;
    return gemm(CY, AY, BX, 0.0, {{CY, 1.0}}, active, step);
  };
  if (!pmul(X2, Y2, X, Y, X, Y, 0.0, {}, {}, nullptr, 0)) return nullptr;
  if (!pmul(X3, Y3, X, Y, X2, Y2, 0.0, {}, {}, nullptr, 0)) return nullptr;
  if (!pmul(X4, Y4, X2, Y2, X2, Y2, 0.0, {}, {}, nullptr, 0)) return nullptr;
  {
    const double beta[4] = {inv_fact(13), inv_fact(14), inv_fact(15), inv_fact(16)};
    const double2* MX[4] = {X, X2, X3, X4};
    const double2* MY[4] = {Y, Y2, Y3, Y4};
    launch_combo(s, npad, nb, XH, inv_fact(12), MX, beta);
    launch_combo(s, npad, nb, YH, 0.0, MY, beta);
    PAIR_CHK(hipGetLastError());
  }
  if (!pmul(XT, YT, X4, Y4, XH, YH, inv_fact(8), {{X, inv_fact(9)}, {X2, inv_fact(10)}, {X3, inv_fact(11)}},
            {{Y, inv_fact(9)}, {Y2, inv_fact(10)}, {Y3, inv_fact(11)}}, nullptr, 0))
    return nullptr;
  if (!pmul(XH, YH, X4, Y4, XT, YT, inv_fact(4), {{X, inv_fact(5)}, {X2, inv_fact(6)}, {X3, inv_fact(7)}},
            {{Y, inv_fact(5)}, {Y2, inv_fact(6)}, {Y3, inv_fact(7)}}, nullptr, 0))
    return nullptr;
  if (!pmul(XT, YT, X4, Y4, XH, YH, 1.0, {{X, 1.0}, {X2, inv_fact(2)}, {X3, inv_fact(3)}},
            {{Y, 1.0}, {Y2, inv_fact(2)}, {Y3, inv_fact(3)}}, nullptr, 0))
    return nullptr;
  double2* bx[2] = {XT, XH};
  double2* by[2] = {YT, YH};
  for (int i = 0; i < smax; i++)
    if (!pmul(bx[(i + 1) & 1], by[(i + 1) & 1], bx[i & 1], by[i & 1], bx[i & 1], by[i & 1], 0.0, {}, {}, d_sq, i))
      return nullptr;
  launch_parity_copy(s, npad, nb, d_sq, XH, XT);
  launch_parity_copy(s, npad, nb, d_sq, YH, YT);
  PAIR_CHK(hipGetLastError());
#undef PAIR_CHK
  return XT;
}

// FP64 form: expm, then psi_k = E^k e_1 by the depth recursion
int32_t expm_depth_fp64(ctem_plan* tp, hipStream_t s, int npad, int nb, double2* A, double2* W,
                        double2* Psi, int64_t spsi, int nz, int32_t* d_sq, int32_t* d_err,
                        const std::vector<int>& h_n, std::vector<int32_t>& h_sq, int depth_grid,
                        double* d_amod = nullptr, std::vector<double>* h_amod = nullptr,
                        ChunkSched* cs = nullptr) {
  int32_t rc;
  double2* E = expm_fp64(tp, s, npad, nb, A, W, d_sq, d_err, h_n, h_sq, rc, d_amod, h_amod, cs);
  if (!E) return rc;
  launch_depth(s, npad, nb, nz, E, 0, 0, Psi, spsi, depth_grid);
  HIPCHK(hipGetLastError());
  return CTEM_OK;
}

// Mixed precision: the same scaling-and-squaring with the exponential kept as E = I + F, so
// that every product acts on the small part F and its rounding error scales with |F|^2:
//   F_0 = the degree-8 Taylor polynomial of exp(B) - I,  B = A / 2^s, ||B||_1 <= theta
//   F_{i+1} = F_i^2 + 2 F_i            (E_{i+1} = E_i^2)
// The depth powers use the same form, psi_{m+j} = psi_j + F_m psi_j with F_m = E^m - I.
// Products run on split-FP16 (or BF16) MFMA with FP32 accumulation (ctem_mixed.hip); the
// depth recursion accumulates in FP64 and the final Lgh/Sgh contraction runs in FP64.
// theta: 1/2 for the split-BF16 options (truncation theta^9/9! = 5.4e-9). 5/4 for split-FP16
// storage: one squaring fewer than theta = 3/4 for about three quarters of the matrices (all of
// bcc Fe at 20 kV, whose norms lie between 31 and 37: 8 products instead of 9). The truncation's
// effect on the intensities, measured in FP64 arithmetic against the exact exponential over
// sampled directions, stays below 5e-6 relative (bcc Fe at 10 and 20 kV, fcc Ni, corundum;
// 2e-9 at theta = 3/4, up to 6e-6 at 3/2 and 1.2e-3 at 3), the size of the split products' own
// rounding error and far below the mixed form's 1e-4 bar.
constexpr double kThetaMixed = 0.5;
constexpr double kThetaSplit = 1.25;
// x1..x7, y2 of the three-product degree-8 Taylor evaluation (the structure of Bader, Blanes and
// Casas, "Computing the matrix exponential with an optimized Taylor polynomial approximation",
// Mathematics 7 (2019) 1174), solved here for the coefficients of B^2..B^8 of exp(B):
// with x1 = 4 x2, B8 matches B^3/6 + ... + B^8/8! (coefficient error < 3e-17)
constexpr float kBbc8[8] = {0.098979145598294002f, 0.024744786399573501f, 0.60892636975088432f,
                            0.59860708413115948f,  0.17640400262879624f,  0.015427061856584176f,
                            0.040505319112355415f, 0.13549236135285087f};

int32_t expm_depth_mixed(ctem_plan* tp, hipStream_t s, int npad, int nb, float2* W32,
                         float* wmax, double2* Psi64, int64_t spsi, int nz, int32_t* d_sq,
                         int32_t* d_err, const std::vector<int>& h_n, std::vector<int32_t>& h_sq,
                         int terms_taylor, int terms_square, int depth_grid, ChunkSched* cs) {
  const int64_t per = (int64_t)npad * npad;
  float2* B = W32;
  float2* G2 = W32 + per * nb;
  float2* G3 = W32 + 2 * per * nb;
  float2* P = W32 + 3 * per * nb;
  float2* F = W32 + 4 * per * nb;
  float2* H = W32 + 5 * per * nb;
  // per-matrix max |entry| of each buffer (split-FP16 operand scaling)
  float* mB = wmax;
  float* mG2 = wmax + nb;
  float* mG3 = wmax + 2 * nb;
  float* mP = wmax + 3 * nb;
  float* mF = wmax + 4 * nb;
  float* mH = wmax + 5 * nb;
  launch_fill_zero(s, wmax, 6 * sizeof(float) * (size_t)nb);
  // B (the first buffer) holds A in FP32 as assembled; scale it by 2^-s in place
  launch_norm_scale32(s, npad, nb, B, kThetaMixed, kSmax, d_sq, d_err, mB);
  HIPCHK(hipGetLastError());
  if (cs->known) {
    h_sq = cs->sq;
  } else {
    h_sq.resize(nb);
    HIPCHK(hipMemcpyAsync(h_sq.data(), d_sq, 4 * (size_t)nb, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  int smax = 0;
  for (int b = 0; b < nb; b++) smax = std::max(smax, h_sq[b]);
  if (!cs->known) { cs->sq = h_sq; cs->smax = smax; cs->known = true; }
  launch_sq_check(s, nb, d_sq, smax, d_err);
  double n3 = 0.0;
  for (int b = 0; b < nb; b++) n3 += (double)h_n[b] * h_n[b] * h_n[b];
  auto launch = [&](const CGemm& g, double flops) -> int32_t {
    GemmTimer t{tp, s};
    t.begin();
    launch_cgemm(s, g);
    t.end();
    HIPCHK(hipGetLastError());
    if (tp) { tp->f_gemm += flops; tp->n_gemm++; }
    return CTEM_OK;
  };
  // C = X*Y + betaI*I + sum beta_j M_j over nb square matrices
  auto gemm = [&](float2* C, float* cmax, const float2* X, const float* xmax, const float2* Y, const float* ymax,
                  float betaI, std::initializer_list<std::pair<const float2*, float>> ms, int terms) {
    CGemm g{};
    g.m = g.n = g.k = npad;
    g.ldx = g.ldy = g.ldc = g.ldm = npad;
    g.X = X; g.sx = per; g.Y = Y; g.sy = per; g.C = C; g.sc = per; g.sm = per;
    g.alpha = 1.0f; g.betaI = betaI;
    int q = 0;
    for (auto& m : ms) { g.M[q] = m.first; g.beta[q] = m.second; q++; }
    g.nb = nb; g.terms = terms; g.err = d_err;
    g.xmax = xmax; g.ymax = ymax; g.cmax = cmax;
    return g;
  };
  int32_t rc;
  const double f3 = 8.0 * n3;
  // degree-8 Taylor polynomial of exp(B) - I from three products
  //   B2 = B*B,  B4 = B2*(x1 B + x2 B2),
  //   B8 = (x3 B2 + B4)*(x4 I + x5 B + x6 B2 + x7 B4),
  //   F0 = B + y2 B2 + B8
  // each product's epilogue also writes the next operand (second output of the GEMM)
  {
    CGemm g = gemm(G2, mG2, B, mB, B, mB, 0.f, {{B, 0.f}}, terms_taylor);
    g.C2 = P; g.cmax2 = mP; g.alpha2 = kBbc8[1]; g.betaI2 = 0.f; g.beta2[0] = kBbc8[0];
    if ((rc = launch(g, f3))) return rc;
  }
  {
    CGemm g = gemm(G3, mG3, G2, mG2, P, mP, 0.f, {{G2, kBbc8[2]}, {B, 0.f}}, terms_taylor);
    g.C2 = H; g.cmax2 = mH; g.alpha2 = kBbc8[6]; g.betaI2 = kBbc8[3];
    g.beta2[0] = kBbc8[5]; g.beta2[1] = kBbc8[4];
    if ((rc = launch(g, f3))) return rc;
  }
  if ((rc = launch(gemm(F, mF, G3, mG3, H, mH, 0.f, {{B, 1.f}, {G2, kBbc8[7]}}, terms_taylor), f3))) return rc;
  float2* buf[2] = {F, H};
  float* bmax[2] = {mF, mH};
  for (int i = 0; i < smax; i++) {
    double fl = 0.0;
    for (int b = 0; b < nb; b++) if (i < h_sq[b]) fl += 8.0 * h_n[b] * h_n[b] * h_n[b];
    float2* X = buf[i & 1];
    float* out_max = bmax[(i + 1) & 1];
    // the output buffer's maxima were those of two steps back
    launch_fill_zero(s, out_max, sizeof(float) * (size_t)nb);
    CGemm g = gemm(buf[(i + 1) & 1], out_max, X, bmax[i & 1], X, bmax[i & 1], 0.f, {{X, 2.f}}, terms_square);
    g.active = d_sq; g.step = i;
    if ((rc = launch(g, fl))) return rc;
  }
  launch_cparity_copy(s, npad, nb, d_sq, H, F);
  HIPCHK(hipGetLastError());
  // depth on E = I + F with FP64 accumulation
  launch_depth(s, npad, nb, nz, F, 1, 1, Psi64, spsi, depth_grid);
  HIPCHK(hipGetLastError());
  return CTEM_OK;
}

// The mixed form with both stages on CTEM_SPLIT_FP16X3: the same Taylor polynomial and
// squarings as expm_depth_mixed, with every matrix after the assembly kept in split form
// (ctem_split.hip), so that the products read their operands without converting them.
// Buffers (nb matrices each, 8 bytes per entry): 0 A as assembled (FP32), read raw by the three
// Taylor products as B = A 2^-s (split on the fly when staged); 1 L; 2 B2; 3 P, later the FP32
// copy of the final F of every matrix (written by its last product) for the depth kernel; 4 R,
// then the squarings alternate between 5 (F0) and 4. wmax holds six FP32 maxima and six
// exponents per matrix (slot 0: L, slot 1: B, slots 2-5 as the buffers), then the max of A.
int32_t expm_depth_split(ctem_plan* tp, hipStream_t s, int npad, int nb, float2* W32, const double* colpart,
                         const float* amax, float* wmax, double2* Psi64, int64_t spsi, int nz,
                         int32_t* d_sq, int32_t* d_err, const std::vector<int>& h_n, std::vector<int32_t>& h_sq,
                         int depth_grid, bool sym, ChunkSched* cs) {
  const int64_t per = (int64_t)npad * npad;
  uint16_t* S[6];
  float* mx[6];
  int32_t* ex[6];
  for (int i = 0; i < 6; i++) {
    S[i] = (uint16_t*)(W32 + i * per * nb);
    mx[i] = wmax + i * nb;
    ex[i] = (int32_t*)(wmax + (6 + i) * nb);
  }
  launch_fill_zero(s, wmax, 12 * sizeof(float) * (size_t)nb);
  // the squaring count, exponent and max of B = A 2^-s; the first products split A themselves
  launch_norm_split(s, npad, nb, colpart, amax, kThetaSplit, kSmax, d_sq, d_err, ex[1], mx[1]);
  HIPCHK(hipGetLastError());
  if (cs->known) {
    h_sq = cs->sq;
  } else {
    h_sq.resize(nb);
    HIPCHK(hipMemcpyAsync(h_sq.data(), d_sq, 4 * (size_t)nb, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  int smax = 0;
  for (int b = 0; b < nb; b++) smax = std::max(smax, h_sq[b]);
  if (!cs->known) { cs->sq = h_sq; cs->smax = smax; cs->known = true; }
  launch_sq_check(s, nb, d_sq, smax, d_err);
  // flops of one complex product: 8 n^3, or 8 n n (n + 1) / 2 for the entries on and above the
  // diagonal of a symmetric result
  auto pflops = [&](int b) { return sym ? 4.0 * h_n[b] * h_n[b] * (h_n[b] + 1.0) : 8.0 * h_n[b] * h_n[b] * h_n[b]; };
  double n3 = 0.0;
  for (int b = 0; b < nb; b++) n3 += pflops(b);
  auto launch = [&](const SGemm& g, double flops) -> int32_t {
    GemmTimer t{tp, s};
    t.begin();
    launch_sgemm(s, g);
    t.end();
    HIPCHK(hipGetLastError());
    if (tp) { tp->f_gemm += flops; tp->n_gemm++; }
    return CTEM_OK;
  };
  // C = X Y (+ terms added below)
  auto mk = [&](int c, int x, int y) {
    SGemm g{};
    g.n = npad; g.nb = nb;
    g.X = S[x]; g.xe = ex[x]; g.xmax = mx[x];
    g.Y = S[y]; g.ye = ex[y]; g.ymax = mx[y];
    g.C = S[c]; g.ce = ex[c]; g.cmax = mx[c];
    g.alpha = 1.0f;
    g.sym = sym ? 1 : 0;
    return g;
  };
  auto add_m = [&](SGemm& g, int q, int m, float beta, float beta2) {
    g.M[q] = S[m]; g.me[q] = ex[m]; g.mmax[q] = mx[m];
    g.beta[q] = beta; g.beta2[q] = beta2;
  };
  auto add_c2 = [&](SGemm& g, int c2, float alpha2, float betaI2) {
    g.C2 = S[c2]; g.ce2 = ex[c2]; g.cmax2 = mx[c2];
    g.alpha2 = alpha2; g.betaI2 = betaI2;
  };
  // B as a raw operand: the assembled A (buffer 0) with B = A 2^-s (exponent and max in slot 1)
  auto raw_b = [&](SGemm& g, int q, float beta, float beta2) {
    g.M[q] = S[0]; g.me[q] = d_sq; g.mmax[q] = mx[1]; g.mraw[q] = 1;
    g.beta[q] = beta; g.beta2[q] = beta2;
  };
  int32_t rc;
  const double f3 = n3;
  // B2 = B B, P = x2 B2 + x1 B
  {
    SGemm g = mk(2, 1, 1);
    g.X = g.Y = S[0]; g.xraw = 1; g.rs = d_sq;
    add_c2(g, 3, kBbc8[1], 0.f);
    raw_b(g, 0, 0.f, kBbc8[0]);
    if ((rc = launch(g, f3))) return rc;
  }
  // B4 = B2 P;  L = B4 + x3 B2,  R = x7 B4 + x4 I + x6 B2 + x5 B  (L into buffer 1, its exponent
  // and max in slot 0: A stays in buffer 0 for the next product)
  {
    SGemm g = mk(0, 2, 3);
    g.C = S[1];
    add_c2(g, 4, kBbc8[6], kBbc8[3]);
    add_m(g, 0, 2, kBbc8[2], kBbc8[5]);
    raw_b(g, 1, 0.f, kBbc8[4]);
    if ((rc = launch(g, f3))) return rc;
  }
  // F0 = L R + B + y2 B2 (the final F where no squaring follows)
  float2* Ff = W32 + 3 * per * nb;
  // matrices too large for the on-chip depth kernel: two squarings more give E^2 - I and
  // E^4 - I (FP32 copies in buffers 6 and 7) for the depth recursion over groups of four depths
  const bool quads = npad > 224;
  float2* Ff2 = quads ? W32 + 6 * per * nb : nullptr;
  float2* Ff4 = quads ? W32 + 7 * per * nb : nullptr;
  const int extra = quads ? 2 : 0;
  {
    SGemm g = mk(5, 0, 4);
    g.X = S[1];
    raw_b(g, 0, 1.f, 0.f);
    add_m(g, 1, 2, kBbc8[7], 0.f);
    g.Cf = Ff; g.f_sq = d_sq; g.f_at = 0; g.Cf2 = Ff2; g.Cf4 = Ff4;
    if ((rc = launch(g, f3))) return rc;
  }
  const int buf[2] = {5, 4};
  for (int i = 0; i < smax + extra; i++) {
    double fl = 0.0;
    for (int b = 0; b < nb; b++) if (i < h_sq[b] + extra) fl += pflops(b);
    const int x = buf[i & 1], out = buf[(i + 1) & 1];
    // the output buffer's maxima were those of two steps back
    launch_fill_zero(s, mx[out], sizeof(float) * (size_t)nb);
    SGemm g = mk(out, x, x);
    add_m(g, 0, x, 2.f, 0.f);
    // every matrix of the chunk in the grid; those whose squarings (plus the extra two) are done
    // leave at once (counts stay on the device)
    g.active = d_sq; g.step = i; g.active_add = extra;
    g.Cf = Ff; g.f_sq = d_sq; g.f_at = i + 1; g.Cf2 = Ff2; g.Cf4 = Ff4;
    if ((rc = launch(g, fl))) return rc;
  }
  // depth on E = I + F with FP64 accumulation (the depth kernel keeps the FP32 matrix on chip,
  // or for larger matrices steps four depths at a time with E^4)
  if (quads)
    launch_depth_quad(s, npad, nb, nz, Ff, Ff2, Ff4, Psi64, spsi, depth_grid);
  else
    launch_depth(s, npad, nb, nz, Ff, 1, 1, Psi64, spsi, depth_grid);
  HIPCHK(hipGetLastError());
  return CTEM_OK;
}

// The Frechet derivative L_exp(A, D0) in the mixed form, from the pair (B, D) = (A 2^-s, D0 2^-s)
// in slots 0 and 1 of W32 (maxima in wmax slots 0 and 1): the degree-8 Taylor polynomial of
// expm_depth_mixed and its derivative in the direction D by the product rule,
//   B2 = B B                 dB2 = B D + D B
//   P  = x2 B2 + x1 B        dP  = x2 dB2 + x1 D
//   L  = B2 P + x3 B2        dL  = B2 dP + dB2 P + x3 dB2
//   R  = x7 B2 P + x4 I + x6 B2 + x5 B
//                            dR  = x7 (B2 dP + dB2 P) + x6 dB2 + x5 D
//   F0 = L R + B + y2 B2     G0  = L dR + dL R + D + y2 dB2
// (the upper-right block of the same polynomial of [[B, D], [0, B]]), then s squarings of the pair
// in the E = I + F form, F' = F F + 2F, G' = F G + G F + 2G, so that every product acts on the
// small part. Products on split-FP16 MFMA from FP32 storage (k_cgemm), nine for the polynomial and
// three per squaring, each matrix leaving the squarings after its own count. W32: 16 slots of nb
// matrices; wmax: 16 maxima per matrix (slots 0 and 1 set, the rest zeroed here). Returns G
// (slot 13) or null with rc_out set.
float2* expm_pair_mixed(hipStream_t s, int npad, int nb, float2* W32, float* wmax, int32_t* d_sq, int32_t* d_err,
                        std::vector<int32_t>& h_sq, int32_t& rc_out) {
  const int64_t per = (int64_t)npad * npad;
  float2* S[16];
  float* m[16];
  for (int i = 0; i < 16; i++) {
    S[i] = W32 + i * per * nb;
    m[i] = wmax + i * nb;
  }
  rc_out = CTEM_OK;
  launch_fill_zero(s, m[2], 14 * sizeof(float) * (size_t)nb);
  h_sq.resize(nb);
  if (hipMemcpyAsync(h_sq.data(), d_sq, 4 * (size_t)nb, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    rc_out = fail(CTEM_E_HIP, "expm_pair_mixed: reading the squaring counts");
    return nullptr;
  }
  int smax = 0;
  for (int b = 0; b < nb; b++) smax = std::max(smax, h_sq[b]);
  // C = X Y + sum beta_j M_j (+ C2 = alpha2 X Y + betaI2 I + sum beta2_j M_j)
  struct Term { int slot; float beta, beta2; };
  auto gemm = [&](int c, bool cmax, int x, int y, std::initializer_list<Term> ms, int c2 = -1, float alpha2 = 0.f,
                  float betaI2 = 0.f, const int32_t* active = nullptr, int step = 0) -> int32_t {
    CGemm g{};
    g.m = g.n = g.k = npad;
    g.ldx = g.ldy = g.ldc = g.ldm = npad;
    g.X = S[x]; g.sx = per; g.Y = S[y]; g.sy = per; g.C = S[c]; g.sc = per; g.sm = per;
    g.alpha = 1.0f;
    int q = 0;
    for (const Term& t : ms) { g.M[q] = S[t.slot]; g.beta[q] = t.beta; g.beta2[q] = t.beta2; q++; }
    g.nb = nb; g.terms = CTEM_SPLIT_FP16X3; g.err = d_err;
    g.xmax = m[x]; g.ymax = m[y]; g.cmax = cmax ? m[c] : nullptr;
    if (c2 >= 0) { g.C2 = S[c2]; g.cmax2 = m[c2]; g.alpha2 = alpha2; g.betaI2 = betaI2; }
    g.active = active; g.step = step;
    launch_cgemm(s, g);
    HIPCHK(hipGetLastError());
    return CTEM_OK;
  };
  const float x1 = kBbc8[0], x2 = kBbc8[1], x3 = kBbc8[2], x4 = kBbc8[3], x5 = kBbc8[4], x6 = kBbc8[5],
              x7 = kBbc8[6], y2 = kBbc8[7];
  enum { sB = 0, sD, sB2, sP, sT, sdB2, sdP, sL, sR, sdL, sdR, sF0, sT3, sG0, sF1, sG1 };
  int32_t rc;
  if ((rc = gemm(sB2, true, sB, sB, {{sB, 0.f, x1}}, sP, x2)) ||                              // B2, P
      (rc = gemm(sT, false, sD, sB, {})) ||                                                 // D B
      (rc = gemm(sdB2, true, sB, sD, {{sT, 1.f, x2}, {sD, 0.f, x1}}, sdP, x2)) ||             // dB2, dP
      (rc = gemm(sL, true, sB2, sP, {{sB2, x3, x6}, {sB, 0.f, x5}}, sR, x7, x4)) ||            // L, R
      (rc = gemm(sT, false, sdB2, sP, {})) ||                                               // dB2 P
      (rc = gemm(sdL, true, sB2, sdP, {{sT, 1.f, x7}, {sdB2, x3, x6}, {sD, 0.f, x5}}, sdR, x7)) ||  // dL, dR
      (rc = gemm(sF0, true, sL, sR, {{sB, 1.f, 0.f}, {sB2, y2, 0.f}})) ||                     // F0
      (rc = gemm(sT3, false, sdL, sR, {})) ||                                               // dL R
      (rc = gemm(sG0, true, sL, sdR, {{sT3, 1.f, 0.f}, {sD, 1.f, 0.f}, {sdB2, y2, 0.f}}))) {  // G0
    rc_out = rc;
    return nullptr;
  }
  const int fb[2] = {sF0, sF1}, gb[2] = {sG0, sG1};
  for (int i = 0; i < smax; i++) {
    const int f = fb[i & 1], fo = fb[(i + 1) & 1], gi = gb[i & 1], go = gb[(i + 1) & 1];
    // the output slots' maxima were those of two steps back
    launch_fill_zero(s, m[fo], sizeof(float) * (size_t)nb);
    launch_fill_zero(s, m[go], sizeof(float) * (size_t)nb);
    if ((rc = gemm(fo, true, f, f, {{f, 2.f, 0.f}}, -1, 0.f, 0.f, d_sq, i)) ||            // F F + 2F
        (rc = gemm(sT, false, gi, f, {}, -1, 0.f, 0.f, d_sq, i)) ||                        // G F
        (rc = gemm(go, true, f, gi, {{sT, 1.f, 0.f}, {gi, 2.f, 0.f}}, -1, 0.f, 0.f, d_sq, i))) {  // F G + G F + 2G
      rc_out = rc;
      return nullptr;
    }
  }
  launch_cparity_copy(s, npad, nb, d_sq, S[sG1], S[sG0]);
  if (hipGetLastError() != hipSuccess) {
    rc_out = fail(CTEM_E_HIP, "expm_pair_mixed: launch failed");
    return nullptr;
  }
  return S[sG0];
}

}  // namespace


// Algorithmic flops of one direction: F = 8n^2w + 8n^3(q+j) + (32/3)n^3 + 16n^2Z + 4n^2P (the
// Bethe product, the Pade powers and squarings, the Pade solve, propagation with psi psi^H
// accumulation, the S contraction), with the Pade degree q and squaring count j that the
// reference's MatrixExponential picks for this matrix: modA = max |a_ij| of
// A = 2 pi i dz DynMat / fl32(2kn), ilev = nint(log10(modA)) + 3 clamped to 1..6
// (math.f90:208-219). amod is max |a_ij| of A.
static void alg_flops(int n, int w, int nz, int nsite, double amod, double /*dz*/, double& tot,
                      double& prod, double& solve) {
  static const int qPade[6] = {2, 3, 4, 4, 4, 4}, jPade[6] = {0, 0, 1, 5, 8, 11};
  int ilev = (amod > 0.0 && std::isfinite(amod)) ? (int)std::lround(std::log10(amod)) + 3 : 1;
  ilev = std::min(6, std::max(1, ilev));
  const double q = qPade[ilev - 1], j = jPade[ilev - 1];
  const double n2 = (double)n * n, n3 = n2 * n;
  const double fp = 8.0 * n3 * (q + j), fs = 32.0 / 3.0 * n3;
  prod += fp;
  solve += fs;
  tot += 8.0 * n2 * w + fp + fs + 16.0 * n2 * nz + 4.0 * n2 * nsite;
}

namespace {

// form of the matrix exponential a run uses (the squaring counts differ between forms)
int run_form(const ctem_plan* p, int32_t precision) {
  if (precision == CTEM_PREC_FP64) return 0;
  return (p->terms_taylor == CTEM_SPLIT_FP16X3 && p->terms_square == CTEM_SPLIT_FP16X3) ? 1 : 2;
}

size_t chunk_budget(const ctem_plan* p) { return p->chunk_bytes; }

// Device memory a run of the plan needs, chunk by chunk; grows the context's work space (and
// the plan's buffers) to it before anything is enqueued, so that enqueueing allocates nothing.
int32_t run_reserve(ctem_plan* p, int32_t precision) {
  ctem_ctx* c = p->ctx;
  const int ns = std::max(1, p->cr->nsite);
  int32_t rc;
  if ((rc = p->dots.ensure(8 * (size_t)std::max(1, p->numk) * ns)) || (rc = p->err.ensure(4))) return rc;
  size_t a = 0, w = 0, psi = 0, sq = 0, part = 0, wmax = 0, col = 0, amod = 0;
  for (auto& kv : p->buckets) {
    const int npad = kv.first;
    const int nbk = (int)kv.second.size();
    const int64_t per = (int64_t)npad * npad;
    const int64_t spsi = (int64_t)npad * p->nz;
    const size_t per_dir = (size_t)(6 * per + spsi) * sizeof(double2) + 64;
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>(nbk, chunk_budget(p) / per_dir));
    const size_t nb = (size_t)std::min(chunk, nbk);
    a = std::max(a, nb * per * sizeof(double2));
    w = std::max(w, 5 * nb * per * sizeof(double2));
    psi = std::max(psi, nb * spsi * sizeof(double2));
    sq = std::max(sq, 4 * nb);
    // partial sums of the contraction: one per tile pair (FP64 form) or per depth part of a
    // direction (mixed form, two parts)
    part = std::max(part, 8 * nb * std::max(2, (npad / 32) * (npad / 32)) * ns);
    wmax = std::max(wmax, 13 * sizeof(float) * nb);
    col = std::max(col, sizeof(double) * nb * npad * ((npad + 63) / 64));
    amod = std::max(amod, 8 * nb);
  }
  const void* before[8] = {c->w.A.p, c->w.W.p, c->w.Psi.p, c->w.sq.p, c->w.partial.p, c->w.wmax.p,
                           c->w.colsum.p, c->w.amod.p};
  if ((rc = c->w.A.ensure(a)) || (rc = c->w.W.ensure(w)) || (rc = c->w.Psi.ensure(psi)) ||
      (rc = c->w.sq.ensure(sq)) || (rc = c->w.partial.ensure(part)) ||
      (precision == CTEM_PREC_MIXED && (rc = c->w.wmax.ensure(wmax))) ||
      (precision == CTEM_PREC_MIXED && (rc = c->w.colsum.ensure(col))) ||
      (precision == CTEM_PREC_FP64 && (rc = c->w.amod.ensure(amod))))
    return rc;
  const void* after[8] = {c->w.A.p, c->w.W.p, c->w.Psi.p, c->w.sq.p, c->w.partial.p, c->w.wmax.p,
                          c->w.colsum.p, c->w.amod.p};
  for (int i = 0; i < 8; i++)
    if (before[i] != after[i]) c->wgen++;
  return CTEM_OK;
}

// Every kernel of one run of the plan, in stream order. With `timed`, the GEMM launches are
// bracketed by events; with `sync_ok`, chunks whose squaring counts are not yet known read them
// back (otherwise such a chunk is an error: the caller warms the schedule first).
int32_t run_enqueue(ctem_plan* p, int32_t precision, bool timed, bool sync_ok, double fa[3]) {
  ctem_ctx* c = p->ctx;
  hipStream_t s = c->stream;
  const int form = run_form(p, precision);
  launch_fill_zero(s, p->err.p, 4);
  BeamView V{p->nns.as<int32_t>(), p->nnw.as<int32_t>(), p->soff.as<int64_t>(), p->woff.as<int64_t>(),
             p->sidx.as<int32_t>(), p->ssg.as<float>(), p->widx.as<int32_t>(), p->wsg.as<float>()};
  DevRefl R = p->cr->refl();
  DevLut L = p->cr->lutv(p->err.as<int32_t>());
  const double dzt = p->dz / p->thick;
  ctem_plan* tp = timed ? p : nullptr;
  size_t off = 0;
  int32_t rc;
  for (auto& kv : p->buckets) {
    const int npad = kv.first;
    const int nbk = (int)kv.second.size();
    const int64_t per = (int64_t)npad * npad;
    const int64_t spsi = (int64_t)npad * p->nz;
    const size_t per_dir = (size_t)(6 * per + spsi) * sizeof(double2) + 64;
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>(nbk, chunk_budget(p) / per_dir));
    for (int c0 = 0; c0 < nbk; c0 += chunk) {
      const int nb = std::min(chunk, nbk - c0);
      ChunkSched& cs = p->sched[std::make_tuple(form, npad, c0)];
      if (!cs.known && !sync_ok) return fail(CTEM_E_STATE, "run schedule not known for a chunk");
      Bucket B{npad, nb, p->dirs.as<int32_t>() + off + c0};
      std::vector<int> hn(nb);
      for (int b = 0; b < nb; b++) hn[b] = p->h_nns[kv.second[c0 + b]];
      const bool split = form == 1;
      // complex-symmetric DynMat (inversion-symmetric LUT): upper tiles only, split form alone
      const bool sym = split && p->symmetry == 1 && p->cr->centro;
      float* amax = nullptr;
      if (precision == CTEM_PREC_MIXED) {
        // six maxima and six exponents per matrix for the products, then the max of A
        amax = c->w.wmax.as<float>() + 12 * (size_t)nb;
        launch_fill_zero(s, amax, sizeof(float) * (size_t)nb);
      }
      if (precision == CTEM_PREC_FP64)
        launch_assemble(s, B, V, R, L, p->kn.as<double>(), p->cr->mlambda, p->cr->dynupz, p->dz, c->w.A.as<double2>());
      else if (split) {
        int nwmax = 0;
        for (int bb = 0; bb < nb; bb++) nwmax = std::max(nwmax, p->h_nnw[kv.second[c0 + bb]]);
        launch_assemble_split(s, B, V, R, L, p->kn.as<double>(), p->cr->mlambda, p->cr->dynupz, p->dz,
                              p->cr->lutmax, nwmax, c->w.W.as<float2>(), c->w.colsum.as<double>(), amax, sym);
      } else
        launch_assemble32(s, B, V, R, L, p->kn.as<double>(), p->cr->mlambda, p->cr->dynupz, p->dz, c->w.W.as<float2>());
      HIPCHK(hipGetLastError());
      if (precision == CTEM_PREC_FP64) {
        const bool first = !cs.known;
        std::vector<double> hmod;
        rc = expm_depth_fp64(tp, s, npad, nb, c->w.A.as<double2>(), c->w.W.as<double2>(), c->w.Psi.as<double2>(), spsi,
                             p->nz, c->w.sq.as<int32_t>(), p->err.as<int32_t>(), hn, p->h_sq, c->cu_count,
                             c->w.amod.as<double>(), &hmod, &cs);
        if (!rc && first && fa) {
          for (int b = 0; b < nb; b++) {
            const int d = kv.second[c0 + b];
            alg_flops(hn[b], p->h_nnw[d], p->nz, p->cr->nsite, hmod[b], p->dz, fa[0], fa[1], fa[2]);
          }
        }
      } else if (split) {
        rc = expm_depth_split(tp, s, npad, nb, c->w.W.as<float2>(), c->w.colsum.as<double>(), amax,
                              c->w.wmax.as<float>(), c->w.Psi.as<double2>(), spsi, p->nz,
                              c->w.sq.as<int32_t>(), p->err.as<int32_t>(), hn, p->h_sq, c->cu_count, sym, &cs);
      } else {
        rc = expm_depth_mixed(tp, s, npad, nb, c->w.W.as<float2>(), c->w.wmax.as<float>(),
                              c->w.Psi.as<double2>(), spsi, p->nz, c->w.sq.as<int32_t>(), p->err.as<int32_t>(), hn,
                              p->h_sq, p->terms_taylor, p->terms_square, c->cu_count, &cs);
      }
      if (rc) return rc;
      launch_lgh_dots(s, B, V, R, L, c->w.Psi.as<double2>(), spsi, p->nz, p->lam.as<double>(), dzt,
                      c->w.partial.as<double>(), p->dots.as<double>(), p->keep_dir, p->keep_lgh.as<double2>(),
                      precision == CTEM_PREC_MIXED);
      HIPCHK(hipGetLastError());
    }
    off += nbk;
  }
  return CTEM_OK;
}

bool sched_complete(const ctem_plan* p, int form) {
  for (auto& kv : p->buckets) {
    const int npad = kv.first;
    const int nbk = (int)kv.second.size();
    const int64_t per = (int64_t)npad * npad;
    const int64_t spsi = (int64_t)npad * p->nz;
    const size_t per_dir = (size_t)(6 * per + spsi) * sizeof(double2) + 64;
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>(nbk, chunk_budget(p) / per_dir));
    for (int c0 = 0; c0 < nbk; c0 += chunk) {
      auto it = p->sched.find(std::make_tuple(form, npad, c0));
      if (it == p->sched.end() || !it->second.known) return false;
    }
  }
  return true;
}

void graph_drop(hipGraphExec_t& g) {
  if (g) (void)hipGraphExecDestroy(g);
  g = nullptr;
}

// what a captured run depends on besides the device data: form, options and the addresses of
// the work space (its generation)
std::vector<uint64_t> graph_key_of(const ctem_plan* p, int32_t precision) {
  return {p->serial, (uint64_t)precision, (uint64_t)p->terms_taylor, (uint64_t)p->terms_square,
          (uint64_t)p->symmetry, p->ctx->wgen, p->bufsig, (uint64_t)(uintptr_t)p->dots.p,
          (uint64_t)(uintptr_t)p->cr.get(), (uint64_t)p->bucket_multiple, (uint64_t)p->chunk_bytes};
}

}  // namespace

// Everything a run needs that is not device work: the crystal, the beam sets and their bucket
// layout, the work space. Returns CTEM_OK when the plan can be enqueued.
static int32_t run_prepare(ctem_plan* p, int32_t precision, int32_t keep_lgh_ik) {
  if (precision != CTEM_PREC_FP64 && precision != CTEM_PREC_MIXED)
    return fail(CTEM_E_ARG, "unknown precision %d", precision);
  if (p->ecp_mode >= 0) return fail(CTEM_E_STATE, "a channeling-pattern plan runs with ctem_ecp_run");
  ctem_ctx* c = p->ctx;
  if (!p->cr) p->cr = c->cr;
  if (!p->cr) return fail(CTEM_E_STATE, "ctem_plan_run: no crystal uploaded");
  if (keep_lgh_ik >= p->numk) return fail(CTEM_E_ARG, "keep_lgh_ik out of range");
  int32_t rc;
  if (!p->selected && (rc = ctem_plan_select(p, nullptr, nullptr))) return rc;
  HIPCHK(hipSetDevice(c->device));
  p->keep_dir = keep_lgh_ik;
  p->keep_n = (keep_lgh_ik >= 0) ? p->h_nns[keep_lgh_ik] : 0;
  if (p->keep_n && (rc = p->keep_lgh.ensure(16 * (size_t)p->keep_n * p->keep_n))) return rc;
  return run_reserve(p, precision);
}

static int32_t run_finish(ctem_plan* p, int32_t precision, const double fa[3], bool have_fa) {
  ctem_ctx* c = p->ctx;
  hipStream_t s = c->stream;
  HIPCHK(hipEventRecord(c->ev[1], s));
  HIPCHK(hipStreamSynchronize(s));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  p->t_total = ms;
  for (size_t i = 0; i + 1 < p->evused; i += 2) {
    float g = 0.0f;
    HIPCHK(hipEventElapsedTime(&g, p->evpool[i], p->evpool[i + 1]));
    p->t_gemm += g;
  }
  int32_t rc;
  if ((rc = check_err_flag(p->err, s, "ctem_plan_run"))) return rc;
  if (precision == CTEM_PREC_FP64 && have_fa) { p->fa_total = fa[0]; p->fa_prod = fa[1]; p->fa_solve = fa[2]; }
  p->ran = true;
  return CTEM_OK;
}

static void reset_timing(ctem_plan* p) {
  p->t_total = p->t_gemm = p->f_gemm = 0.0;
  p->n_gemm = 0;
  p->evused = 0;
}

extern "C" int32_t ctem_plan_run(ctem_plan* p, int32_t precision, int32_t keep_lgh_ik) {
  if (!p) return fail(CTEM_E_ARG, "null plan");
  int32_t rc;
  if ((rc = run_prepare(p, precision, keep_lgh_ik))) return rc;
  ctem_ctx* c = p->ctx;
  hipStream_t s = c->stream;
  const int form = run_form(p, precision);
  reset_timing(p);
  double fa[3] = {0.0, 0.0, 0.0};
  const bool have_fa = precision == CTEM_PREC_FP64 && !sched_complete(p, form);
  // graph mode, once every chunk's squaring counts are known and no Lgh is kept: the run is
  // captured once and then replayed as one launch
  if (p->graph_mode && keep_lgh_ik < 0 && sched_complete(p, form)) {
    const std::vector<uint64_t> key = graph_key_of(p, precision);
    if (!p->graph || p->graph_key != key) {
      graph_drop(p->graph);
      HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      rc = run_enqueue(p, precision, false, false, nullptr);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(s, &g);
      if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
      if (e != hipSuccess) return fail(CTEM_E_HIP, "graph capture: %s", hipGetErrorString(e));
      e = hipGraphInstantiate(&p->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) { p->graph = nullptr; return fail(CTEM_E_HIP, "graph instantiate: %s", hipGetErrorString(e)); }
      p->graph_key = key;
    }
    HIPCHK(hipEventRecord(c->ev[0], s));
    HIPCHK(hipGraphLaunch(p->graph, s));
    rc = run_finish(p, precision, fa, false);
  } else {
    HIPCHK(hipEventRecord(c->ev[0], s));
    if ((rc = run_enqueue(p, precision, true, true, fa))) return rc;
    rc = run_finish(p, precision, fa, have_fa);
  }
  if (rc == CTEM_E_STATE && strstr(g_err.c_str(), "schedule")) {
    // a matrix needed more squarings than an earlier run of the plan: run again with the counts
    // read back
    for (auto it = p->sched.begin(); it != p->sched.end();)
      it = (std::get<0>(it->first) == form) ? p->sched.erase(it) : std::next(it);
    graph_drop(p->graph);
    reset_timing(p);
    fa[0] = fa[1] = fa[2] = 0.0;
    HIPCHK(hipEventRecord(c->ev[0], s));
    if ((rc = run_enqueue(p, precision, true, true, fa))) return rc;
    rc = run_finish(p, precision, fa, precision == CTEM_PREC_FP64);
  }
  return rc;
}

// No counterpart in the reference: tuning of the batched launch (see the header).
extern "C" int32_t ctem_plan_set_tuning(ctem_plan* p, int32_t bucket_multiple, double chunk_gib) {
  if (!p) return fail(CTEM_E_ARG, "null plan");
  if (bucket_multiple != 32 && bucket_multiple != 64)
    return fail(CTEM_E_ARG, "bucket multiple must be 32 or 64 (got %d)", bucket_multiple);
  if (!(chunk_gib >= 0.25 && chunk_gib <= 256.0)) return fail(CTEM_E_ARG, "chunk size must be 0.25..256 GiB");
  const size_t bytes = (size_t)(chunk_gib * (double)(1ull << 30));
  if (bucket_multiple != p->bucket_multiple || bytes != p->chunk_bytes) {
    p->bucket_multiple = bucket_multiple;
    p->chunk_bytes = bytes;
    p->sched.clear();
    graph_drop(p->graph);
    p->selected = false;  // the bucket layout follows at the next run
  }
  return CTEM_OK;
}

// No counterpart in the reference: graph mode of a plan (see the header).
extern "C" int32_t ctem_plan_set_graph(ctem_plan* p, int32_t on) {
  if (!p) return fail(CTEM_E_ARG, "null plan");
  if (on != 0 && on != 1) return fail(CTEM_E_ARG, "graph mode must be 0 or 1");
  p->graph_mode = on;
  if (!on) graph_drop(p->graph);
  return CTEM_OK;
}

// No counterpart in the reference: the plans of an energy series as one captured graph.
extern "C" int32_t ctem_plans_run(ctem_plan* const* plans, int32_t nplans, int32_t precision) {
  if (!plans || nplans < 1) return fail(CTEM_E_ARG, "ctem_plans_run: empty plan list");
  ctem_ctx* c = plans[0] ? plans[0]->ctx : nullptr;
  if (!c) return fail(CTEM_E_ARG, "ctem_plans_run: null plan");
  for (int i = 0; i < nplans; i++)
    if (!plans[i] || plans[i]->ctx != c) return fail(CTEM_E_ARG, "ctem_plans_run: plans of one context only");
  int32_t rc;
  // warm: every plan's bucket layout, work space and squaring counts (eager runs where needed)
  for (int i = 0; i < nplans; i++) {
    ctem_plan* p = plans[i];
    if ((rc = run_prepare(p, precision, -1))) return rc;
    if (!sched_complete(p, run_form(p, precision)))
      if ((rc = ctem_plan_run(p, precision, -1))) return rc;
  }
  for (int i = 0; i < nplans; i++)
    if ((rc = run_reserve(plans[i], precision))) return rc;
  std::vector<uint64_t> key{(uint64_t)nplans};
  for (int i = 0; i < nplans; i++) {
    auto k = graph_key_of(plans[i], precision);
    key.insert(key.end(), k.begin(), k.end());
  }
  hipStream_t s = c->stream;
  HIPCHK(hipSetDevice(c->device));
  if (!c->mgraph || c->mgraph_key != key) {
    graph_drop(c->mgraph);
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    rc = CTEM_OK;
    for (int i = 0; i < nplans && !rc; i++) {
      plans[i]->keep_dir = -1;
      rc = run_enqueue(plans[i], precision, false, false, nullptr);
    }
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return fail(CTEM_E_HIP, "graph capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&c->mgraph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) { c->mgraph = nullptr; return fail(CTEM_E_HIP, "graph instantiate: %s", hipGetErrorString(e)); }
    c->mgraph_key = key;
  }
  for (int i = 0; i < nplans; i++) reset_timing(plans[i]);
  HIPCHK(hipEventRecord(c->ev[0], s));
  HIPCHK(hipGraphLaunch(c->mgraph, s));
  HIPCHK(hipEventRecord(c->ev[1], s));
  HIPCHK(hipStreamSynchronize(s));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  for (int i = 0; i < nplans; i++) {
    ctem_plan* p = plans[i];
    p->t_total = (i == 0) ? ms : 0.0;
    if ((rc = check_err_flag(p->err, s, "ctem_plans_run"))) {
      graph_drop(c->mgraph);
      return rc;
    }
    p->keep_n = 0;
    p->ran = true;
  }
  return CTEM_OK;
}

extern "C" int32_t ctem_plan_set_mixed_terms(ctem_plan* p, int32_t taylor_terms, int32_t square_terms) {
  if (!p) return fail(CTEM_E_ARG, "null plan");
  auto ok = [](int32_t t) { return t == CTEM_SPLIT_BF16X3 || t == CTEM_SPLIT_BF16X6 || t == CTEM_SPLIT_FP16X3; };
  if (!ok(taylor_terms) || !ok(square_terms))
    return fail(CTEM_E_ARG, "split scheme must be CTEM_SPLIT_BF16X3 (3), CTEM_SPLIT_BF16X6 (6) or CTEM_SPLIT_FP16X3 (16)");
  p->terms_taylor = taylor_terms;
  p->terms_square = square_terms;
  return CTEM_OK;
}

extern "C" int32_t ctem_plan_set_symmetry(ctem_plan* p, int32_t mode, int32_t* in_effect) {
  if (!p) return fail(CTEM_E_ARG, "null plan");
  if (mode != 0 && mode != 1) return fail(CTEM_E_ARG, "symmetry mode must be 0 (off) or 1 (auto)");
  if (!p->cr) p->cr = p->ctx->cr;
  if (!p->cr) return fail(CTEM_E_STATE, "ctem_plan_set_symmetry: no crystal uploaded");
  p->symmetry = mode;
  if (in_effect) *in_effect = (mode == 1 && p->cr->centro) ? 1 : 0;
  return CTEM_OK;
}

extern "C" int32_t ctem_plan_get_dots(ctem_plan* p, double* dots) {
  if (!p || !p->ran) return fail(CTEM_E_STATE, "ctem_plan_get_dots: run the plan first");
  if (!dots) return fail(CTEM_E_ARG, "null output");
  const size_t n = (size_t)p->numk * p->cr->nsite;
  if (n) HIPCHK(hipMemcpy(dots, p->dots.p, 8 * n, hipMemcpyDeviceToHost));
  return CTEM_OK;
}

extern "C" int32_t ctem_plan_get_svals(ctem_plan* p, float* svals) {
#pragma clang fp contract(off)
  if (!p || !p->ran) return fail(CTEM_E_STATE, "ctem_plan_get_svals: run the plan first");
  if (!svals) return fail(CTEM_E_ARG, "null output");
  const size_t n = (size_t)p->numk * p->cr->nsite;
  std::vector<double> d(n);
  if (n) HIPCHK(hipMemcpy(d.data(), p->dots.p, 8 * n, hipMemcpyDeviceToHost));
  const float nat = p->cr->nat;
  for (size_t i = 0; i < n; i++) svals[i] = (float)d[i] / nat;
  return CTEM_OK;
}

extern "C" int32_t ctem_plan_get_lgh(ctem_plan* p, double* lgh, int32_t cap, int32_t* n) {
  if (!p || !p->ran || p->keep_dir < 0) return fail(CTEM_E_STATE, "ctem_plan_get_lgh: no Lgh kept");
  if (!n) return fail(CTEM_E_ARG, "null output");
  *n = p->keep_n;
  if (!lgh) return CTEM_OK;
  if (cap < p->keep_n * p->keep_n) return fail(CTEM_E_ARG, "capacity too small");
  HIPCHK(hipMemcpy(lgh, p->keep_lgh.p, 16 * (size_t)p->keep_n * p->keep_n, hipMemcpyDeviceToHost));
  return CTEM_OK;
}

// ------------------------------------------------------------------------------------------
// reverse mode (FP64): gradient of f(dots) with respect to the LUT and the site B factors
// The reverse pass (both forms). FP64: forward and Frechet exponential in FP64 (the pair form of
// expm_pair_fp64). Mixed: the forward is the plan's own mixed form (so psi, and the intensities
// the caller's loss saw, are those of ctem_plan_run in mixed precision), the backward depth
// recursion and E_bar stay FP64, and the Frechet exponential runs on split-FP16 products
// (expm_pair_mixed). Squaring stages are recomputed, not stored.
static int32_t plan_adjoint(ctem_plan* p, bool mixed, const double* gdots, double* glut, double* gb) {
  const char* who = mixed ? "ctem_plan_adjoint_mixed" : "ctem_plan_adjoint";
  if (!p || !gdots || !glut) return fail(CTEM_E_ARG, "%s: null argument", who);
  ctem_ctx* c = p->ctx;
  if (!p->cr) p->cr = c->cr;
  if (!p->cr) return fail(CTEM_E_STATE, "%s: no crystal uploaded", who);
  if (p->cr->nsite > 0 && !gb) return fail(CTEM_E_ARG, "%s: gb is NULL", who);
  if (p->ecp_mode >= 0) return fail(CTEM_E_STATE, "%s: not for a channeling-pattern plan", who);
  int32_t rc;
  if (!p->selected && (rc = ctem_plan_select(p, nullptr, nullptr))) return rc;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const int nk = p->numk, ns = std::max(1, p->cr->nsite), nz = p->nz;
  for (int i = 0; i < nk; i++)
    if (npad_of(p->h_nns[i]) > 512) return fail(CTEM_E_RANGE, "%s: at most 512 strong beams per direction", who);
  const int nrep = adj_lut_replicas(p->cr->ncell);
  const size_t glut_bytes = 16 * (size_t)p->cr->ncell * nrep;
  if ((rc = upload(c->w.g_dots, gdots, (size_t)nk * p->cr->nsite, s)) || (rc = c->w.g_lut.ensure(glut_bytes)) ||
      (rc = c->w.g_dots2.ensure(8 * (size_t)nk * ns)) || (rc = p->err.ensure(4)) ||
      (rc = p->dots.ensure(8 * (size_t)std::max(1, nk) * ns)))
    return rc;
  HIPCHK(hipMemsetAsync(c->w.g_lut.p, 0, glut_bytes, s));
  HIPCHK(hipMemsetAsync(c->w.g_dots2.p, 0, 8 * (size_t)nk * ns, s));
  HIPCHK(hipMemsetAsync(p->err.p, 0, 4, s));
  std::map<int, std::vector<int32_t>> buckets;
  for (int i = 0; i < nk; i++) buckets[npad_of(p->h_nns[i])].push_back(i);
  std::vector<int32_t> order;
  for (auto& kv : buckets) order.insert(order.end(), kv.second.begin(), kv.second.end());
  if ((rc = upload(c->w.dirs, order.data(), order.size(), s))) return rc;
  BeamView V{p->nns.as<int32_t>(), p->nnw.as<int32_t>(), p->soff.as<int64_t>(), p->woff.as<int64_t>(),
             p->sidx.as<int32_t>(), p->ssg.as<float>(), p->widx.as<int32_t>(), p->wsg.as<float>()};
  DevRefl R = p->cr->refl();
  DevLut L = p->cr->lutv(p->err.as<int32_t>());
  DevLut LdB = L;
  LdB.slut = p->cr->st->dslut.as<double2>();
  const double dzt = p->dz / p->thick;
  const int form = run_form(p, CTEM_PREC_MIXED);
  const bool split = mixed && form == 1;
  const bool sym = split && p->symmetry == 1 && p->cr->centro;
  const size_t budget = (size_t)16 << 30;
  size_t off = 0;
  std::vector<int32_t> h_sq, h_sq2;
  for (auto& kv : buckets) {
    const int npad = kv.first;
    const int nbk = (int)kv.second.size();
    const int64_t per = (int64_t)npad * npad;
    const int64_t spsi = (int64_t)npad * nz;
    const size_t per_dir = (size_t)(21 * per + 4 * spsi) * sizeof(double2) + 256;
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>(nbk, budget / per_dir));
    for (int c0 = 0; c0 < nbk; c0 += chunk) {
      const int nb = std::min(chunk, nbk - c0);
      // buffers shared with ctem_plan_run: a move changes the addresses its captured graphs hold
      const void* before[7] = {c->w.A.p, c->w.W.p, c->w.Psi.p, c->w.sq.p, c->w.partial.p, c->w.wmax.p, c->w.colsum.p};
      auto grow = [&]() -> int32_t {
        int32_t r;
        if ((r = c->w.A.ensure((size_t)nb * per * 16)) || (r = c->w.W.ensure((size_t)5 * nb * per * 16)) ||
            (r = c->w.Psi.ensure((size_t)nb * spsi * 16)) || (r = c->w.sq.ensure(4 * (size_t)nb)) ||
            (r = c->w.partial.ensure(8 * (size_t)nb * (npad / 32) * (npad / 32) * ns)) ||
            (r = c->w.aM.ensure((size_t)nb * per * 16)) || (r = c->w.aQ.ensure((size_t)nb * spsi * 16)) ||
            (r = c->w.aPhi.ensure((size_t)nb * spsi * 16)) || (r = c->w.aHm.ensure((size_t)nb * spsi * 16)) ||
            (r = c->w.aEbar.ensure((size_t)nb * per * 16)) || (r = c->w.aBlk.ensure((size_t)2 * nb * per * 16)) ||
            (r = c->w.aW2.ensure((size_t)10 * nb * per * 16)) || (r = c->w.aDbar.ensure((size_t)nb * per * 16)) ||
            (r = c->w.aSig.ensure(8 * (size_t)nb)))
          return r;
        if (mixed && ((r = c->w.wmax.ensure(13 * sizeof(float) * (size_t)nb)) ||
                      (r = c->w.colsum.ensure(sizeof(double) * (size_t)nb * npad * ((npad + 63) / 64))) ||
                      (r = c->w.aMax.ensure(16 * sizeof(float) * (size_t)nb))))
          return r;
        return CTEM_OK;
      };
      rc = grow();
      const void* after[7] = {c->w.A.p, c->w.W.p, c->w.Psi.p, c->w.sq.p, c->w.partial.p, c->w.wmax.p, c->w.colsum.p};
      for (int i = 0; i < 7; i++)
        if (before[i] != after[i]) c->wgen++;
      if (rc) return rc;
      Bucket B{npad, nb, c->w.dirs.as<int32_t>() + off + c0};
      std::vector<int> hn(nb);
      int max_nw = 0;
      for (int b = 0; b < nb; b++) {
        const int d = kv.second[c0 + b];
        hn[b] = p->h_nns[d];
        max_nw = std::max(max_nw, p->h_nnw[d]);
      }
      const double* kn = p->kn.as<double>();
      // forward: E and psi_1..psi_Z
      double2* E = nullptr;
      if (mixed) {
        float2* W32 = c->w.W.as<float2>();
        float* wmax = c->w.wmax.as<float>();
        ChunkSched cs;
        if (split) {
          float* amax = wmax + 12 * (size_t)nb;
          launch_fill_zero(s, amax, sizeof(float) * (size_t)nb);
          launch_assemble_split(s, B, V, R, L, kn, p->cr->mlambda, p->cr->dynupz, p->dz, p->cr->lutmax, max_nw, W32,
                                c->w.colsum.as<double>(), amax, sym);
          HIPCHK(hipGetLastError());
          rc = expm_depth_split(nullptr, s, npad, nb, W32, c->w.colsum.as<double>(), amax, wmax, c->w.Psi.as<double2>(),
                                spsi, nz, c->w.sq.as<int32_t>(), p->err.as<int32_t>(), hn, h_sq, c->cu_count, sym, &cs);
        } else {
          launch_assemble32(s, B, V, R, L, kn, p->cr->mlambda, p->cr->dynupz, p->dz, W32);
          HIPCHK(hipGetLastError());
          rc = expm_depth_mixed(nullptr, s, npad, nb, W32, wmax, c->w.Psi.as<double2>(), spsi, nz, c->w.sq.as<int32_t>(),
                                p->err.as<int32_t>(), hn, h_sq, p->terms_taylor, p->terms_square, c->cu_count, &cs);
        }
        if (rc) return rc;
        // the final F = E - I in FP32: buffer 3 of the split form, buffer 4 of the other
        E = c->w.A.as<double2>();
        launch_adj_e64(s, npad, nb, W32 + (split ? 3 : 4) * per * nb, E);
      } else {
        double2* A = c->w.A.as<double2>();
        launch_assemble(s, B, V, R, L, kn, p->cr->mlambda, p->cr->dynupz, p->dz, A);
        HIPCHK(hipGetLastError());
        E = expm_fp64(nullptr, s, npad, nb, A, c->w.W.as<double2>(), c->w.sq.as<int32_t>(), p->err.as<int32_t>(), hn,
                      h_sq, rc);
        if (!E) return rc;
        launch_depth(s, npad, nb, nz, E, 0, 0, c->w.Psi.as<double2>(), spsi, c->cu_count);
      }
      HIPCHK(hipGetLastError());
      // B: d dots_p / d B_p = Re sum Lgh dS_p/dB_p
      launch_lgh_dots(s, B, V, R, LdB, c->w.Psi.as<double2>(), spsi, nz, p->lam.as<double>(), dzt,
                      c->w.partial.as<double>(), c->w.g_dots2.as<double>(), -1, nullptr, 0);
      HIPCHK(hipGetLastError());
      // M = c (Lbar + Lbar^H), Q = M Psi
      launch_adj_m(s, B, V, R, L, c->w.g_dots.as<double>(), dzt, c->w.aM.as<double2>());
      HIPCHK(hipGetLastError());
      {
        ZGemm g{};
        g.m = npad; g.n = nz; g.k = npad;
        g.ldx = npad; g.ldy = npad; g.ldc = npad; g.ldm = npad;
        g.X = c->w.aM.as<double2>(); g.sx = per;
        g.Y = c->w.Psi.as<double2>(); g.sy = spsi;
        g.C = c->w.aQ.as<double2>(); g.sc = spsi;
        g.alpha = 1.0; g.nb = nb;
        launch_zgemm(s, g);
        HIPCHK(hipGetLastError());
      }
      // phi_k backward, Ebar = Phi [psi_0 .. psi_{Z-1}]^H
      launch_adj_depth(s, npad, nb, nz, E, c->w.aQ.as<double2>(), p->lam.as<double>(), c->w.aPhi.as<double2>());
      launch_adj_hm(s, npad, nb, nz, c->w.Psi.as<double2>(), spsi, c->w.aHm.as<double2>());
      HIPCHK(hipGetLastError());
      {
        ZGemm g{};
        g.m = npad; g.n = npad; g.k = nz;
        g.ldx = npad; g.ldy = nz; g.ldc = npad; g.ldm = npad;
        g.X = c->w.aPhi.as<double2>(); g.sx = spsi;
        g.Y = c->w.aHm.as<double2>(); g.sy = spsi;
        g.C = c->w.aEbar.as<double2>(); g.sc = per;
        g.alpha = 1.0; g.nb = nb;
        launch_zgemm(s, g);
        HIPCHK(hipGetLastError());
      }
      if (mixed) {
        // A_bar = L_exp(A, Ebar^H)^H: the pair (A, sigma Ebar^H) from A rebuilt unscaled in FP32
        float2* PW = c->w.aW2.as<float2>();
        float* pm = c->w.aMax.as<float>();
        launch_fill_zero(s, pm, 2 * sizeof(float) * (size_t)nb);
        launch_assemble32(s, B, V, R, L, kn, p->cr->mlambda, p->cr->dynupz, p->dz, PW);
        launch_adj_pair32(s, npad, nb, PW, c->w.aEbar.as<double2>(), kThetaSplit, kSmax, c->w.sq.as<int32_t>(),
                          p->err.as<int32_t>(), c->w.aSig.as<double>(), PW + per * nb, pm, pm + nb);
        HIPCHK(hipGetLastError());
        float2* G = expm_pair_mixed(s, npad, nb, PW, pm, c->w.sq.as<int32_t>(), p->err.as<int32_t>(), h_sq2, rc);
        if (!G) return rc;
        launch_adj_extract32(s, B, V, kn, p->dz, G, c->w.aSig.as<double>(), c->w.aDbar.as<double2>());
      } else {
        // A_bar = L_exp(A^H, Ebar), the upper-right block of expm([[A^H, sigma Ebar], [0, A^H]]) taken in
        // the pair form; A rebuilt unscaled
        double2* A = c->w.A.as<double2>();
        launch_assemble(s, B, V, R, L, kn, p->cr->mlambda, p->cr->dynupz, p->dz, A);
        double2* PX = c->w.aBlk.as<double2>();
        double2* PY = PX + per * nb;
        launch_adj_pair(s, npad, nb, A, c->w.aEbar.as<double2>(), c->w.aSig.as<double>(), PX, PY);
        HIPCHK(hipGetLastError());
        double2* T2 = expm_pair_fp64(s, npad, nb, PX, PY, c->w.aW2.as<double2>(), c->w.sq.as<int32_t>(),
                                     p->err.as<int32_t>(), h_sq2, rc);
        if (!T2) return rc;
        launch_adj_extract(s, B, V, kn, p->dz, T2 + 5 * per * nb, npad, c->w.aSig.as<double>(),
                           c->w.aDbar.as<double2>());
      }
      HIPCHK(hipGetLastError());
      // LUT entries
      launch_adj_lut(s, B, V, R, L, p->cr->mlambda, c->w.aDbar.as<double2>(), c->w.g_lut.as<double>(), max_nw, nrep);
      HIPCHK(hipGetLastError());
    }
    off += nbk;
  }
  std::vector<double> d2((size_t)nk * ns);
  launch_adj_lut_reduce(s, p->cr->ncell, nrep, c->w.g_lut.as<double>());
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(glut, c->w.g_lut.p, 16 * (size_t)p->cr->ncell, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(d2.data(), c->w.g_dots2.p, 8 * d2.size(), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if ((rc = check_err_flag(p->err, s, who))) return rc;
  for (int q = 0; q < p->cr->nsite; q++) {
    double acc = 0.0;
    for (int i = 0; i < nk; i++) acc += gdots[(size_t)i * p->cr->nsite + q] * d2[(size_t)i * p->cr->nsite + q];
    gb[q] = acc;
  }
  return CTEM_OK;
}

extern "C" int32_t ctem_plan_adjoint(ctem_plan* p, const double* gdots, double* glut, double* gb) {
  return plan_adjoint(p, false, gdots, glut, gb);
}

extern "C" int32_t ctem_plan_adjoint_mixed(ctem_plan* p, const double* gdots, double* glut, double* gb) {
  return plan_adjoint(p, true, gdots, glut, gb);
}

extern "C" int32_t ctem_plan_dynmat(ctem_plan* p, int32_t ik, double* dynmat, int32_t cap, int32_t* n) {
  if (!p || !n) return fail(CTEM_E_ARG, "null argument");
  int32_t rc;
  if (!p->selected && (rc = ctem_plan_select(p, nullptr, nullptr))) return rc;
  if (ik < 0 || ik >= p->numk) return fail(CTEM_E_ARG, "direction out of range");
  ctem_ctx* c = p->ctx;
  hipStream_t s = c->stream;
  const int nn = p->h_nns[ik];
  *n = nn;
  if (!dynmat) return CTEM_OK;
  if (cap < nn * nn) return fail(CTEM_E_ARG, "capacity too small");
  const int npad = npad_of(nn);
  if ((rc = c->s_A.ensure(sizeof(double2) * (size_t)npad * npad)) || (rc = c->s_dirs.ensure(4)) ||
      (rc = c->s_err.ensure(4)))
    return rc;
  HIPCHK(hipMemcpyAsync(c->s_dirs.p, &ik, 4, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemsetAsync(c->s_err.p, 0, 4, s));
  BeamView V{p->nns.as<int32_t>(), p->nnw.as<int32_t>(), p->soff.as<int64_t>(), p->woff.as<int64_t>(),
             p->sidx.as<int32_t>(), p->ssg.as<float>(), p->widx.as<int32_t>(), p->wsg.as<float>()};
  Bucket B{npad, 1, c->s_dirs.as<int32_t>()};
  launch_dynmat(s, B, V, p->cr->refl(), p->cr->lutv(c->s_err.as<int32_t>()), p->cr->mlambda, p->cr->dynupz, c->s_A.as<double2>());
  HIPCHK(hipGetLastError());
  if ((rc = check_err_flag(c->s_err, s, "ctem_plan_dynmat"))) return rc;
  std::vector<double2> h((size_t)npad * npad);
  HIPCHK(hipMemcpyAsync(h.data(), c->s_A.p, h.size() * sizeof(double2), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int i = 0; i < nn; i++)
    for (int j = 0; j < nn; j++) {
      dynmat[2 * (i * nn + j)] = h[(size_t)j * npad + i].x;
      dynmat[2 * (i * nn + j) + 1] = h[(size_t)j * npad + i].y;
    }
  return CTEM_OK;
}

extern "C" int32_t ctem_plan_scatter(ctem_plan* p, const int32_t* ij, int32_t npx, int32_t npy,
                                     int32_t isym, float* sr) {
  if (!p || !p->ran) return fail(CTEM_E_STATE, "ctem_plan_scatter: run the plan first");
  if (!ij || !sr || npx < 0 || npy < 0) return fail(CTEM_E_ARG, "bad arguments");
  ctem_ctx* c = p->ctx;
  hipStream_t s = c->stream;
  if (isym < 1 || isym > 13) return fail(CTEM_E_ARG, "ctem_plan_scatter: Laue group %d not in 1..13", isym);
  const size_t plane = (size_t)(2 * npx + 1) * (2 * npy + 1);
  const size_t tot = plane * std::max(1, p->cr->nsite);
  DBuf d_sr, d_owner;
  int32_t rc;
  if ((rc = upload(p->ij, ij, (size_t)2 * p->numk, s)) || (rc = d_sr.ensure(4 * tot)) ||
      (rc = d_owner.ensure(4 * plane)))
    return rc;
  HIPCHK(hipMemcpyAsync(d_sr.p, sr, 4 * tot, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemsetAsync(d_owner.p, 0xff, 4 * plane, s));
  launch_scatter(s, p->numk, p->ij.as<int32_t>(), p->dots.as<double>(), p->cr->nsite, p->cr->nat, npx, npy,
                 isym, d_owner.as<int32_t>(), d_sr.as<float>());
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(sr, d_sr.p, 4 * tot, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  d_sr.release();
  d_owner.release();
  if (e != hipSuccess) return fail(CTEM_E_HIP, "scatter: %s", hipGetErrorString(e));
  return CTEM_OK;
}

extern "C" int32_t ctem_plan_timing(ctem_plan* p, double* total_ms, double* gemm_ms, double* gemm_flops,
                                    int32_t* gemm_launches) {
  if (!p || !p->ran) return fail(CTEM_E_STATE, "ctem_plan_timing: run the plan first");
  if (total_ms) *total_ms = p->t_total;
  if (gemm_ms) *gemm_ms = p->t_gemm;
  if (gemm_flops) *gemm_flops = p->f_gemm;
  if (gemm_launches) *gemm_launches = p->n_gemm;
  return CTEM_OK;
}

extern "C" int32_t ctem_plan_alg_flops(ctem_plan* p, double* total, double* products, double* solve) {
  if (!p || p->fa_total < 0) return fail(CTEM_E_STATE, "ctem_plan_alg_flops: run the plan in FP64 first");
  if (total) *total = p->fa_total;
  if (products) *products = p->fa_prod;
  if (solve) *solve = p->fa_solve;
  return CTEM_OK;
}

// ------------------------------------------------------------------------------------------
// CalcLgh3 drop-in (one direction, FP64 on the GPU)
namespace {
__global__ void k_scale_dmat(int nn, int npad, const double2* __restrict__ D, double pre,
                             double2* __restrict__ A) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)npad * npad) return;
  const int i = (int)(e % npad), j = (int)(e / npad);
  double2 v = make_double2(0.0, 0.0);
  if (i < nn && j < nn) {
    const double2 d = D[(int64_t)j * nn + i];
    v = make_double2(-pre * d.y, pre * d.x);
  }
  A[e] = v;
}
}  // namespace

namespace {
__global__ void k_fill_zero(uint32_t* __restrict__ p, int64_t n) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) p[e] = 0u;
}
__global__ void k_widen_lam(int n, const float* __restrict__ f, double* __restrict__ d) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) d[e] = (double)f[e];
}
// row-major nn x nn -> column-major
__global__ void k_transpose_c(int nn, const double2* __restrict__ R, double2* __restrict__ Cm) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nn * nn) return;
  const int i = (int)(e % nn), j = (int)(e / nn);  // Cm(i, j) at j*nn + i
  Cm[e] = R[(int64_t)i * nn + j];
}
}  // namespace

extern "C" int32_t ctem_calclgh3_static(ctem_ctx* c, const double* dmat, double* lgh, double thick, double kn,
                                        int32_t nn, int32_t gzero, double depthstep, const float* lambdae,
                                        int32_t izz) {
  if (!c || !dmat || !lgh || !lambdae) return fail(CTEM_E_ARG, "ctem_calclgh3_static: null argument");
  if (nn < 1 || izz < 1 || !(thick > 0) || !(depthstep > 0) || !(kn > 0))
    return fail(CTEM_E_ARG, "ctem_calclgh3_static: bad sizes");
  if (gzero != 1) return fail(CTEM_E_RANGE, "ctem_calclgh3_static: only gzero = 1 (the transmitted beam first) is supported");
  if (nn > kStaticMaxN || izz > kStaticMaxZ)
    return fail(CTEM_E_RANGE, "ctem_calclgh3_static: at most %d strong beams and %d depth steps (got %d, %d); "
                "use ctem_calclgh3", kStaticMaxN, kStaticMaxZ, nn, izz);
  if (!c->scratch) return fail(CTEM_E_STATE, "ctem_calclgh3_static: no work space on this context");
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const int npad = npad_of(nn);
  const int64_t per = (int64_t)npad * npad;
  const int64_t spsi = (int64_t)npad * izz;
  // carve the work space (all offsets multiples of 256 bytes)
  unsigned char* base = c->scratch;
  double2* A = (double2*)base;                          // 1 matrix
  double2* W = (double2*)(base + kStaticMat);           // 5 matrices
  double2* Psi = (double2*)(base + 6 * kStaticMat);     // npad x izz <= 1 matrix
  double2* Lr = (double2*)(base + 7 * kStaticMat);      // DynMat in, then Lgh (row-major)
  unsigned char* small = base + 8 * kStaticMat;         // 64 KiB of small buffers
  int32_t* d_sq = (int32_t*)(small);
  int32_t* d_err = (int32_t*)(small + 256);
  int32_t* d_dirs = (int32_t*)(small + 512);
  int32_t* d_nns = (int32_t*)(small + 768);
  int64_t* d_off = (int64_t*)(small + 1024);
  double* d_dots = (double*)(small + 1280);
  int32_t* d_zero = (int32_t*)(small + 2048);           // 4 npad + 16 bytes <= 1040
  double* d_part = (double*)(small + 4096);             // 8 (npad/32)^2 <= 512 bytes
  float* d_lamf = (float*)(small + 8192);               // 4 izz <= 1 KiB
  double* d_lam = (double*)(small + 12288);             // 8 izz <= 2 KiB
  // the zeroed small buffers: counts, flags, the zero index list, partial sums
  hipLaunchKernelGGL(k_fill_zero, dim3(16), dim3(256), 0, s, (uint32_t*)small, (int64_t)(8192 / 4));
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(d_lamf, lambdae, 4 * (size_t)izz, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_widen_lam, dim3((izz + 255) / 256), dim3(256), 0, s, izz, (const float*)d_lamf, d_lam);
  HIPCHK(hipMemcpyAsync(Lr, dmat, 16 * (size_t)nn * nn, hipMemcpyHostToDevice, s));
  const double pre = 2.0 * kPi * depthstep / (double)(float)(2.0 * kn);
  hipLaunchKernelGGL(k_scale_dmat, dim3((unsigned)((per + 255) / 256)), dim3(256), 0, s, nn, npad,
                     (const double2*)Lr, pre, A);
  HIPCHK(hipGetLastError());
  std::vector<int> hn{nn};
  std::vector<int32_t> hsq;
  int32_t rc = expm_depth_fp64(nullptr, s, npad, 1, A, W, Psi, spsi, izz, d_sq, d_err, hn, hsq, c->cu_count);
  if (rc) return rc;
  const int32_t nnv[2] = {nn, 0};
  HIPCHK(hipMemcpyAsync(d_nns, nnv, 8, hipMemcpyHostToDevice, s));
  BeamView V{d_nns, d_nns + 1, d_off, d_off + 1, d_zero, nullptr, nullptr, nullptr};
  DevRefl R{};
  R.hkl = d_zero;
  DevLut L{};
  L.nsite = 0;
  L.dim[0] = L.dim[1] = L.dim[2] = 1;
  Bucket B{npad, 1, d_dirs};
  launch_lgh_dots(s, B, V, R, L, Psi, spsi, izz, d_lam, depthstep / thick, d_part, d_dots, 0, Lr, 0);
  HIPCHK(hipGetLastError());
  // column-major for the Fortran caller, into the first work matrix (free after the depth stage)
  hipLaunchKernelGGL(k_transpose_c, dim3((unsigned)(((int64_t)nn * nn + 255) / 256)), dim3(256), 0, s, nn,
                     (const double2*)Lr, W);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(lgh, W, 16 * (size_t)nn * nn, hipMemcpyDeviceToHost, s));
  int32_t h = 0;
  HIPCHK(hipMemcpyAsync(&h, d_err, sizeof h, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (h & 2) return fail(CTEM_E_NUMERIC, "ctem_calclgh3_static: non-finite dynamical matrix");
  if (h & 4) return fail(CTEM_E_NUMERIC, "ctem_calclgh3_static: matrix norm needs more squarings than allowed");
  if (h) return fail(CTEM_E_NUMERIC, "ctem_calclgh3_static: error flag %d", h);
  return CTEM_OK;
}

extern "C" int32_t ctem_calclgh3(ctem_ctx* c, const double* dmat, double* lgh, double thick, double kn,
                                 int32_t nn, int32_t gzero, double depthstep, const float* lambdae,
                                 int32_t izz) {
  if (!c || !dmat || !lgh || !lambdae) return fail(CTEM_E_ARG, "ctem_calclgh3: null argument");
  if (nn < 1 || izz < 1 || !(thick > 0) || !(depthstep > 0) || !(kn > 0))
    return fail(CTEM_E_ARG, "ctem_calclgh3: bad sizes");
  if (gzero != 1) return fail(CTEM_E_RANGE, "ctem_calclgh3: only gzero = 1 (the transmitted beam first) is supported");
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const int npad = npad_of(nn);
  const int64_t per = (int64_t)npad * npad;
  const int64_t spsi = (int64_t)npad * izz;
  int32_t rc;
  if ((rc = c->s_A.ensure(per * sizeof(double2))) || (rc = c->s_W.ensure(5 * per * sizeof(double2))) ||
      (rc = c->s_Psi.ensure(spsi * sizeof(double2))) || (rc = c->s_sq.ensure(4)) || (rc = c->s_err.ensure(4)) ||
      (rc = c->s_lgh.ensure(std::max<size_t>(16 * (size_t)nn * nn, per * sizeof(double2)))) ||
      (rc = c->s_dirs.ensure(4)) || (rc = c->s_nns.ensure(8)) || (rc = c->s_zero.ensure(4 * (size_t)npad + 16)) ||
      (rc = c->s_off.ensure(16)) || (rc = c->s_part.ensure(8 * (size_t)(npad / 32) * (npad / 32))) ||
      (rc = c->s_dots.ensure(8)))
    return rc;
  std::vector<double> lamd(lambdae, lambdae + izz);
  if ((rc = upload(c->s_lam, lamd.data(), lamd.size(), s))) return rc;
  HIPCHK(hipMemcpyAsync(c->s_lgh.p, dmat, 16 * (size_t)nn * nn, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemsetAsync(c->s_err.p, 0, 4, s));
  const double pre = 2.0 * kPi * depthstep / (double)(float)(2.0 * kn);
  hipLaunchKernelGGL(k_scale_dmat, dim3((unsigned)((per + 255) / 256)), dim3(256), 0, s, nn, npad,
                     c->s_lgh.as<double2>(), pre, c->s_A.as<double2>());
  HIPCHK(hipGetLastError());
  std::vector<int> hn{nn};
  std::vector<int32_t> hsq;
  rc = expm_depth_fp64(nullptr, s, npad, 1, c->s_A.as<double2>(), c->s_W.as<double2>(), c->s_Psi.as<double2>(),
                       spsi, izz, c->s_sq.as<int32_t>(), c->s_err.as<int32_t>(), hn, hsq, c->cu_count);
  if (rc) return rc;
  // Lgh through the contraction kernel with no sites: a one-entry plan view
  const int32_t zero4[4] = {0, 0, 0, 0};
  const int32_t nnv[2] = {nn, 0};
  const int64_t offs[2] = {0, 0};
  HIPCHK(hipMemsetAsync(c->s_zero.p, 0, 4 * (size_t)npad + 16, s));
  HIPCHK(hipMemcpyAsync(c->s_dirs.p, zero4, 4, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(c->s_nns.p, nnv, 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(c->s_off.p, offs, 16, hipMemcpyHostToDevice, s));
  BeamView V{c->s_nns.as<int32_t>(), c->s_nns.as<int32_t>() + 1, c->s_off.as<int64_t>(), c->s_off.as<int64_t>() + 1,
             c->s_zero.as<int32_t>(), nullptr, nullptr, nullptr};
  DevRefl R{};
  R.hkl = c->s_zero.as<int32_t>();
  DevLut L{};
  L.nsite = 0;
  L.dim[0] = L.dim[1] = L.dim[2] = 1;
  Bucket B{npad, 1, c->s_dirs.as<int32_t>()};
  launch_lgh_dots(s, B, V, R, L, c->s_Psi.as<double2>(), spsi, izz, c->s_lam.as<double>(), depthstep / thick,
                  c->s_part.as<double>(), c->s_dots.as<double>(), 0, c->s_lgh.as<double2>(), 0);
  HIPCHK(hipGetLastError());
  std::vector<double> h(2 * (size_t)nn * nn);
  HIPCHK(hipMemcpyAsync(h.data(), c->s_lgh.p, 16 * (size_t)nn * nn, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if ((rc = check_err_flag(c->s_err, s, "ctem_calclgh3"))) return rc;
  // row-major from the kernel -> column-major for the Fortran caller
  for (int i = 0; i < nn; i++)
    for (int j = 0; j < nn; j++) {
      lgh[2 * ((size_t)j * nn + i)] = h[2 * ((size_t)i * nn + j)];
      lgh[2 * ((size_t)j * nn + i) + 1] = h[2 * ((size_t)i * nn + j) + 1];
    }
  return CTEM_OK;
}

// ------------------------------------------------------------------------------------------
// electron channeling patterns (include/ctem_ecp.h)

// The channeling structure matrix Sgh(d) for every d = g-h of the crystal's difference box, all
// sites in one table (CTEMECP.f90:318-350): per site Znsq = float(Z)**2 * occ and
// DBWF = Znsq * exp(-B kkl) with kkl = 0.25 CalcLength(float(d), 'r')**2, all float32; per
// orbit position the float32 complex product cmplx(DBWF, 0) * cmplx(cos(arg), sin(arg)) with
// arg = tpi * sum(d * r) (tpi = fl32(2 pi), the sum in double), added in double site after
// site. *s0 receives Sgh(000).
static void build_slut_ecp(const Crystal& cr, const float* occ, std::vector<double>& out, double* s0) {
#pragma clang fp contract(off)
  const int32_t* lo = cr.slut_lo;
  const int32_t* dim = cr.slut_dim;
  const size_t ncell = (size_t)dim[0] * dim[1] * dim[2];
  out.assign(2 * ncell, 0.0);
  const float tpi = (float)(2.0 * kPi);
  std::vector<float> znsq(cr.nsite);
  for (int p = 0; p < cr.nsite; p++) {
    const float zf = (float)cr.h_site_z[p];
    znsq[p] = (zf * zf) * occ[p];
  }
  for (int l = 0; l < dim[2]; l++)
    for (int kk = 0; kk < dim[1]; kk++)
      for (int h = 0; h < dim[0]; h++) {
        const int d[3] = {h + lo[0], kk + lo[1], l + lo[2]};
        const float df[3] = {(float)d[0], (float)d[1], (float)d[2]};
        const float len = std::sqrt(dot_r(df, df, cr.h_rmt));
        const float kkl = 0.25f * (len * len);
        double sr = 0.0, si = 0.0;
        size_t obase = 0;
        for (int p = 0; p < cr.nsite; p++) {
          const float dbwf = znsq[p] * ::expf(-cr.h_site_b[p] * kkl);
          const double* orb = cr.h_orbit.data() + 3 * obase;
          for (int a = 0; a < cr.h_norbit[p]; a++) {
            double sd = 0.0;
            for (int t = 0; t < 3; t++) sd += (double)d[t] * orb[3 * a + t];
            const double arg = (double)tpi * sd;
            double sn, cs;
            ::sincos(arg, &sn, &cs);
            sr += (double)(dbwf * (float)cs);
            si += (double)(dbwf * (float)sn);
          }
          obase += cr.h_norbit[p];
        }
        const size_t cell = (size_t)h + (size_t)dim[0] * ((size_t)kk + (size_t)dim[1] * l);
        out[2 * cell] = sr;
        out[2 * cell + 1] = si;
        if (d[0] == 0 && d[1] == 0 && d[2] == 0) *s0 = sr;
      }
}

// E = MatrixExponential(A, z0, 'Pade') (math.f90:187-315) for the nb matrices of A (holding
// A z0, overwritten with the scaled B); W holds 5 nb npad^2 entries of work space, d_q and d_j nb
// integers each. Returns the slice of W that holds E.
static double2* pade_fp64(ctem_plan* tp, hipStream_t s, int npad, int nb, double2* A, double2* W, int32_t* d_q,
                          int32_t* d_j, int32_t* d_err, int32_t& rc_out) {
  rc_out = CTEM_OK;
  const int64_t per = (int64_t)npad * npad;
  double2* C0 = W;
  double2* C1 = W + per * nb;
  double2* N = W + 2 * per * nb;
  double2* D = W + 3 * per * nb;
  double2* T = W + 4 * per * nb;
  auto prod = [&](double2* C, const double2* X, const double2* Y, const int32_t* active, int step) -> bool {
    ZGemm g{};
    g.m = g.n = g.k = npad;
    g.ldx = g.ldy = g.ldc = g.ldm = npad;
    g.X = X; g.sx = per; g.Y = Y; g.sy = per; g.C = C; g.sc = per; g.sm = per;
    g.alpha = 1.0; g.betaI = 0.0;
    g.active = active; g.step = step; g.nb = nb;
    GemmTimer t{tp, s};
    t.begin();
    launch_zgemm(s, g);
    t.end();
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { rc_out = fail(CTEM_E_HIP, "Pade product: %s", hipGetErrorString(e)); return false; }
    return true;
  };
  launch_pade_prep(s, npad, nb, A, N, D, C0, d_q, d_j, d_err);
  double2* cb[2] = {C0, C1};
  for (int icnt = 1; icnt <= 4; icnt++) {
    launch_pade_accum(s, npad, nb, icnt, d_q, cb[(icnt - 1) & 1], N, D);
    // C_{icnt+1} = C_icnt B for the matrices with icnt < q (active = q - 1, step = icnt - 1)
    if (icnt < 4 && !prod(cb[icnt & 1], cb[(icnt - 1) & 1], A, d_q, icnt - 1)) return nullptr;
  }
  launch_zsolve(s, npad, nb, D, N, d_err);
  std::vector<int32_t> hj(nb);
  hipError_t e = hipMemcpyAsync(hj.data(), d_j, 4 * (size_t)nb, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) { rc_out = fail(CTEM_E_HIP, "Pade squaring counts: %s", hipGetErrorString(e)); return nullptr; }
  int jmax = 0;
  for (int b = 0; b < nb; b++) jmax = std::max(jmax, hj[b]);
  double2* buf[2] = {N, T};
  for (int i = 0; i < jmax; i++)
    if (!prod(buf[(i + 1) & 1], buf[i & 1], buf[i & 1], d_j, i)) return nullptr;
  launch_parity_copy(s, npad, nb, d_j, T, N);
  return N;
}

extern "C" int32_t ctem_ecp_plan_create(ctem_ctx* c, int32_t numk, const double* k, const double* kn,
                                        const ctem_ecp_params* prm, ctem_plan** out) {
#pragma clang fp contract(off)
  if (!c || !prm || !out) return fail(CTEM_E_ARG, "ctem_ecp_plan_create: null argument");
  *out = nullptr;
  if (!c->cr) return fail(CTEM_E_STATE, "ctem_ecp_plan_create: no crystal uploaded");
  const Crystal& cr = *c->cr;
  if (prm->mode != CTEM_ECP_BLOCHWV && prm->mode != CTEM_ECP_SCATMAT)
    return fail(CTEM_E_ARG, "ctem_ecp_plan_create: unknown mode %d", prm->mode);
  if (prm->numthick < 1 || prm->numthick > 4096) return fail(CTEM_E_ARG, "ctem_ecp_plan_create: 1..4096 thicknesses");
  if (cr.nsite < 1 || !prm->site_occ) return fail(CTEM_E_ARG, "ctem_ecp_plan_create: site occupancies needed");
  for (int p = 0; p < cr.nsite; p++)
    if (!std::isfinite(prm->site_occ[p])) return fail(CTEM_E_ARG, "ctem_ecp_plan_create: non-finite occupancy");
  for (int i = 0; i < 3; i++)
    if (!std::isfinite(prm->fn[i])) return fail(CTEM_E_ARG, "ctem_ecp_plan_create: non-finite foil normal");
  const int nt = prm->numthick;
  // thick = startthick + thickinc * (/ (i-1, i=1,nt) /) (CTEMECP.f90:270), float32
  std::vector<float> thick(nt);
  for (int i = 0; i < nt; i++) thick[i] = prm->startthick + prm->thickinc * (float)i;
  std::vector<int32_t> store;
  int numt = 0;
  if (prm->mode == CTEM_ECP_BLOCHWV) {
    if (!(prm->startthick > 0.0f) || !(prm->thickinc > 0.0f) || !std::isfinite(thick[nt - 1]))
      return fail(CTEM_E_ARG, "ctem_ecp_plan_create: Blochwv needs startthick > 0 and thickinc > 0");
  } else {
    if (!(prm->zintstep > 0.0f)) return fail(CTEM_E_ARG, "ctem_ecp_plan_create: zintstep must be positive");
    // tvals = nint(thick/zintstep), numt = nint(maxval(thick)/zintstep); step i stores the running
    // sum at thickness tval when i == tvals(tval), then tval advances (CTEMECP.f90:596-616)
    float tmax = thick[0];
    for (int i = 1; i < nt; i++) tmax = std::max(tmax, thick[i]);
    const float qmax = tmax / prm->zintstep;
    if (!(qmax < 65536.5f)) return fail(CTEM_E_RANGE, "ctem_ecp_plan_create: more than 65536 ScatMat steps");
    numt = (int)std::lround(qmax);
    if (numt < 1) return fail(CTEM_E_ARG, "ctem_ecp_plan_create: no ScatMat step below the largest thickness");
    std::vector<int> tvals(nt);
    for (int i = 0; i < nt; i++) tvals[i] = (int)std::lround(thick[i] / prm->zintstep);
    store.assign(numt, -1);
    int tval = 0;
    for (int i = 1; i <= numt; i++)
      if (tval < nt && i == tvals[tval]) store[i - 1] = tval++;
  }
  static const float one = 1.0f;
  int32_t rc = ctem_plan_create(c, numk, k, kn, 1, &one, 1.0, 1.0, out);
  if (rc) return rc;
  ctem_plan* p = *out;
  p->ecp_mode = prm->mode;
  p->ecp_nt = nt;
  p->ecp_numt = numt;
  p->ecp_t0 = prm->startthick;
  p->ecp_dt = prm->thickinc;
  p->ecp_zint = prm->zintstep;
  int nat = 0;
  for (int q = 0; q < cr.nsite; q++) nat += cr.h_norbit[q];
  p->ecp_nat = (double)(float)nat;
  std::vector<double> slut;
  build_slut_ecp(cr, prm->site_occ, slut, &p->ecp_s0);
  hipStream_t s = c->stream;
  if ((rc = upload(p->fn, prm->fn, 3, s)) || (rc = upload(p->ecp_slut, slut.data(), slut.size(), s)) ||
      (rc = upload(p->ecp_thick, thick.data(), thick.size(), s)) ||
      (numt > 0 && (rc = upload(p->ecp_store, store.data(), store.size(), s))) ||
      (rc = p->ecp_sr.ensure(sizeof(float) * (size_t)nt * std::max(numk, 1)))) {
    ctem_plan_destroy(p);
    *out = nullptr;
    return rc;
  }
  if (hipStreamSynchronize(s) != hipSuccess) {
    ctem_plan_destroy(p);
    *out = nullptr;
    return fail(CTEM_E_HIP, "ECP plan upload failed");
  }
  return CTEM_OK;
}

extern "C" int32_t ctem_ecp_run(ctem_plan* p) {
  if (!p) return fail(CTEM_E_ARG, "null plan");
  if (p->ecp_mode < 0) return fail(CTEM_E_STATE, "ctem_ecp_run: not a channeling-pattern plan");
  ctem_ctx* c = p->ctx;
  int32_t rc;
  if (!p->selected && (rc = ctem_plan_select(p, nullptr, nullptr))) return rc;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const bool bloch = p->ecp_mode == CTEM_ECP_BLOCHWV;
  const int nt = p->ecp_nt, numt = p->ecp_numt;
  for (auto& kv : p->buckets)
    if (bloch && kv.first > 512) return fail(CTEM_E_RANGE, "ctem_ecp_run: Blochwv supports up to 512 strong beams");
  // work space per direction: ScatMat A, 5 matrices of the exponential, Psi at every step and the
  // per-tile depth sums; Blochwv the assembled A, the two block matrices of size 2 npad and the
  // 5 x 2 matrices of their exponentials
  auto per_dir = [&](int npad) -> size_t {
    const size_t per = (size_t)npad * npad;
    if (bloch) return (per + 8 * per + 40 * per) * sizeof(double2) + 64;
    const size_t nzp = (size_t)(numt + 31) / 32 * 32;
    return (6 * per + (size_t)npad * numt) * sizeof(double2) + (size_t)(npad / 32) * nzp * 8 + 64;
  };
  size_t a = 0, w = 0, ps = 0, sq = 0, part = 0;
  for (auto& kv : p->buckets) {
    const int npad = kv.first, nbk = (int)kv.second.size();
    const size_t per = (size_t)npad * npad;
    const size_t nb = (size_t)std::max(1, (int)std::min<size_t>(nbk, std::max<size_t>(1, chunk_budget(p) / per_dir(npad))));
    if (bloch) {
      a = std::max(a, 8 * per * nb * sizeof(double2));
      w = std::max(w, 40 * per * nb * sizeof(double2));
      ps = std::max(ps, per * nb * sizeof(double2));
      sq = std::max(sq, 8 * nb);
    } else {
      const size_t nzp = (size_t)(numt + 31) / 32 * 32;
      a = std::max(a, per * nb * sizeof(double2));
      w = std::max(w, 5 * per * nb * sizeof(double2));
      ps = std::max(ps, (size_t)npad * numt * nb * sizeof(double2));
      sq = std::max(sq, 8 * nb);
      part = std::max(part, (size_t)(npad / 32) * nzp * nb * 8);
    }
  }
  const void* before[4] = {c->w.A.p, c->w.W.p, c->w.Psi.p, c->w.sq.p};
  if ((rc = c->w.A.ensure(a)) || (rc = c->w.W.ensure(w)) || (rc = c->w.Psi.ensure(ps)) || (rc = c->w.sq.ensure(sq)) ||
      (part && (rc = p->ecp_part.ensure(part))))
    return rc;
  const void* after[4] = {c->w.A.p, c->w.W.p, c->w.Psi.p, c->w.sq.p};
  for (int i = 0; i < 4; i++)
    if (before[i] != after[i]) c->wgen++;
  reset_timing(p);
  HIPCHK(hipEventRecord(c->ev[0], s));
  launch_fill_zero(s, p->err.p, 4);
  launch_fill_zero(s, p->ecp_sr.p, sizeof(float) * (size_t)nt * std::max(p->numk, 1));
  BeamView V{p->nns.as<int32_t>(), p->nnw.as<int32_t>(), p->soff.as<int64_t>(), p->woff.as<int64_t>(),
             p->sidx.as<int32_t>(), p->ssg.as<float>(), p->widx.as<int32_t>(), p->wsg.as<float>()};
  DevRefl R = p->cr->refl();
  DevLut L = p->cr->lutv(p->err.as<int32_t>());
  DevLut LE = L;
  LE.slut = p->ecp_slut.as<double2>();
  LE.nsite = 1;
  size_t off = 0;
  for (auto& kv : p->buckets) {
    const int npad = kv.first;
    const int nbk = (int)kv.second.size();
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>(nbk, chunk_budget(p) / per_dir(npad)));
    for (int c0 = 0; c0 < nbk; c0 += chunk) {
      const int nb = std::min(chunk, nbk - c0);
      Bucket B{npad, nb, p->dirs.as<int32_t>() + off + c0};
      std::vector<int> hn;
      std::vector<int32_t> hsq;
      if (!bloch) {
        for (int b = 0; b < nb; b++) hn.push_back(p->h_nns[kv.second[c0 + b]]);
        // Minp = DynMat * dcmplx(0, cPi mLambda), Azz = MatrixExponential(Minp, zintstep, 'Pade')
        launch_assemble_scatmat(s, B, V, R, L, p->cr->mlambda, p->cr->dynupz, (double)p->ecp_zint, c->w.A.as<double2>());
        HIPCHK(hipGetLastError());
        double2* E = pade_fp64(p, s, npad, nb, c->w.A.as<double2>(), c->w.W.as<double2>(), c->w.sq.as<int32_t>(),
                               c->w.sq.as<int32_t>() + nb, p->err.as<int32_t>(), rc);
        if (!E) return rc;
        const int64_t spsi = (int64_t)npad * numt;
        launch_depth(s, npad, nb, numt, E, 0, 0, c->w.Psi.as<double2>(), spsi, c->cu_count);
        launch_ecp_depth_dots(s, B, V, R, LE, c->w.Psi.as<double2>(), spsi, numt, p->ecp_part.as<double>());
        launch_ecp_scatmat_out(s, B, numt, p->ecp_part.as<double>(), p->ecp_store.as<int32_t>(), p->ecp_nat, p->numk,
                               p->ecp_sr.as<float>());
        HIPCHK(hipGetLastError());
      } else {
        for (int b = 0; b < nb; b++) hn.push_back(2 * p->h_nns[kv.second[c0 + b]]);
        for (int b = 0; b < nb; b++) hn.push_back(2 * p->h_nns[kv.second[c0 + b]]);
        // A1 = 2 pi i DynMat / fl32(2 kn); eps scales Sgh^T to entries of order 1/npad
        launch_assemble(s, B, V, R, L, p->kn.as<double>(), p->cr->mlambda, p->cr->dynupz, 1.0, c->w.Psi.as<double2>());
        const double eps = std::ldexp(1.0, -std::ilogb(std::max(p->ecp_s0, 1e-300) * (double)npad));
        launch_ecp_augment(s, B, V, R, LE, c->w.Psi.as<double2>(), eps, (double)p->ecp_t0, (double)p->ecp_dt,
                           c->w.A.as<double2>());
        HIPCHK(hipGetLastError());
        double2* E = expm_fp64(p, s, 2 * npad, 2 * nb, c->w.A.as<double2>(), c->w.W.as<double2>(), c->w.sq.as<int32_t>(),
                               p->err.as<int32_t>(), hn, hsq, rc);
        if (!E) return rc;
        launch_ecp_bloch_steps(s, B, V, E, nt, p->ecp_thick.as<float>(), eps, p->ecp_nat, p->numk, p->ecp_sr.as<float>());
        HIPCHK(hipGetLastError());
      }
    }
    off += nbk;
  }
  HIPCHK(hipEventRecord(c->ev[1], s));
  HIPCHK(hipStreamSynchronize(s));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  p->t_total = ms;
  for (size_t i = 0; i + 1 < p->evused; i += 2) {
    float g = 0.0f;
    HIPCHK(hipEventElapsedTime(&g, p->evpool[i], p->evpool[i + 1]));
    p->t_gemm += g;
  }
  if ((rc = check_err_flag(p->err, s, "ctem_ecp_run"))) return rc;
  p->ran = true;
  return CTEM_OK;
}

extern "C" int32_t ctem_ecp_get_sr(ctem_plan* p, float* sr) {
  if (!p || !sr) return fail(CTEM_E_ARG, "ctem_ecp_get_sr: null argument");
  if (p->ecp_mode < 0 || !p->ran) return fail(CTEM_E_STATE, "ctem_ecp_get_sr: run ctem_ecp_run first");
  const size_t n = (size_t)p->ecp_nt * p->numk;
  if (n) {
    HIPCHK(hipMemcpyAsync(sr, p->ecp_sr.p, sizeof(float) * n, hipMemcpyDeviceToHost, p->ctx->stream));
    HIPCHK(hipStreamSynchronize(p->ctx->stream));
  }
  return CTEM_OK;
}

extern "C" int32_t ctem_ecp_image(ctem_plan* p, const int32_t* ij, int32_t npx, float* img) {
  if (!p || !img || (p->numk > 0 && !ij)) return fail(CTEM_E_ARG, "ctem_ecp_image: null argument");
  if (npx < 0 || npx > 4096) return fail(CTEM_E_ARG, "ctem_ecp_image: npx must be 0..4096");
  const int64_t w = 2 * (int64_t)npx + 1;
  std::vector<float> sr((size_t)p->ecp_nt * p->numk);
  int32_t rc;
  if ((rc = ctem_ecp_get_sr(p, sr.data()))) return rc;
  std::fill(img, img + (size_t)p->ecp_nt * w * w, 0.0f);
  for (int ik = 0; ik < p->numk; ik++) {
    const int64_t x = ij[2 * ik] + npx, y = ij[2 * ik + 1] + npx;
    if (x < 0 || x >= w || y < 0 || y >= w) return fail(CTEM_E_RANGE, "ctem_ecp_image: direction %d outside the image", ik);
    for (int it = 0; it < p->ecp_nt; it++) img[((size_t)it * w + y) * w + x] = sr[(size_t)it * p->numk + ik];
  }
  return CTEM_OK;
}
