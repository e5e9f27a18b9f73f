// Monte Carlo electron trajectories (include/ctem_mc.h): the trajectory loop of the reference's
// CTEMMC program (single_run, src_programs/CTEMMC.f90:304-526) on the GPU.
//
// k_mc: one lane per electron at a time. The loop is flat -- one trajectory step per iteration,
// and a lane whose electron has ended (stopped below 1 keV, left the surface, or 5000 steps)
// starts its next electron in the next iteration -- so the lanes of a wave stay busy however the
// trajectory lengths differ. Every step is the program's double-precision expression in its
// order (contraction off; division and square root correctly rounded), with the device's log,
// sin, cos, acos, atan. Hits go to the histograms with integer atomics, so the result does not
// depend on the order in which lanes finish.
//
// Random numbers: rng.f95's generator (lag-3 subtraction modulo m = 2147483579 and the 69069
// congruential generator, float32 output). CTEM_MC_BLOCKED: electron e starts at draw 16384 e of the
// stream seeded with primeseed + 1; the host computes each lane's first state by jumping the
// generator (the lag-3 part is linear modulo m, a 3 x 3 matrix power; the congruential part is an
// affine map), and a lane moves on to its next electron (e + G) by one more jump of G 16384 draws.
// CTEM_MC_SEQUENTIAL: one lane draws the stream continuously, electron after electron.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ctem_mc.h"

#pragma clang fp contract(off)

namespace {

thread_local std::string g_mcerr;

int32_t mc_fail(int32_t code, const std::string& m) {
  g_mcerr = m;
  return code;
}

constexpr int64_t kM = 2147483579LL;       // modulus of the lag-3 subtraction
constexpr uint64_t kBlock = 16384;         // draws per electron in the blocked mode (>= 1 + 3 * 5000)
constexpr int kNumSteps = 5000;

// jump of the generator: the lag-3 state (s1, s2, s3) -> P (s1, s2, s3) mod m, s4 -> A s4 + C
struct Jump {
  int64_t P[9];
  uint32_t A, C;
};

void m3mul(const int64_t a[9], const int64_t b[9], int64_t c[9]) {
  int64_t t[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      int64_t s = 0;
      for (int k = 0; k < 3; k++) s = (s + a[3 * i + k] * b[3 * k + j]) % kM;
      t[3 * i + j] = s;
    }
  for (int i = 0; i < 9; i++) c[i] = t[i];
}

// the jump by n draws: one draw maps (s1, s2, s3) to (s2, s3, s1 - s3 mod m) and s4 to 69069 s4 + 1013904243
Jump jump_of(uint64_t n) {
  Jump J{{1, 0, 0, 0, 1, 0, 0, 0, 1}, 1u, 0u};
  int64_t B[9] = {0, 1, 0, 0, 0, 1, 1, 0, kM - 1};
  uint32_t a = 69069u, c = 1013904243u;
  while (n) {
    if (n & 1) {
      m3mul(B, J.P, J.P);
      J.C = a * J.C + c;
      J.A = a * J.A;
    }
    m3mul(B, B, B);
    c = a * c + c;
    a = a * a;
    n >>= 1;
  }
  return J;
}

void apply_host(const Jump& J, int32_t s[4]) {
  int64_t w[3];
  for (int i = 0; i < 3; i++)
    w[i] = ((J.P[3 * i] * s[0]) % kM + (J.P[3 * i + 1] * s[1]) % kM + (J.P[3 * i + 2] * s[2]) % kM) % kM;
  for (int i = 0; i < 3; i++) s[i] = (int32_t)w[i];
  s[3] = (int32_t)(J.A * (uint32_t)s[3] + J.C);
}

}  // namespace

namespace ctem {

struct McDev {
  double Emin, pre, prealpha, J, predEds, tano, delta, cxstart, czstart, sPi2, Ehistmin, Ebinsize, depthstep,
      EkeV, Ze, tpi;
  int32_t numEbins, numzbins, nx, nzx, nzw, numsx;
};

struct JumpDev {
  int64_t P[9];
  uint32_t A, C;
};

// rng_uniform (rng.f95): u = fl32(0.5 + 0.23283064e-9 * imz)
__device__ __forceinline__ float mc_uniform(int32_t s[4]) {
  int32_t imz = s[0] - s[2];
  if (imz < 0) imz += 2147483579;
  s[0] = s[1];
  s[1] = s[2];
  s[2] = imz;
  s[3] = (int32_t)(69069u * (uint32_t)s[3] + 1013904243u);
  imz = (int32_t)((uint32_t)imz + (uint32_t)s[3]);
  return (float)(0.5 + 0.23283064e-9 * (double)imz);
}

__device__ __forceinline__ void mc_jump(const JumpDev& J, int32_t s[4]) {
  uint64_t w[3];
  for (int i = 0; i < 3; i++) {
    const uint64_t t = (uint64_t)J.P[3 * i] * (uint64_t)s[0] + (uint64_t)J.P[3 * i + 1] * (uint64_t)s[1] +
                       (uint64_t)J.P[3 * i + 2] * (uint64_t)s[2];
    w[i] = t % (uint64_t)kM;
  }
  for (int i = 0; i < 3; i++) s[i] = (int32_t)w[i];
  s[3] = (int32_t)(J.A * (uint32_t)s[3] + J.C);
}

// compensated sum of three doubles in element order (the Fortran runtime's SUM)
__device__ __forceinline__ double sum3_kahan_d(double a, double b, double c) {
  double s = 0.0, comp = 0.0;
  const double v[3] = {a, b, c};
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const double y = v[i] - comp;
    const double t = s + y;
    comp = (t - s) - y;
    s = t;
  }
  return s;
}

__global__ __launch_bounds__(256) void k_mc(McDev P, const int32_t* __restrict__ lane_state, JumpDev JG,
                                            int64_t first, int64_t count, int sequential,
                                            int32_t* __restrict__ ae, int32_t* __restrict__ az,
                                            unsigned long long* __restrict__ dropped,
                                            unsigned long long* __restrict__ steps) {
  const int64_t G = (int64_t)gridDim.x * blockDim.x;
  const int64_t lane = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sequential && lane > 0) return;
  const int64_t end = first + count;
  int64_t e = first + lane;
  int32_t cur[4], st[4];
  for (int i = 0; i < 4; i++) cur[i] = st[i] = lane_state[4 * (sequential ? 0 : lane) + i];
  const double scaled = 1.0e8, min_energy = 1.0, presig = 1.5273987e19;
  double Ec = 0.0, alpha = 0.0, step = 0.0;
  double xyz[3] = {0.0, 0.0, 0.0}, cxyz[3] = {0.0, 0.0, 0.0};
  int traj = 0;
  bool fresh = true;
  unsigned long long drop = 0, nstep = 0;
  while (true) {
    if (fresh) {
      if (e >= end) break;
      if (!sequential) {
        for (int i = 0; i < 4; i++) st[i] = cur[i];
        mc_jump(JG, cur);
      }
      // the electron's entry (CTEMMC.f90:416-431)
      Ec = P.EkeV;
      alpha = P.prealpha / Ec;
      step = Ec * (Ec + 1024.0) / P.Ze / (Ec + 511.0);
      const double sige = presig * step * step * alpha * (1.0 + alpha);
      const double lambda = P.pre * sige;
      const double rr = (double)mc_uniform(st);
      step = -lambda * log(rr);
      cxyz[0] = P.cxstart; cxyz[1] = 0.0; cxyz[2] = P.czstart;
      for (int i = 0; i < 3; i++) xyz[i] = 0.0 + step * scaled * cxyz[i];
      traj = 0;
      fresh = false;
      continue;
    }
    bool done = traj >= kNumSteps;
    if (!done) {
      nstep++;
      // one step (:434-521)
      const double dEds = P.predEds * log(Ec * P.J + 1.0) / Ec;
      const double dE = dEds * step;
      Ec = Ec + dE;
      if (Ec < min_energy) {
        done = true;
      } else {
        double rr = (double)mc_uniform(st);
        const double cphi = 1.0 - 2.0 * alpha * rr / (1.0 + alpha - rr);
        const double sphi = sin(acos(cphi));
        rr = (double)mc_uniform(st);
        const double psi = P.tpi * rr;
        const double spsi = sin(psi), cpsi = cos(psi);
        double cxyzp[3];
        if (fabs(cxyz[2]) > 0.99999) {
          cxyzp[0] = sphi * cpsi;
          cxyzp[1] = sphi * spsi;
          cxyzp[2] = (cxyz[2] / fabs(cxyz[2])) * cphi;
        } else {
          const double dsq = __dsqrt_rn(1.0 - cxyz[2] * cxyz[2]);
          const double dsqi = 1.0 / dsq;
          cxyzp[0] = sphi * (cxyz[0] * cxyz[2] * cpsi - cxyz[1] * spsi) * dsqi + cxyz[0] * cphi;
          cxyzp[1] = sphi * (cxyz[1] * cxyz[2] * cpsi + cxyz[0] * spsi) * dsqi + cxyz[1] * cphi;
          cxyzp[2] = -sphi * cpsi * dsq + cxyz[2] * cphi;
        }
        const double ss = sum3_kahan_d(cxyzp[0] * cxyzp[0], cxyzp[1] * cxyzp[1], cxyzp[2] * cxyzp[2]);
        const double dd = 1.0 / __dsqrt_rn(ss);
        for (int i = 0; i < 3; i++) cxyzp[i] = cxyzp[i] * dd;
        alpha = P.prealpha / Ec;
        step = Ec * (Ec + 1024.0) / P.Ze / (Ec + 511.0);
        const double sige = presig * step * step * alpha * (1.0 + alpha);
        const double lambda = P.pre * sige;
        rr = (double)mc_uniform(st);
        step = -lambda * log(rr);
        double xyzn[3];
        for (int i = 0; i < 3; i++) xyzn[i] = xyz[i] + step * scaled * cxyzp[i];
        const double znmax = xyzn[1] * P.tano;
        if (xyzn[2] > znmax) {
          // the exit direction on the square Lambert grid (Lambert2DSquareInverseDouble,
          // Lambert.f90:452-483, times delta) and the histograms (:489-514)
          double res0 = 0.0, res1 = 0.0;
          const double s2 = (cxyzp[0] * cxyzp[0] + cxyzp[1] * cxyzp[1]) + cxyzp[2] * cxyzp[2];
          if (!(fabs(1.0 - s2) > 1.0e-12) && fabs(cxyzp[2]) != 1.0) {
            if (fabs(cxyzp[1]) <= fabs(cxyzp[0])) {
              const double q = fabs(cxyzp[0]) / cxyzp[0] * __dsqrt_rn(2.0 * (1.0 - cxyzp[2]));
              res0 = q * P.sPi2;
              res1 = q * atan(cxyzp[1] / cxyzp[0]) / P.sPi2;
            } else {
              const double q = fabs(cxyzp[1]) / cxyzp[1] * __dsqrt_rn(2.0 * (1.0 - cxyzp[2]));
              res0 = q * atan(cxyzp[0] / cxyzp[1]) / P.sPi2;
              res1 = q * P.sPi2;
            }
          }
          const double dx0 = P.delta * res0, dx1 = P.delta * res1;
          const int ix = (int)round(dx1), iy = (int)round(-dx0);
          if (abs(ix) <= P.nx && abs(iy) <= P.nx && Ec > P.Emin) {
            const int iE = (int)round((Ec - P.Ehistmin) / P.Ebinsize) + 1;
            if (iE < 1 || iE > P.numEbins) {
              drop++;
            } else {
              const double edis = fabs(xyz[2] / cxyzp[2]);
              const int iz = (int)round(edis * 0.1 / P.depthstep) + 1;
              if (iz > 0 && iz <= P.numzbins) {
                const int px = (int)roundf(__fdiv_rn((float)ix, 10.0f));
                const int py = (int)roundf(__fdiv_rn((float)iy, 10.0f));
                if (abs(px) <= P.nzx && abs(py) <= P.nzx)
                  atomicAdd(&az[(((int64_t)(py + P.nzx) * P.nzw + (px + P.nzx)) * P.numzbins + (iz - 1)) * P.numEbins +
                                (iE - 1)],
                            1);
                else
                  drop++;
              }
              atomicAdd(&ae[((int64_t)(iy + P.nx) * P.numsx + (ix + P.nx)) * P.numEbins + (iE - 1)], 1);
            }
          }
          done = true;
        } else {
          for (int i = 0; i < 3; i++) { cxyz[i] = cxyzp[i]; xyz[i] = xyzn[i]; }
          traj++;
        }
      }
    }
    if (done) {
      fresh = true;
      e += sequential ? 1 : G;
    }
  }
  if (drop) atomicAdd(dropped, drop);
  atomicAdd(steps, nstep);
}

}  // namespace ctem

struct ctem_mc {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  ctem_mc_params p{};
  ctem::McDev d{};
  int32_t seed_state[4] = {0, 0, 0, 0};
  int32_t* ae = nullptr;
  int32_t* az = nullptr;
  int32_t* lanes = nullptr;
  unsigned long long* dropped = nullptr;
  unsigned long long* steps = nullptr;
  size_t ne = 0, nz = 0;
  int lanes_n = 0;
  double last_ms = 0.0;
};

#define MCHIP(expr)                                                                        \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return mc_fail(CTEM_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

extern "C" CTEM_API const char* ctem_mc_last_error(void) { return g_mcerr.c_str(); }

extern "C" CTEM_API int32_t ctem_mc_create(int32_t device, const ctem_mc_params* p, ctem_mc** out) {
#pragma clang fp contract(off)
  if (!p || !out) return mc_fail(CTEM_E_ARG, "ctem_mc_create: null argument");
  *out = nullptr;
  if (p->numsx < 1 || (p->numsx % 2) == 0 || p->numsx > 20001)
    return mc_fail(CTEM_E_ARG, "ctem_mc_create: numsx must be odd, 1..20001");
  if (!(p->Ebinsize > 0.0) || !(p->depthstep > 0.0) || !(p->EkeV > p->Ehistmin) || !(p->depthmax >= 0.0))
    return mc_fail(CTEM_E_ARG, "ctem_mc_create: need Ebinsize > 0, depthstep > 0, EkeV > Ehistmin");
  if (!(p->Ze > 0.0) || !(p->density > 0.0) || !(p->at_wt > 0.0) || p->num_el < 0)
    return mc_fail(CTEM_E_ARG, "ctem_mc_create: Ze, density, at_wt must be positive");
  if (p->primeseed < 0 || p->primeseed >= 2147483578)
    return mc_fail(CTEM_E_ARG, "ctem_mc_create: primeseed + 1 must lie in 0..2147483578 (the generator's modulus)");
  const double nEd = (p->EkeV - p->Ehistmin) / p->Ebinsize, nZd = p->depthmax / p->depthstep;
  if (!(nEd < 4096.0) || !(nZd < 4096.0)) return mc_fail(CTEM_E_RANGE, "ctem_mc_create: more than 4096 bins");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return mc_fail(CTEM_E_HIP, "no HIP device available");
  if (device < 0 || device >= ndev) return mc_fail(CTEM_E_ARG, "ctem_mc_create: device out of range");
  MCHIP(hipSetDevice(device));
  ctem_mc* m = new ctem_mc();
  m->device = device;
  m->p = *p;
  // the derived constants of single_run (:367-392), on the host with the C library as the
  // reference computes them
  ctem::McDev& d = m->d;
  const double cPi = 3.141592653589, cAvogadro = 6.0221415e23, cDtoR = 0.017453293;
  const double dpi = 4.0 * std::atan(1.0);
  const int nx = (p->numsx - 1) / 2;
  d.numEbins = (int)((p->EkeV - p->Ehistmin) / p->Ebinsize) + 1;
  d.numzbins = (int)(p->depthmax / p->depthstep) + 1;
  d.nx = nx;
  d.nzx = nx / 10;
  d.nzw = 2 * (nx / 10) + 1;
  d.numsx = p->numsx;
  d.Emin = p->Ehistmin - p->Ebinsize / 2.0;
  d.pre = p->at_wt / cAvogadro / p->density;
  d.prealpha = 3.4e-3 * std::pow(p->Ze, (double)0.67f);  // Ze**(0.67), a default-real exponent
  double J = (9.76 * p->Ze + 58.5 / std::pow(p->Ze, 0.19)) * 1.0e-3 / (double)1.166f;
  d.J = 1.0 / J;
  d.predEds = -78500.0 * p->density * p->Ze / p->at_wt;
  d.tano = std::tan(p->omega * cDtoR);
  d.delta = (double)nx / std::sqrt(dpi * 0.5);
  d.cxstart = std::cos((90.0 - p->sig) * cDtoR);
  d.czstart = -std::sin((90.0 - p->sig) * cDtoR);
  d.sPi2 = std::sqrt(dpi) * 0.5;
  d.Ehistmin = p->Ehistmin;
  d.Ebinsize = p->Ebinsize;
  d.depthstep = p->depthstep;
  d.EkeV = p->EkeV;
  d.Ze = p->Ze;
  d.tpi = 2.0 * cPi;
  m->seed_state[0] = p->primeseed + 1;
  m->seed_state[1] = 362436069;
  m->seed_state[2] = 16163801;
  m->seed_state[3] = 1131199299;
  m->ne = (size_t)d.numEbins * p->numsx * p->numsx;
  m->nz = (size_t)d.numEbins * d.numzbins * d.nzw * d.nzw;
  hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&m->ev[0]);
  if (e == hipSuccess) e = hipEventCreate(&m->ev[1]);
  if (e == hipSuccess) e = hipMalloc((void**)&m->ae, sizeof(int32_t) * m->ne);
  if (e == hipSuccess) e = hipMalloc((void**)&m->az, sizeof(int32_t) * m->nz);
  if (e == hipSuccess) e = hipMalloc((void**)&m->dropped, sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMalloc((void**)&m->steps, sizeof(unsigned long long));
  if (e != hipSuccess) {
    ctem_mc_destroy(m);
    return mc_fail(CTEM_E_HIP, std::string("ctem_mc_create: ") + hipGetErrorString(e));
  }
  int32_t rc = ctem_mc_reset(m);
  if (rc) { ctem_mc_destroy(m); return rc; }
  *out = m;
  return CTEM_OK;
}

extern "C" CTEM_API int32_t ctem_mc_dims(const ctem_mc* m, int32_t* numEbins, int32_t* numzbins, int32_t* nx,
                                         int32_t* nzw) {
  if (!m) return mc_fail(CTEM_E_ARG, "ctem_mc_dims: null object");
  if (numEbins) *numEbins = m->d.numEbins;
  if (numzbins) *numzbins = m->d.numzbins;
  if (nx) *nx = m->d.nx;
  if (nzw) *nzw = m->d.nzw;
  return CTEM_OK;
}

extern "C" CTEM_API int32_t ctem_mc_reset(ctem_mc* m) {
  if (!m) return mc_fail(CTEM_E_ARG, "ctem_mc_reset: null object");
  MCHIP(hipSetDevice(m->device));
  MCHIP(hipMemsetAsync(m->ae, 0, sizeof(int32_t) * m->ne, m->stream));
  MCHIP(hipMemsetAsync(m->az, 0, sizeof(int32_t) * m->nz, m->stream));
  MCHIP(hipMemsetAsync(m->dropped, 0, sizeof(unsigned long long), m->stream));
  MCHIP(hipMemsetAsync(m->steps, 0, sizeof(unsigned long long), m->stream));
  MCHIP(hipStreamSynchronize(m->stream));
  return CTEM_OK;
}

extern "C" CTEM_API int32_t ctem_mc_run(ctem_mc* m, int32_t mode, int64_t first, int64_t count) {
  if (!m) return mc_fail(CTEM_E_ARG, "ctem_mc_run: null object");
  if (mode != CTEM_MC_BLOCKED && mode != CTEM_MC_SEQUENTIAL) return mc_fail(CTEM_E_ARG, "ctem_mc_run: unknown mode");
  if (first < 0 || count < 0 || first > ((int64_t)1 << 46) - count)
    return mc_fail(CTEM_E_ARG, "ctem_mc_run: electron range must lie in 0 .. 2^46");
  if (mode == CTEM_MC_SEQUENTIAL && first != 0)
    return mc_fail(CTEM_E_ARG, "ctem_mc_run: the program's single stream starts at electron 0");
  MCHIP(hipSetDevice(m->device));
  hipDeviceProp_t prop;
  MCHIP(hipGetDeviceProperties(&prop, m->device));
  // lanes: four 256-lane workgroups per CU, no more than the electrons
  const int64_t want = mode == CTEM_MC_SEQUENTIAL ? 1 : std::min<int64_t>((int64_t)prop.multiProcessorCount * 4 * 256,
                                                                          std::max<int64_t>(count, 1));
  const int blocks = (int)((want + 255) / 256);
  const int G = blocks * 256;
  std::vector<int32_t> st(4 * (size_t)G);
  Jump JG{};
  {
    int32_t s[4] = {m->seed_state[0], m->seed_state[1], m->seed_state[2], m->seed_state[3]};
    if (mode == CTEM_MC_BLOCKED) {
      apply_host(jump_of((uint64_t)first * kBlock), s);
      const Jump J1 = jump_of(kBlock);
      for (int l = 0; l < G; l++) {
        for (int i = 0; i < 4; i++) st[4 * (size_t)l + i] = s[i];
        apply_host(J1, s);
      }
      JG = jump_of((uint64_t)G * kBlock);
    } else {
      for (int i = 0; i < 4; i++) st[i] = s[i];
    }
  }
  if (m->lanes_n < G) {
    if (m->lanes) (void)hipFree(m->lanes);
    m->lanes = nullptr;
    m->lanes_n = 0;
    MCHIP(hipMalloc((void**)&m->lanes, sizeof(int32_t) * 4 * (size_t)G));
    m->lanes_n = G;
  }
  MCHIP(hipMemcpyAsync(m->lanes, st.data(), sizeof(int32_t) * 4 * (size_t)(mode == CTEM_MC_SEQUENTIAL ? 1 : G),
                       hipMemcpyHostToDevice, m->stream));
  ctem::JumpDev jd;
  for (int i = 0; i < 9; i++) jd.P[i] = JG.P[i];
  jd.A = JG.A;
  jd.C = JG.C;
  MCHIP(hipEventRecord(m->ev[0], m->stream));
  if (count > 0)
    hipLaunchKernelGGL(ctem::k_mc, dim3(blocks), dim3(256), 0, m->stream, m->d, m->lanes, jd, first, count,
                       mode == CTEM_MC_SEQUENTIAL ? 1 : 0, m->ae, m->az, m->dropped, m->steps);
  MCHIP(hipGetLastError());
  MCHIP(hipEventRecord(m->ev[1], m->stream));
  MCHIP(hipStreamSynchronize(m->stream));
  float ms = 0.0f;
  MCHIP(hipEventElapsedTime(&ms, m->ev[0], m->ev[1]));
  m->last_ms = ms;
  return CTEM_OK;
}

extern "C" CTEM_API int32_t ctem_mc_get(ctem_mc* m, int32_t* accum_e, int32_t* accum_z, int64_t* dropped) {
  if (!m) return mc_fail(CTEM_E_ARG, "ctem_mc_get: null object");
  MCHIP(hipSetDevice(m->device));
  if (accum_e) MCHIP(hipMemcpyAsync(accum_e, m->ae, sizeof(int32_t) * m->ne, hipMemcpyDeviceToHost, m->stream));
  if (accum_z) MCHIP(hipMemcpyAsync(accum_z, m->az, sizeof(int32_t) * m->nz, hipMemcpyDeviceToHost, m->stream));
  unsigned long long dr = 0;
  MCHIP(hipMemcpyAsync(&dr, m->dropped, sizeof dr, hipMemcpyDeviceToHost, m->stream));
  MCHIP(hipStreamSynchronize(m->stream));
  if (dropped) *dropped = (int64_t)dr;
  return CTEM_OK;
}

extern "C" CTEM_API int32_t ctem_mc_device(ctem_mc* m, int32_t** accum_e, int32_t** accum_z) {
  if (!m) return mc_fail(CTEM_E_ARG, "ctem_mc_device: null object");
  if (accum_e) *accum_e = m->ae;
  if (accum_z) *accum_z = m->az;
  return CTEM_OK;
}

extern "C" CTEM_API int32_t ctem_mc_steps(ctem_mc* m, int64_t* steps) {
  if (!m || !steps) return mc_fail(CTEM_E_ARG, "ctem_mc_steps: null argument");
  MCHIP(hipSetDevice(m->device));
  unsigned long long v = 0;
  MCHIP(hipMemcpyAsync(&v, m->steps, sizeof v, hipMemcpyDeviceToHost, m->stream));
  MCHIP(hipStreamSynchronize(m->stream));
  *steps = (int64_t)v;
  return CTEM_OK;
}

extern "C" CTEM_API int32_t ctem_mc_last_ms(const ctem_mc* m, double* ms) {
  if (!m || !ms) return mc_fail(CTEM_E_ARG, "ctem_mc_last_ms: null argument");
  *ms = m->last_ms;
  return CTEM_OK;
}

extern "C" CTEM_API int32_t ctem_mc_destroy(ctem_mc* m) {
  if (!m) return CTEM_OK;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  if (m->ae) (void)hipFree(m->ae);
  if (m->az) (void)hipFree(m->az);
  if (m->lanes) (void)hipFree(m->lanes);
  if (m->dropped) (void)hipFree(m->dropped);
  if (m->steps) (void)hipFree(m->steps);
  for (auto& e : m->ev)
    if (e) (void)hipEventDestroy(e);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return CTEM_OK;
}
