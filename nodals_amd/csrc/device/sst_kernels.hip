// Batched pointwise SST and wall-law kernels (one thread per point / face).
#include "sst_kernels.hpp"

#include <mutex>

#include "common.hpp"
#include "sst_physics.hpp"

namespace nodals {

namespace {

__global__ void k_sst_local(long long n, const double* __restrict__ dist, const double* __restrict__ xy,
                            double radius, double nu, const double* __restrict__ gu, const double* __restrict__ k,
                            const double* __restrict__ om, const double* __restrict__ gk,
                            const double* __restrict__ gw, double kFloor, double omFloor, int deviatoric,
                            double* __restrict__ out, int* __restrict__ flags) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d = xy ? sst::pipeDistance(xy[2 * i], xy[2 * i + 1], radius) : sst::floorDistance(dist[i], radius);
  double g[9], a[3], b[3];
  for (int j = 0; j < 9; ++j) g[j] = gu[9 * i + j];
  for (int j = 0; j < 3; ++j) {
    a[j] = gk[3 * i + j];
    b[j] = gw[3 * i + j];
  }
  const sst::Local q = sst::local(d, nu, g, k[i], om[i], a, b, kFloor, omFloor, deviatoric != 0);
  const double v[14] = {q.nut, q.F1, q.F2, q.sigmaK, q.sigmaW, q.beta, q.gamma, q.strain,
                        q.pk, q.srcOmega, q.reactK, q.reactOmega, q.omegaProd, q.omegaCross};
  for (int j = 0; j < 14; ++j) out[14 * i + j] = v[j];
  flags[2 * i] = q.kFloored;
  flags[2 * i + 1] = q.omegaFloored;
}

__global__ void k_spalding(long long n, const double* __restrict__ slip, const double* __restrict__ y, double nu,
                           double* __restrict__ out, int* __restrict__ ok) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double ut = 0, yp = 0, beta = 0;
  ok[i] = sst::spaldingRoot(slip[i], y[i], nu, ut, yp, beta) ? 1 : 0;
  out[3 * i] = ut;
  out[3 * i + 1] = yp;
  out[3 * i + 2] = beta;
}

__global__ void k_profile5(long long n, const double* __restrict__ U, const double* __restrict__ Us,
                           const double* __restrict__ y, double nu, double* __restrict__ out, int* __restrict__ valid) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double u[5], us[5], yy[5];
  for (int m = 0; m < 5; ++m) {
    u[m] = U[5 * i + m];
    us[m] = Us[5 * i + m];
    yy[m] = y[5 * i + m];
  }
  const sst::Profile p = sst::fitProfile5(u, us, yy, nu);
  const double v[6] = {p.uTau, p.yPlus, p.beta, p.uMid, p.fitRms, p.span};
  for (int j = 0; j < 6; ++j) out[6 * i + j] = v[j];
  valid[i] = p.valid;
}

__global__ void k_fit_tangential(long long n, const double* __restrict__ samples, const double* __restrict__ y,
                                 const double* __restrict__ nrm, const double* __restrict__ pic, double nu, int count,
                                 int zeroLimit, double* __restrict__ out, int* __restrict__ valid) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double sx[5], sy[5], sz[5];
  for (int m = 0; m < 5; ++m) {
    sx[m] = samples[15 * i + 3 * m];
    sy[m] = samples[15 * i + 3 * m + 1];
    sz[m] = samples[15 * i + 3 * m + 2];
  }
  const sst::VecFit f = sst::fitTangential(sx, sy, sz, y + 5 * i, nrm + 3 * i, pic + 3 * i, nu, count, zeroLimit != 0);
  out[3 * i] = f.beta;
  out[3 * i + 1] = f.uTau;
  out[3 * i + 2] = f.yPlus;
  valid[i] = f.valid;
}

__global__ void k_omega_wall(long long n, const double* __restrict__ in, const int* __restrict__ shared,
                             double* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* a = in + 15 * i;
  const sst::OmegaWall q = sst::omegaWall(a[0], a[1], a[2], a[3], a[4], a[5], a[13], a[6], a[7], a[8], a[9], a[10],
                                          a[11], a[12], shared[i] != 0);
  const double v[8] = {q.omegaTarget, q.omegaVis, q.omegaLog, q.lamFrac, q.rey, q.yPlusK, q.pkAuto, q.nutSpalding};
  for (int j = 0; j < 8; ++j) out[8 * i + j] = v[j];
}

// Host arrays in, host arrays out, through device copies on a private stream.
struct Staging {
  hipStream_t s = nullptr;
  Staging() {
    NODALS_HIP(hipSetDevice(0));
    NODALS_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  }
  ~Staging() {
    if (s) (void)hipStreamDestroy(s);
  }
  template <class T>
  DBuf<T> in(const T* h, std::size_t n) {
    DBuf<T> d;
    d.alloc(std::max<std::size_t>(n, 1));
    if (n) NODALS_HIP(hipMemcpyAsync(d.data(), h, n * sizeof(T), hipMemcpyHostToDevice, s));
    return d;
  }
  template <class T>
  void out(T* h, const DBuf<T>& d, std::size_t n) {
    if (n) NODALS_HIP(hipMemcpyAsync(h, d.data(), n * sizeof(T), hipMemcpyDeviceToHost, s));
  }
  void sync() { NODALS_HIP(hipStreamSynchronize(s)); }
};

// Staging memory of the allocation-free five-point fit (profile5BatchNoAlloc):
// U, U_signed and y ([n][5] doubles each), out ([n][6] doubles) and valid
// ([n] ints), packed in that order for n <= kNoAllocMaxFaces, in device
// memory the code object reserves on every GPU it is loaded on.  Host code
// never passes `scratch` itself to a kernel or copy: each device's address of
// it is taken once with hipGetSymbolAddress and used from then on.
constexpr long long kNoAllocMaxFaces = 65536;
constexpr std::size_t kNoAllocBytes = (std::size_t)kNoAllocMaxFaces * (21 * sizeof(double) + sizeof(int));
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
unsigned char* g_scratchPtr[kMaxDevices] = {};

// The calling thread's current device's address of `scratch`, looked up on
// the first call on that device.
unsigned char* scratchOnCurrentDevice() {
  int dev = -1;
  NODALS_HIP(hipGetDevice(&dev));
  require(dev >= 0 && dev < kMaxDevices, "wall_profile5_noalloc: device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratchMu);
  if (!g_scratchPtr[dev]) {
    void* ptr = nullptr;
    NODALS_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    require(ptr != nullptr, "wall_profile5_noalloc: no device address for the staging array");
    g_scratchPtr[dev] = static_cast<unsigned char*>(ptr);
  }
  return g_scratchPtr[dev];
}

}  // namespace

void sstLocalBatch(long long n, const double* dist, const double* xy, double radius, double nu, const double* gu,
                   const double* k, const double* omega, const double* gk, const double* gw, double kFloor,
                   double omegaFloor, int deviatoric, double* out, int* flags) {
  require(n >= 0 && (dist || xy) && gu && k && omega && gk && gw && out && flags, "sst_local: null or bad arguments");
  if (n == 0) return;
  Staging S;
  DBuf<double> dd = dist ? S.in(dist, (std::size_t)n) : DBuf<double>();
  DBuf<double> dxy = xy ? S.in(xy, 2 * (std::size_t)n) : DBuf<double>();
  DBuf<double> dg = S.in(gu, 9 * (std::size_t)n), dk = S.in(k, (std::size_t)n), dw = S.in(omega, (std::size_t)n);
  DBuf<double> dgk = S.in(gk, 3 * (std::size_t)n), dgw = S.in(gw, 3 * (std::size_t)n);
  DBuf<double> dout;
  dout.alloc(14 * (std::size_t)n);
  DBuf<int> dfl;
  dfl.alloc(2 * (std::size_t)n);
  hipLaunchKernelGGL(k_sst_local, blocksFor(n), 256, 0, S.s, n, dist ? dd.data() : nullptr,
                     xy ? dxy.data() : nullptr, radius, nu, dg.data(), dk.data(), dw.data(), dgk.data(), dgw.data(),
                     kFloor, omegaFloor, deviatoric, dout.data(), dfl.data());
  NODALS_LAUNCHED();
  S.out(out, dout, 14 * (std::size_t)n);
  S.out(flags, dfl, 2 * (std::size_t)n);
  S.sync();
}

void spaldingBatch(long long n, const double* slip, const double* y, double nu, double* out, int* ok) {
  require(n >= 0 && slip && y && out && ok, "spalding: null or bad arguments");
  if (n == 0) return;
  Staging S;
  DBuf<double> ds = S.in(slip, (std::size_t)n), dy = S.in(y, (std::size_t)n), dout;
  dout.alloc(3 * (std::size_t)n);
  DBuf<int> dok;
  dok.alloc((std::size_t)n);
  hipLaunchKernelGGL(k_spalding, blocksFor(n), 256, 0, S.s, n, ds.data(), dy.data(), nu, dout.data(), dok.data());
  NODALS_LAUNCHED();
  S.out(out, dout, 3 * (std::size_t)n);
  S.out(ok, dok, (std::size_t)n);
  S.sync();
}

void profile5Batch(long long n, const double* U, const double* Us, const double* y, double nu, double* out,
                   int* valid) {
  require(n >= 0 && U && Us && y && out && valid, "profile5: null or bad arguments");
  if (n == 0) return;
  Staging S;
  DBuf<double> du = S.in(U, 5 * (std::size_t)n), dus = S.in(Us, 5 * (std::size_t)n), dy = S.in(y, 5 * (std::size_t)n);
  DBuf<double> dout;
  dout.alloc(6 * (std::size_t)n);
  DBuf<int> dv;
  dv.alloc((std::size_t)n);
  hipLaunchKernelGGL(k_profile5, blocksFor(n), 256, 0, S.s, n, du.data(), dus.data(), dy.data(), nu, dout.data(),
                     dv.data());
  NODALS_LAUNCHED();
  S.out(out, dout, 6 * (std::size_t)n);
  S.out(valid, dv, (std::size_t)n);
  S.sync();
}

void profile5BatchNoAlloc(long long n, const double* U, const double* Us, const double* y, double nu, double* out,
                          int* valid) {
  require(n >= 0 && U && Us && y && out && valid, "wall_profile5_noalloc: null or bad arguments");
  require(n <= kNoAllocMaxFaces, "wall_profile5_noalloc: more than 65536 faces (the staging array's capacity)");
  if (n == 0) return;
  unsigned char* base = scratchOnCurrentDevice();
  const std::size_t n5 = 5 * (std::size_t)n;
  double* dU = reinterpret_cast<double*>(base);
  double* dUs = dU + n5;
  double* dy = dUs + n5;
  double* dout = dy + n5;
  int* dvalid = reinterpret_cast<int*>(dout + 6 * (std::size_t)n);
  // Every output element is written by the kernel, so nothing left from an
  // earlier call needs clearing.  Null stream: a call is complete on return.
  NODALS_HIP(hipMemcpyAsync(dU, U, n5 * sizeof(double), hipMemcpyHostToDevice, nullptr));
  NODALS_HIP(hipMemcpyAsync(dUs, Us, n5 * sizeof(double), hipMemcpyHostToDevice, nullptr));
  NODALS_HIP(hipMemcpyAsync(dy, y, n5 * sizeof(double), hipMemcpyHostToDevice, nullptr));
  hipLaunchKernelGGL(k_profile5, blocksFor(n), 256, 0, nullptr, n, dU, dUs, dy, nu, dout, dvalid);
  NODALS_LAUNCHED();
  NODALS_HIP(hipMemcpyAsync(out, dout, 6 * (std::size_t)n * sizeof(double), hipMemcpyDeviceToHost, nullptr));
  NODALS_HIP(hipMemcpyAsync(valid, dvalid, (std::size_t)n * sizeof(int), hipMemcpyDeviceToHost, nullptr));
  NODALS_HIP(hipStreamSynchronize(nullptr));
}

void fitTangentialBatch(long long n, const double* samples, const double* y, const double* nrm, const double* pic,
                        double nu, int count, int zeroLimit, double* out, int* valid) {
  require(n >= 0 && samples && y && nrm && pic && out && valid, "fit_tangential: null or bad arguments");
  require(count >= 1 && count <= 5, "fit_tangential: count must be in 1..5");
  if (n == 0) return;
  Staging S;
  DBuf<double> ds = S.in(samples, 15 * (std::size_t)n), dy = S.in(y, 5 * (std::size_t)n);
  DBuf<double> dn = S.in(nrm, 3 * (std::size_t)n), dp = S.in(pic, 3 * (std::size_t)n), dout;
  dout.alloc(3 * (std::size_t)n);
  DBuf<int> dv;
  dv.alloc((std::size_t)n);
  hipLaunchKernelGGL(k_fit_tangential, blocksFor(n), 256, 0, S.s, n, ds.data(), dy.data(), dn.data(), dp.data(), nu,
                     count, zeroLimit, dout.data(), dv.data());
  NODALS_LAUNCHED();
  S.out(out, dout, 3 * (std::size_t)n);
  S.out(valid, dv, (std::size_t)n);
  S.sync();
}

void omegaWallBatch(long long n, const double* in, const int* shared, double* out) {
  require(n >= 0 && in && shared && out, "omega_wall: null or bad arguments");
  if (n == 0) return;
  Staging S;
  DBuf<double> di = S.in(in, 15 * (std::size_t)n), dout;
  DBuf<int> dsh = S.in(shared, (std::size_t)n);
  dout.alloc(8 * (std::size_t)n);
  hipLaunchKernelGGL(k_omega_wall, blocksFor(n), 256, 0, S.s, n, di.data(), dsh.data(), dout.data());
  NODALS_LAUNCHED();
  S.out(out, dout, 8 * (std::size_t)n);
  S.sync();
}

}  // namespace nodals
