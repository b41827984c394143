// Test-only host build of the product's pointwise SST / wall-law functions
// (nodals_amd/csrc/device/sst_physics.hpp) behind the same flat C interface
// as ref_sst_wrap.cpp, so the tests can show that the restatement the GPU
// kernels run reproduces the reference cores bit for bit when both are
// compiled for the host (the GPU build differs only by its math library).
#include "../nodals_amd/csrc/device/sst_physics.hpp"

extern "C" {

void port_sst_local(long long n, const double* distance, double radius, double nu, const double* gu, const double* k,
                    const double* omega, const double* gk, const double* gw, double kFloor, double omegaFloor,
                    int deviatoric, double* out, int* flags) {
  using namespace nodals::sst;
  for (long long i = 0; i < n; ++i) {
    const Local q = local(floorDistance(distance[i], radius), nu, gu + 9 * i, k[i], omega[i], gk + 3 * i, gw + 3 * i,
                          kFloor, omegaFloor, deviatoric != 0);
    const double v[14] = {q.nut, q.F1, q.F2, q.sigmaK, q.sigmaW, q.beta, q.gamma, q.strain,
                          q.pk, q.srcOmega, q.reactK, q.reactOmega, q.omegaProd, q.omegaCross};
    for (int j = 0; j < 14; ++j) out[14 * i + j] = v[j];
    flags[2 * i] = q.kFloored;
    flags[2 * i + 1] = q.omegaFloored;
  }
}

void port_spalding(long long n, const double* slip, const double* y, double nu, double* out, int* ok) {
  for (long long i = 0; i < n; ++i) {
    double ut = 0, yp = 0, beta = 0;
    ok[i] = nodals::sst::spaldingRoot(slip[i], y[i], nu, ut, yp, beta) ? 1 : 0;
    out[3 * i] = ut;
    out[3 * i + 1] = yp;
    out[3 * i + 2] = beta;
  }
}

void port_profile5(long long n, const double* U, const double* Us, const double* y, double nu, double* out,
                   int* valid) {
  for (long long i = 0; i < n; ++i) {
    const nodals::sst::Profile p = nodals::sst::fitProfile5(U + 5 * i, Us + 5 * i, y + 5 * i, nu);
    const double v[6] = {p.uTau, p.yPlus, p.beta, p.uMid, p.fitRms, p.span};
    for (int j = 0; j < 6; ++j) out[6 * i + j] = v[j];
    valid[i] = p.valid;
  }
}

void port_of_auto(long long n, const double* in, const int* shared, double* out) {
  for (long long i = 0; i < n; ++i) {
    const double* a = in + 15 * i;
    const nodals::sst::OmegaWall q = nodals::sst::omegaWall(a[0], a[1], a[2], a[3], a[4], a[5], a[13], a[6], a[7],
                                                            a[8], a[9], a[10], a[11], a[12], shared[i] != 0);
    const double v[8] = {q.omegaTarget, q.omegaVis, q.omegaLog, q.lamFrac, q.rey, q.yPlusK, q.pkAuto, q.nutSpalding};
    for (int j = 0; j < 8; ++j) out[8 * i + j] = v[j];
  }
}

void port_fit_tangential(long long n, const double* samples, const double* y, const double* nrm, const double* pic,
                         double nu, int count, int zeroLimit, double* out, int* valid) {
  for (long long i = 0; i < n; ++i) {
    double sx[5], sy[5], sz[5];
    for (int m = 0; m < 5; ++m) {
      sx[m] = samples[15 * i + 3 * m];
      sy[m] = samples[15 * i + 3 * m + 1];
      sz[m] = samples[15 * i + 3 * m + 2];
    }
    const nodals::sst::VecFit f =
        nodals::sst::fitTangential(sx, sy, sz, y + 5 * i, nrm + 3 * i, pic + 3 * i, nu, count, zeroLimit != 0);
    out[3 * i] = f.beta;
    out[3 * i + 1] = f.uTau;
    out[3 * i + 2] = f.yPlus;
    valid[i] = f.valid;
  }
}

}  // extern "C"
