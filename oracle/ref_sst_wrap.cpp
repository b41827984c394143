// Test-only C wrapper around the reference's portable SST and wall-law cores.
//
// Compiled by oracle/Makefile, together with the reference's own
// modules/physics/sst2003m_host.cpp and modules/wall/portable/wall_host.cpp
// where they lie, into oracle/_ref/libnodals_ref_sst.so.  Nothing of the
// reference is copied here: the wrapper loops the reference functions over
// arrays so the tests can compare the HIP kernels and the numpy restatement
// against the reference itself.
#include "physics/sst2003m_distance_core.hpp"
#include "physics/sst2003m_host.hpp"
#include "wall/portable/spalding_core.hpp"
#include "wall/portable/wall_host.hpp"

#include <cstdint>

extern "C" {

// out[n][14]: nut F1 F2 sigmaK sigmaW beta gamma strain pk srcOmega reactK
// reactOmega omegaProductionAudit omegaCrossAudit; flags[n][2]: k/omega
// floored.  pipe != 0: distance from (x, y) and radius (evaluate_host);
// otherwise distance[] (evaluate_at_distance).
void ref_sst_local(long long n, int pipe, const double* x, const double* y, const double* distance, double radius,
                   double nu, const double* gu, const double* k, const double* omega, const double* gk,
                   const double* gw, double kFloor, double omegaFloor, int deviatoric, double* out, int* flags) {
  using namespace nodals::sst2003m;
  for (long long i = 0; i < n; ++i) {
    LocalState<double> q;
    if (pipe) {
      HostInput in;
      in.x = x[i];
      in.y = y[i];
      in.radius = radius;
      in.nu = nu;
      for (int j = 0; j < 9; ++j) in.gu[j] = gu[9 * i + j];
      in.k = k[i];
      in.omega = omega[i];
      for (int j = 0; j < 3; ++j) {
        in.gradK[j] = gk[3 * i + j];
        in.gradOmega[j] = gw[3 * i + j];
      }
      in.kFloor = kFloor;
      in.omegaFloor = omegaFloor;
      in.deviatoricProduction = deviatoric;
      q = evaluate_host(in);
    } else {
      q = evaluate_at_distance<double>(distance[i], radius, nu, gu + 9 * i, k[i], omega[i], gk + 3 * i, gw + 3 * i,
                                       kFloor, omegaFloor, deviatoric);
    }
    const double v[14] = {q.nut, q.F1, q.F2, q.sigmaK, q.sigmaW, q.beta, q.gamma, q.strain, q.pk, q.srcOmega,
                          q.reactK, q.reactOmega, q.omegaProductionAudit, q.omegaCrossAudit};
    for (int j = 0; j < 14; ++j) out[14 * i + j] = v[j];
    flags[2 * i] = q.kFloored ? 1 : 0;
    flags[2 * i + 1] = q.omegaFloored ? 1 : 0;
  }
}

// Spalding root per sample: out[n][3] = ut, y+, beta; ok[n].
void ref_spalding(long long n, const double* slip, const double* y, double nu, double* out, int* ok) {
  for (long long i = 0; i < n; ++i) {
    double ut = 0, yp = 0, beta = 0;
    ok[i] = nodals::wall::spalding(slip[i], y[i], nu, ut, yp, beta) ? 1 : 0;
    out[3 * i] = ut;
    out[3 * i + 1] = yp;
    out[3 * i + 2] = beta;
  }
}

// Five-point profile fit per face: U, Usigned, y are [n][5]; out[n][6] =
// uTau yPlus beta uMid fitRms pointUTauRelSpan; valid[n].
void ref_profile5(long long n, const double* U, const double* Us, const double* y, double nu, double* out,
                  int* valid) {
  for (long long i = 0; i < n; ++i) {
    const nodals::wall::HostProfile p = nodals::wall::fit_profile5_host(U + 5 * i, Us + 5 * i, y + 5 * i, nu);
    const double v[6] = {p.uTau, p.yPlus, p.beta, p.uMid, p.fitRms, p.pointUTauRelSpan};
    for (int j = 0; j < 6; ++j) out[6 * i + j] = v[j];
    valid[i] = p.valid ? 1 : 0;
  }
}

// OpenFOAM-style omega wall target.  in[n][15] = y nu Ut uTau k omega kFloor
// omegaFloor kappa E beta1 reBlend prodLimit pk nut; shared[n];
// out[n][8] = omegaTarget omegaVis omegaLog lamFrac rey yPlusK pkAuto
// nutSpalding.
void ref_of_auto(long long n, const double* in, const int* shared, double* out) {
  for (long long i = 0; i < n; ++i) {
    const double* a = in + 15 * i;
    nodals::wall::HostOFInput q{a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10], a[11], a[12],
                                a[13], a[14], shared[i]};
    const nodals::wall::HostOFState s = nodals::wall::evaluate_of_auto_host(q);
    const double v[8] = {s.omegaTarget, s.omegaVis, s.omegaLog, s.lamFrac, s.rey, s.yPlusK, s.pkAuto, s.nutSpalding};
    for (int j = 0; j < 8; ++j) out[8 * i + j] = v[j];
  }
}

}  // extern "C"
