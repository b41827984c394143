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
#include "wall/portable/spalding_active.hpp"
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

// Spalding fit over the active points of a five-slot sample set
// (active::fit_values, modules/wall/portable/spalding_active.hpp:59-112,
// the fit of spalding_match_profile5_tet_vector): U (speeds) and y are
// [n][5]; out[n][2] = uTau yPlus; valid[n].
void ref_fit_values(long long n, const double* U, const double* y, int count, double nu, double* out, int* valid) {
  for (long long i = 0; i < n; ++i) {
    double ut = 0, yp = 0, beta = 0, uMid = 0, rms = 0, span = 0;
    valid[i] = nodals::wall::active::fit_values(U + 5 * i, U + 5 * i, y + 5 * i, count, nu, ut, yp, beta, uMid, rms,
                                                span)
                   ? 1
                   : 0;
    out[2 * i] = ut;
    out[2 * i + 1] = yp;
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

// Tangential wall fit of an active-point set (vector3::fit_active,
// modules/wall/portable/spalding_active.hpp:129-154), stationary wall:
// samples[n][5][3], y[n][5], nrm[n][3], pic[n][3]; out[n][3] = beta uTau
// yPlus; valid[n].  The reference's host build returns the fit's own Robin
// coefficient only when its device constant STAGE27_SPALDING_ZERO_LIMIT is
// set; with zeroLimit the same selection (active::select_picard_beta with
// its viscous-limit argument set) is applied here to the same fit_values
// results, through the reference's own functions.
void ref_fit_tangential(long long n, const double* samples, const double* y, const double* nrm, const double* pic,
                        double nu, int count, int zeroLimit, double* out, int* valid) {
  using namespace nodals::wall;
  using vector3::V3;
  for (long long i = 0; i < n; ++i) {
    V3 s[5];
    double yy[5];
    for (int m = 0; m < 5; ++m) {
      s[m] = V3{samples[15 * i + 3 * m], samples[15 * i + 3 * m + 1], samples[15 * i + 3 * m + 2]};
      yy[m] = y[5 * i + m];
    }
    const V3 nv{nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]}, pv{pic[3 * i], pic[3 * i + 1], pic[3 * i + 2]};
    vector3::Result r = vector3::fit_active(s, yy, nv, pv, V3{0., 0., 0.}, nu, count);
    if (zeroLimit && count < 5) {
      double U[5] = {}, Y[5] = {};
      const unsigned mask = active::active_slot_mask(count);
      for (int m = 0; m < 5; ++m) {
        Y[m] = yy[m];
        if (mask & (1u << m)) U[m] = vector3::norm(vector3::tangent_projection(s[m], nv));
      }
      double ut = 0, yp = 0, betaFit = 0, uMid = 0, rms = 0, span = 0;
      if (active::fit_values(U, U, Y, count, nu, ut, yp, betaFit, uMid, rms, span)) {
        const double speed = vector3::norm(vector3::tangent_projection(pv, nv));
        const double floor = 1.e-6 * std::fmax(std::fabs(uMid), 1.);
        r.beta = active::select_picard_beta(ut, speed, floor, betaFit, count, true);
        r.valid = std::isfinite(r.beta) && r.beta > 0. && std::isfinite(rms) && std::isfinite(yp);
        r.uTau = ut;
        r.yPlus = yp;
      }
    }
    out[3 * i] = r.beta;
    out[3 * i + 1] = r.uTau;
    out[3 * i + 2] = r.yPlus;
    valid[i] = r.valid ? 1 : 0;
  }
}

}  // extern "C"
