// nodals_sst: the native SST solver executable that the SST case launcher
// runs (turbulence/sst_gpu_v2/nodals.py:432-509 argv, :308-431 environment;
// the reference's build/nodals_sst, puretet_profile5_eq5_hxt5b/hxt5b_main.cu),
// on the path the spiral case selects: tetrahedral HXT1 (or OpenFOAM
// polyMesh) mesh, k-omega SST with Spalding/Nitsche wall functions, M3 block
// momentum, corrected SIMPLE coupling, PCG-AMG pressure.
//
// Output lines the launcher and its tests read: NODALS_SST_RESIDUALS every
// --sst-g2-print-every iterations, NODALS_SST_G2_VTU_CHECKPOINT,
// STAGE31_RESTART_SAVE= / STAGE31_RESTART_LOAD=, NODALS_SST_G2_FINAL and
// NODALS_SOFTWARE_COMPLETION=COMPLETE iterations=N converged=0|1; exit 0
// after a completed run (converged or not), 90 on an error (the message
// follows NODALS_SOFTWARE_ERROR).
//
// Checkpoints (STAGE31_CHECKPOINT_EVERY / _DIR, restart from
// STAGE31_RESTART_FROM) are this build's own format, named
// NODALS_STAGE31_STATE_it<N>.ckpt as the launcher expects: the magic
// "NODALSCK", uint32 version 1, int32 iteration, int64 free velocity DOFs,
// cells and k-omega entities, the 64-character input fingerprint
// (STAGE31_INPUT_FINGERPRINT, or blanks), then u [nv][3], p [nc], k [ns] and
// omega [ns] as float64.  A restart checks the sizes and, when both carry
// one, the fingerprint.
//
// Everything goes through the C ABI of libnodals_hip.so.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nodals_hip.h"

namespace {

struct Fail : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void check(int rc) {
  if (rc != 0) throw Fail(nodals_last_error());
}

std::string envOr(const char* name, const std::string& dflt) {
  const char* v = std::getenv(name);
  return v ? std::string(v) : dflt;
}

// Flags of the launcher's argv that this solver accepts and does not need:
// the G2 path selectors it always runs, the integral-profile and audit
// controls of the reference's other wall closures, the coupled-FGMRES
// momentum controls (the M3 sweeps run instead) and reporting cadences it
// does not use.
const std::set<std::string> kAccepted = {
    "--vtu", "--sst-g2", "--sst-g2-start", "--simple-coupling", "--max-outer", "--sst-g2-physical-gate",
    "--sst-g2-metrics-every", "--sst-g2-plateau-tol", "--sst-g2-plateau-samples",
    "--sst-shared-spalding-wall-state", "--sst-profile-integral-utau", "--sst-profile-integral-nq",
    "--sst-profile-integral-closure", "--sst-profile-average-closure", "--sst-profile-average-wall-law",
    "--sst-profile-match-y", "--sst-fgs-check-every", "--momentum-work", "--mom-max", "--mom-rtol", "--mom-atol",
    "--mom-fgmres-restart", "--profile-audit", "--sst-k-solver", "--sst-omega-solver", "--sst-of-production",
    "--amg-fine-mode", "--gamma", "--sst-g2-max", "--sst-g2-print-every", "--sst-g2-cont-atol"};
// Flags passed on to the H8 context unchanged.
const std::set<std::string> kH8 = {"--re", "--bulk", "--alpha-u", "--alpha-p", "--simple-tol", "--p-rtol",
                                   "--p-atol", "--p-max", "--amg-hierarchy", "--cf-coarsening", "--cf-strength",
                                   "--cf-interp", "--cf-theta", "--amg-terminal", "--amg-smoother",
                                   "--amg-mcgs-order", "--amg-mcgs-fine-sweeps", "--amg-mcgs-coarse-sweeps",
                                   "--amg-mcgs-omega", "--amg-power-its", "--amg-jacobi-omega", "--amg-precision",
                                   "--cf-pmax", "--amg-lambda-safety", "--amg-lambda-low-fraction"};
// Flags passed on to the k-omega setup unchanged.
const std::set<std::string> kTurb = {"--sst-intensity", "--sst-length-ratio", "--sst-alpha-k", "--sst-alpha-omega",
                                     "--sst-k-fgs-sweeps", "--sst-omega-fgs-sweeps", "--sst-omega-wall-mode",
                                     "--sst-wall-penalty-gamma", "--sst-profile-fit-y1", "--sst-profile-fit-y2",
                                     "--sst-profile-fit-y3", "--sst-profile-fit-y4", "--sst-profile-fit-y5",
                                     "--sst-of-kappa", "--sst-of-E", "--sst-of-beta1", "--sst-of-re-blend",
                                     "--sst-gs-omega", "--sst-k-floor", "--sst-omega-floor"};

struct Args {
  std::map<std::string, std::string> kv;
  std::string get(const std::string& k, const std::string& d = "") const {
    auto it = kv.find(k);
    return it == kv.end() ? d : it->second;
  }
};

Args parse(int argc, char** argv) {
  Args a;
  for (int i = 1; i < argc; ++i) {
    const std::string f = argv[i];
    if (f.rfind("--", 0) != 0 || i + 1 >= argc) throw Fail("expected --flag value, got " + f);
    if (!kAccepted.count(f) && !kH8.count(f) && !kTurb.count(f) && f != "--mesh" && f != "--pressure-solver" &&
        f != "--amg-jacobi-fine-sweeps" && f != "--amg-jacobi-coarse-sweeps")
      throw Fail("unknown option " + f);
    a.kv[f] = argv[++i];
  }
  return a;
}

void require(bool ok, const std::string& what) {
  if (!ok) throw Fail(what);
}

struct Context {
  nodals_h8* h = nullptr;
  int64_t nv = 0, nc = 0, ns = 0, nf = 0;
  ~Context() {
    if (h) nodals_h8_destroy(h);
  }
};

std::vector<double> getD(nodals_h8* h, const char* name, int64_t n) {
  std::vector<double> v((std::size_t)n);
  check(nodals_h8_get(h, name, v.data(), n));
  return v;
}

std::string ckptPath(const std::string& dir, int it) {
  return dir + "/NODALS_STAGE31_STATE_it" + std::to_string(it) + ".ckpt";
}

void saveCheckpoint(Context& c, const std::string& path, int it, const std::string& fingerprint) {
  std::vector<double> u((std::size_t)c.nv * 3);
  for (int d = 0; d < 3; ++d) {
    const std::vector<double> ud = getD(c.h, ("u" + std::to_string(d)).c_str(), c.nv);
    for (int64_t i = 0; i < c.nv; ++i) u[3 * (std::size_t)i + d] = ud[(std::size_t)i];
  }
  const std::vector<double> p = getD(c.h, "p", c.nc), k = getD(c.h, "k", c.ns), w = getD(c.h, "omega", c.ns);
  std::FILE* f = std::fopen(path.c_str(), "wb");
  require(f != nullptr, "cannot write checkpoint " + path);
  const uint32_t version = 1;
  const int32_t it32 = it;
  char fp[64];
  std::memset(fp, ' ', 64);
  std::memcpy(fp, fingerprint.data(), std::min<std::size_t>(64, fingerprint.size()));
  bool ok = std::fwrite("NODALSCK", 1, 8, f) == 8 && std::fwrite(&version, 4, 1, f) == 1 &&
            std::fwrite(&it32, 4, 1, f) == 1 && std::fwrite(&c.nv, 8, 1, f) == 1 && std::fwrite(&c.nc, 8, 1, f) == 1 &&
            std::fwrite(&c.ns, 8, 1, f) == 1 && std::fwrite(fp, 1, 64, f) == 64;
  ok = ok && std::fwrite(u.data(), 8, u.size(), f) == u.size() && std::fwrite(p.data(), 8, p.size(), f) == p.size() &&
       std::fwrite(k.data(), 8, k.size(), f) == k.size() && std::fwrite(w.data(), 8, w.size(), f) == w.size();
  ok = (std::fclose(f) == 0) && ok;
  require(ok, "error writing checkpoint " + path);
  std::printf("STAGE31_RESTART_SAVE=PASS iteration=%d path=%s fields=U_P_K_OMEGA\n", it, path.c_str());
  std::fflush(stdout);
}

// The wall path's face fits (friction velocity and y+ per wall face, the fit
// of the last assembly, one iteration behind the saved state) for the next
// segment's per-face sample heights: NODALS_FACE_UTAU_it<N>.csv, written to a
// temporary name and renamed.
void exportFaceUtau(Context& c, const std::string& dir, int it) {
  const std::vector<double> fit = getD(c.h, "wall_fit", 3 * c.nf), faces = getD(c.h, "turb_wall_faces", 2 * c.nf);
  const std::string path = dir + "/NODALS_FACE_UTAU_it" + std::to_string(it) + ".csv", tmp = path + ".tmp";
  std::FILE* f = std::fopen(tmp.c_str(), "w");
  require(f != nullptr, "cannot write " + path);
  std::fprintf(f, "# NODALS_FACE_UTAU_V1 iteration=%d faces=%lld fit=F2_lagged\n", it, (long long)c.nf);
  int64_t bad = 0;
  for (int64_t i = 0; i < c.nf; ++i) {
    const double beta = fit[3 * (std::size_t)i], ut = fit[3 * (std::size_t)i + 1], yp = fit[3 * (std::size_t)i + 2];
    if (!(beta > 0 && std::isfinite(beta) && std::isfinite(ut) && ut >= 0 && std::isfinite(yp))) ++bad;
    std::fprintf(f, "%lld,%d,%d,%.17g,%.17g\n", (long long)i, (int)faces[2 * (std::size_t)i],
                 (int)faces[2 * (std::size_t)i + 1], ut, yp);
  }
  const bool closed = std::fclose(f) == 0;
  if (bad || !closed) {
    std::remove(tmp.c_str());
    throw Fail(bad ? "invalid wall fits on " + std::to_string(bad) + " faces; no friction-velocity export"
                   : "error writing " + path);
  }
  require(std::rename(tmp.c_str(), path.c_str()) == 0, "cannot rename " + tmp);
  std::printf("NODALS_FACE_UTAU_EXPORT=PASS iteration=%d faces=%lld path=%s status=PASS\n", it, (long long)c.nf,
              path.c_str());
  std::fflush(stdout);
}

int loadCheckpoint(Context& c, const std::string& path, const std::string& fingerprint) {
  std::FILE* f = std::fopen(path.c_str(), "rb");
  require(f != nullptr, "cannot open checkpoint " + path);
  char magic[8];
  uint32_t version = 0;
  int32_t it = 0;
  int64_t nv = 0, nc = 0, ns = 0;
  char fp[64];
  bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "NODALSCK", 8) == 0 &&
            std::fread(&version, 4, 1, f) == 1 && version == 1 && std::fread(&it, 4, 1, f) == 1 &&
            std::fread(&nv, 8, 1, f) == 1 && std::fread(&nc, 8, 1, f) == 1 && std::fread(&ns, 8, 1, f) == 1 &&
            std::fread(fp, 1, 64, f) == 64;
  if (!ok) {
    std::fclose(f);
    throw Fail("not a checkpoint of this solver: " + path);
  }
  if (nv != c.nv || nc != c.nc || ns != c.ns) {
    std::fclose(f);
    throw Fail("checkpoint sizes do not match the mesh: " + path);
  }
  const std::string stored(fp, 64);
  if (!fingerprint.empty() && stored.find_first_not_of(' ') != std::string::npos &&
      stored.substr(0, fingerprint.size()) != fingerprint) {
    std::fclose(f);
    throw Fail("checkpoint input fingerprint does not match STAGE31_INPUT_FINGERPRINT: " + path);
  }
  std::vector<double> u((std::size_t)nv * 3), p((std::size_t)nc), k((std::size_t)ns), w((std::size_t)ns);
  ok = std::fread(u.data(), 8, u.size(), f) == u.size() && std::fread(p.data(), 8, p.size(), f) == p.size() &&
       std::fread(k.data(), 8, k.size(), f) == k.size() && std::fread(w.data(), 8, w.size(), f) == w.size();
  std::fclose(f);
  require(ok, "short checkpoint " + path);
  for (int d = 0; d < 3; ++d) {
    std::vector<double> ud((std::size_t)nv);
    for (int64_t i = 0; i < nv; ++i) ud[(std::size_t)i] = u[3 * (std::size_t)i + d];
    check(nodals_h8_set(c.h, ("u" + std::to_string(d)).c_str(), ud.data(), nv));
  }
  check(nodals_h8_set(c.h, "p", p.data(), nc));
  check(nodals_h8_set(c.h, "k", k.data(), ns));
  check(nodals_h8_set(c.h, "omega", w.data(), ns));
  std::printf("STAGE31_RESTART_LOAD=PASS path=%s iteration=%d next=%d fields=U_P_K_OMEGA\n", path.c_str(), it, it + 1);
  std::fflush(stdout);
  return it;
}

int run(int argc, char** argv) {
  const Args A = parse(argc, argv);
  require(!A.get("--mesh").empty(), "--mesh required");
  require(A.get("--sst-g2", "1") == "1", "only the G2 SST path (--sst-g2 1) is supported");
  require(A.get("--simple-coupling", "corrected") == "corrected", "only --simple-coupling corrected is supported");
  require(A.get("--sst-k-solver", "fgs-fixed") == "fgs-fixed" && A.get("--sst-omega-solver", "fgs-fixed") == "fgs-fixed",
          "only the fgs-fixed k and omega solvers are supported");
  require(A.get("--sst-of-production", "0") == "0", "the OF-auto wall production term is not supported");
  require(A.get("--amg-fine-mode", "explicit") == "explicit", "only --amg-fine-mode explicit is supported");
  const std::string psolver = A.get("--pressure-solver", "pcg_amg");
  require(psolver == "pcg_amg" || psolver == "pcg", "only the pcg_amg pressure solver is supported");
  const int maxIt = std::atoi(A.get("--sst-g2-max", A.get("--max-outer", "10")).c_str());
  require(maxIt >= 1, "--sst-g2-max must be >= 1");

  // H8 context: wall-function duct
  std::vector<std::string> h8 = {"nodals_sst", "--mesh", A.get("--mesh"), "--duct", "1", "--wall-model", "spalding",
                                 "--run-mode", "converge", "--max-outer", std::to_string(maxIt),
                                 "--pressure-solver", "pcg", "--verbose", "1", "--tag", "sst",
                                 "--print-every", A.get("--sst-g2-print-every", "10")};
  if (!A.get("--sst-g2-cont-atol").empty()) {
    h8.push_back("--sst-cont-atol");
    h8.push_back(A.get("--sst-g2-cont-atol"));
  }
  const std::string diameter = envOr("SST26_SPIRAL_DIAMETER", "");
  if (!diameter.empty()) {
    h8.push_back("--duct-diameter");
    h8.push_back(diameter);
  }
  const std::string vin = envOr("SST23_INLET_VELOCITY", "");
  if (!vin.empty()) {
    h8.push_back("--inlet-velocity");
    h8.push_back(vin);
  }
  const std::string smoother = A.get("--amg-smoother", "jacobi");
  for (const auto& kv : A.kv)
    if (kH8.count(kv.first)) {
      h8.push_back(kv.first);
      h8.push_back(kv.second);
    }
  // the case's fine/coarse AMG sweeps apply to the chosen smoother
  const std::string fineSw = A.get("--amg-jacobi-fine-sweeps", envOr("P_AMG_FINE_SWEEPS", "1"));
  const std::string coarseSw = A.get("--amg-jacobi-coarse-sweeps", envOr("P_AMG_COARSE_SWEEPS", "1"));
  if (smoother == "mcgs") {
    if (!A.kv.count("--amg-mcgs-fine-sweeps")) h8.insert(h8.end(), {"--amg-mcgs-fine-sweeps", fineSw});
    if (!A.kv.count("--amg-mcgs-coarse-sweeps")) h8.insert(h8.end(), {"--amg-mcgs-coarse-sweeps", coarseSw});
  } else {
    h8.insert(h8.end(), {"--amg-jacobi-fine-sweeps", fineSw, "--amg-jacobi-coarse-sweeps", coarseSw});
  }
  std::vector<const char*> hv;
  for (const auto& s : h8) hv.push_back(s.c_str());
  Context c;
  check(nodals_h8_create((int)hv.size(), hv.data(), &c.h));

  // k-omega setup
  std::vector<std::string> tu;
  for (const auto& kv : A.kv)
    if (kTurb.count(kv.first)) {
      tu.push_back(kv.first);
      tu.push_back(kv.second);
    }
  // wall sample placement (the reference binary's default is the endpoint
  // walk; the launcher asks for ray tracing)
  const std::string sampling = envOr("SST_WALL_SAMPLING_MODE", "legacy_endpoint");
  require(sampling == "ray_trace" || sampling == "legacy_endpoint",
          "SST_WALL_SAMPLING_MODE=" + sampling + " is not supported (ray_trace or legacy_endpoint)");
  tu.insert(tu.end(), {"--sst-wall-sampling", sampling, "--sst-wall-active-points", envOr("SST24B_ACTIVE_POINTS", "1"),
                       "--sst-wall-ray-max-crossings", envOr("SST_WALL_RAY_MAX_CROSSINGS", "4096"),
                       "--sst-wall-ray-bary-tolerance", envOr("SST_WALL_RAY_BARY_TOLERANCE", "1e-9"),
                       "--sst-wall-ray-local-adapt", envOr("STAGE27_LOCAL_RAY_ADAPT", "0")});
  // per-face sample heights from the previous segment's wall fits
  const std::string faceCsv = envOr("NODALS_FACE_YPLUS_UTAU_CSV", "");
  if (!faceCsv.empty()) {
    require(sampling == "ray_trace", "NODALS_FACE_YPLUS_UTAU_CSV requires SST_WALL_SAMPLING_MODE=ray_trace");
    const std::string base = envOr("STAGE32_SPALDING_SAMPLE_SCALE", "");
    require(!base.empty(), "NODALS_FACE_YPLUS_UTAU_CSV requires STAGE32_SPALDING_SAMPLE_SCALE");
    tu.insert(tu.end(), {"--sst-face-utau-csv", faceCsv, "--sst-spalding-sample-scale", base,
                         "--sst-face-yplus-sp", envOr("NODALS_FACE_YPLUS_SP", "30"),
                         "--sst-face-yplus-omega", envOr("NODALS_FACE_YPLUS_OMEGA", "30"),
                         "--sst-face-yplus-min-scale", envOr("NODALS_FACE_YPLUS_MIN_SCALE", "0.02"),
                         "--sst-face-yplus-max-scale", envOr("NODALS_FACE_YPLUS_MAX_SCALE", "0.35")});
  }
  std::vector<const char*> tv;
  for (const auto& s : tu) tv.push_back(s.c_str());
  check(nodals_h8_turb_init(c.h, (int)tv.size(), tv.empty() ? nullptr : tv.data()));
  {
    double rs[7];
    check(nodals_h8_get(c.h, "turb_sample_stats", rs, 7));
    std::printf("SST_WALL_SAMPLING_RUNTIME method=%s rayContainment=%s activeMask=0x%x status=PASS\n",
                sampling.c_str(), sampling == "ray_trace" ? "VERIFIED" : "NOT_CHECKED", (unsigned)rs[5]);
    std::printf("SST_WALL_RAY_CACHE samplePoints=%lld crossedTet=%lld maxCrossings=%d shortened=%lld "
                "minRetainedFrac=%.9g status=PASS\n",
                (long long)rs[0], (long long)rs[1], (int)rs[2], (long long)rs[3], rs[4]);
    if (!faceCsv.empty()) {
      double fs[10];
      check(nodals_h8_get(c.h, "turb_face_yplus_stats", fs, 10));
      std::printf("NODALS_FACE_YPLUS_LOADED faces=%lld targetSp=%s targetOmega=%s minScale=%s maxScale=%s "
                  "fallback=%lld clampSp=%lld clampOmega=%lld spY3Min=%.12e spY3Mean=%.12e spY3Max=%.12e "
                  "omegaY3Min=%.12e omegaY3Mean=%.12e omegaY3Max=%.12e input=%s source=LAGGED_F2_FACE_UTAU "
                  "status=PASS\n",
                  (long long)fs[0], envOr("NODALS_FACE_YPLUS_SP", "30").c_str(),
                  envOr("NODALS_FACE_YPLUS_OMEGA", "30").c_str(), envOr("NODALS_FACE_YPLUS_MIN_SCALE", "0.02").c_str(),
                  envOr("NODALS_FACE_YPLUS_MAX_SCALE", "0.35").c_str(), (long long)fs[1], (long long)fs[2],
                  (long long)fs[3], fs[4], fs[5], fs[6], fs[7], fs[8], fs[9], faceCsv.c_str());
    }
  }

  // wall closure from the launcher's environment
  const std::string conv = envOr("NODALS_MOMENTUM_CONVECTION_FORM", "advective");
  require(conv == "advective" || conv == "skew_symmetric",
          "NODALS_MOMENTUM_CONVECTION_FORM=" + conv + " is not supported (advective or skew_symmetric)");
  const std::string m3 = envOr("NODALS_M3_MODE", "fgs");
  require(m3 == "fgs" || m3 == "sgs", "NODALS_M3_MODE=" + m3 + " is not supported (fgs or sgs)");
  const std::string start = envOr("STAGE27_INITIAL_VELOCITY", "plug");
  std::vector<std::string> wa = {"--gamma", A.get("--gamma", envOr("SST22D4_GAMMA_N", "1e-6")),
                                 "--wall-active-points", envOr("SST24B_ACTIVE_POINTS", "1"),
                                 "--wall-zero-limit", envOr("STAGE27_SPALDING_ZERO_LIMIT", "1"),
                                 "--convection-form", conv,
                                 "--stress-form", envOr("NODALS_MOMENTUM_STRESS_FORM", "vector_laplacian"),
                                 "--m3-mode", m3,
                                 "--m3-sweeps", envOr("NODALS_M3_SWEEPS", "5"),
                                 "--m3-omega", envOr("NODALS_M3_OMEGA", "0.7"),
                                 "--initial-velocity", start == "potential" ? "potential" : "plug"};
  std::vector<const char*> wv;
  for (const auto& s : wa) wv.push_back(s.c_str());
  check(nodals_h8_wall_init(c.h, (int)wv.size(), wv.data()));

  int64_t info[8];
  check(nodals_h8_info(c.h, info));
  c.nc = info[0];
  c.nv = info[1];
  double tst[7];
  check(nodals_h8_get(c.h, "turb_inlet_state", tst, 7));
  c.nf = (int64_t)tst[3];
  c.ns = (int64_t)tst[4];
  std::printf("NODALS_SST_NATIVE mesh=%s cells=%lld freeVelocityDofs=%lld wallFaces=%lld start=%s "
              "convection=%s m3=%s status=PASS\n",
              A.get("--mesh").c_str(), (long long)c.nc, (long long)c.nv, (long long)c.nf, start.c_str(), conv.c_str(),
              m3.c_str());
  std::fflush(stdout);

  const std::string fingerprint = envOr("STAGE31_INPUT_FINGERPRINT", "");
  int it0 = 0;
  const std::string restart = envOr("STAGE31_RESTART_FROM", "");
  if (!restart.empty()) it0 = loadCheckpoint(c, restart, fingerprint);
  const int ckEvery = std::atoi(envOr("STAGE31_CHECKPOINT_EVERY", "0").c_str());
  const std::string ckDir = envOr("STAGE31_CHECKPOINT_DIR", ".");
  const int vtuEvery = std::atoi(envOr("SST_G2_VTU_EVERY", "0").c_str());
  const std::string vtuDir = envOr("SST_G2_VTU_DIR", ".");
  const std::string exportFlag = envOr("NODALS_FACE_YPLUS_EXPORT", "0");
  require(exportFlag == "0" || exportFlag == "1", "NODALS_FACE_YPLUS_EXPORT must be 0 or 1");
  const bool exportFaces = exportFlag == "1";
  auto save = [&](int at) {
    saveCheckpoint(c, ckptPath(ckDir, at), at, fingerprint);
    if (exportFaces) exportFaceUtau(c, ckDir, at);
  };

  nodals_h8_run_result r{};
  int last = it0;
  bool converged = false;
  double loopMs = 0.0;
  for (int it = it0 + 1; it <= maxIt && !converged;) {
    int stop = maxIt;
    if (ckEvery > 0) stop = std::min(stop, (it / ckEvery + (it % ckEvery ? 1 : 0)) * ckEvery);
    if (vtuEvery > 0) stop = std::min(stop, (it / vtuEvery + (it % vtuEvery ? 1 : 0)) * vtuEvery);
    check(nodals_h8_run_window(c.h, it, stop - it + 1, &r));
    loopMs += r.loop_ms;
    last = r.outer;
    converged = r.converged != 0;
    if (ckEvery > 0 && last % ckEvery == 0 && !converged) save(last);
    if (vtuEvery > 0 && last % vtuEvery == 0) {
      const std::string p = vtuDir + "/flow_sst_it" + std::to_string(last) + ".vtu";
      check(nodals_h8_write_vtu(c.h, p.c_str()));
      std::printf("NODALS_SST_G2_VTU_CHECKPOINT it=%d path=%s\n", last, p.c_str());
      std::fflush(stdout);
    }
    it = last + 1;
  }
  // the final (or converged) state, unless the cadence has just written it
  if (ckEvery > 0 && last > it0 && (converged || last % ckEvery != 0)) save(last);
  const std::string vtu = A.get("--vtu");
  if (!vtu.empty()) {
    check(nodals_h8_write_vtu(c.h, vtu.c_str()));
    std::printf("NODALS_SST_G2_VTU_CHECKPOINT it=%d path=%s final=1\n", last, vtu.c_str());
  }
  // wall metrics: area-weighted mean wall shear and y+
  const std::vector<double> shear = getD(c.h, "wall_shear", 4 * c.nf), fit = getD(c.h, "wall_fit", 3 * c.nf),
                            area = getD(c.h, "wall_area", c.nf);
  double at = 0, tau = 0, yp = 0;
  for (int64_t f = 0; f < c.nf; ++f) {
    at += area[(std::size_t)f];
    tau += shear[4 * (std::size_t)f + 3];
    yp += fit[3 * (std::size_t)f + 2] * area[(std::size_t)f];
  }
  std::printf("NODALS_SST_G2_FINAL it=%d converged=%d contPost=%.12e dU=%.6e dK=%.6e dOmega=%.6e "
              "pressureDrop=%.12e wallShearMean=%.12e yPlusMean=%.6f avgPressureIts=%.3f loopMs=%.3f status=%s\n",
              last, converged ? 1 : 0, r.final_cont, r.final_du, r.final_dk, r.final_dw, r.pressure_drop,
              at > 0 ? tau / at : 0.0, at > 0 ? yp / at : 0.0, r.avg_pressure_its, loopMs,
              converged ? "PASS" : "UNCONVERGED");
  std::printf("NODALS_SOFTWARE_COMPLETION=COMPLETE iterations=%d converged=%d\n", last, converged ? 1 : 0);
  std::fflush(stdout);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::printf("NODALS_SOFTWARE_ERROR what=%s\n", e.what());
    std::fflush(stdout);
    return 90;
  }
}
