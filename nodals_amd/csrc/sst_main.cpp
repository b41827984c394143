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
// STAGE31_RESTART_FROM) are the reference's Stage31 files
// NODALS_STAGE31_STATE_it<N>.ckpt (layout at saveCheckpoint), with the
// fields in the mesh file's velocity DOF numbering; a restart checks the
// format, sizes, fingerprint (when STAGE31_INPUT_FINGERPRINT is set),
// payload checksum and finiteness.  With NODALS_FACE_YPLUS_EXPORT=1 each
// checkpoint also writes the wall faces' fitted friction velocities.
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

// Stage31 warm-restart file (modules/coupling/stage31_restart.hpp:38-117):
// a 144-byte header (magic "NODALS_S31_R1" in 16 bytes, uint32 format 1,
// uint32 value size 8, uint64 nv np nvert iteration, double prevMeanNut,
// 65-byte SHA-256 hex fingerprint, padding to 8, uint64 FNV-1a 64 hash of the
// payload), then FP64 u0 u1 u2 [nv], p [np], k omega [nv].  nv counts the
// mesh file's velocity DOFs (vertices first), so a state is stored in the
// file's own numbering; fixed DOFs carry their boundary values.
constexpr std::size_t kS31Header = 144;

std::uint64_t fnv1a(std::uint64_t h, const std::vector<double>& v) {
  const unsigned char* b = reinterpret_cast<const unsigned char*>(v.data());
  for (std::size_t i = 0; i < v.size() * sizeof(double); ++i) {
    h ^= b[i];
    h *= 1099511628211ull;
  }
  return h;
}

// (cell, basis) -> file DOF, free DOF (or -(fixed slot + 1)) and scalar entity
struct DofMaps {
  std::vector<int32_t> file, ref;
  std::vector<double> ent;
  int64_t nvel = 0, nvert = 0;
};

DofMaps dofMaps(Context& c) {
  DofMaps m;
  int64_t counts[2];
  check(nodals_h8_get(c.h, "mesh_dof_counts", counts, 2));
  m.nvel = counts[0];
  m.nvert = counts[1];
  m.file.resize((std::size_t)c.nc * 8);
  m.ref.resize((std::size_t)c.nc * 8);
  check(nodals_h8_get(c.h, "mesh_dof", m.file.data(), c.nc * 8));
  check(nodals_h8_get(c.h, "cell_ref", m.ref.data(), c.nc * 8));
  m.ent = getD(c.h, "turb_cell_ent", c.nc * 8);
  return m;
}

void saveCheckpoint(Context& c, const std::string& path, int it, const std::string& fingerprint) {
  const DofMaps m = dofMaps(c);
  std::vector<double> uf[3];
  for (int d = 0; d < 3; ++d) uf[d] = getD(c.h, ("u" + std::to_string(d)).c_str(), c.nv);
  int64_t info[8];
  check(nodals_h8_info(c.h, info));
  const std::vector<double> fixed = getD(c.h, "fixed_value", 3 * info[7]);
  const std::vector<double> p = getD(c.h, "p", c.nc), ks = getD(c.h, "k", c.ns), ws = getD(c.h, "omega", c.ns);
  std::vector<double> U[3], K((std::size_t)m.nvel), W((std::size_t)m.nvel);
  std::vector<char> seen((std::size_t)m.nvel, 0);
  for (int d = 0; d < 3; ++d) U[d].assign((std::size_t)m.nvel, 0.0);
  for (std::size_t q = 0; q < m.file.size(); ++q) {
    const std::size_t f = (std::size_t)m.file[q], e = (std::size_t)m.ent[q];
    const int32_t r = m.ref[q];
    for (int d = 0; d < 3; ++d) U[d][f] = r >= 0 ? uf[d][(std::size_t)r] : fixed[3 * (std::size_t)(-r - 1) + d];
    K[f] = ks[e];
    W[f] = ws[e];
    seen[f] = 1;
  }
  for (char x : seen) require(x != 0, "a mesh velocity DOF belongs to no cell; cannot write " + path);
  const std::vector<double>* const fields[6] = {&U[0], &U[1], &U[2], &p, &K, &W};
  std::uint64_t hash = 14695981039346656037ull;
  for (const auto* v : fields) hash = fnv1a(hash, *v);
  unsigned char hdr[kS31Header] = {};
  std::memcpy(hdr, "NODALS_S31_R1", 13);
  const uint32_t format = 1, bytes = 8;
  const uint64_t nv = (uint64_t)m.nvel, np = (uint64_t)c.nc, nvert = (uint64_t)m.nvert, iter = (uint64_t)it;
  const double prevMeanNut = -1.0;  // no plateau history is kept here
  std::memcpy(hdr + 16, &format, 4);
  std::memcpy(hdr + 20, &bytes, 4);
  std::memcpy(hdr + 24, &nv, 8);
  std::memcpy(hdr + 32, &np, 8);
  std::memcpy(hdr + 40, &nvert, 8);
  std::memcpy(hdr + 48, &iter, 8);
  std::memcpy(hdr + 56, &prevMeanNut, 8);
  std::memcpy(hdr + 64, fingerprint.data(), std::min<std::size_t>(64, fingerprint.size()));
  std::memcpy(hdr + 136, &hash, 8);
  const std::string tmp = path + ".partial";
  std::FILE* f = std::fopen(tmp.c_str(), "wb");
  require(f != nullptr, "cannot write checkpoint " + tmp);
  bool ok = std::fwrite(hdr, 1, kS31Header, f) == kS31Header;
  for (const auto* v : fields) ok = ok && std::fwrite(v->data(), 8, v->size(), f) == v->size();
  ok = (std::fclose(f) == 0) && ok;
  require(ok, "error writing checkpoint " + tmp);
  require(std::rename(tmp.c_str(), path.c_str()) == 0, "cannot rename " + tmp);
  std::printf("STAGE31_RESTART_SAVE=PASS iteration=%d path=%s fields=FULL_BF3_U_P_K_OMEGA checksum=PASS\n", it,
              path.c_str());
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
  const DofMaps m = dofMaps(c);
  std::FILE* f = std::fopen(path.c_str(), "rb");
  require(f != nullptr, "Stage31 checkpoint not found: " + path);
  unsigned char hdr[kS31Header];
  uint32_t format = 0, bytes = 0;
  uint64_t nv = 0, np = 0, nvert = 0, iter = 0, hash = 0;
  double prevMeanNut = 0;
  bool ok = std::fread(hdr, 1, kS31Header, f) == kS31Header && std::memcmp(hdr, "NODALS_S31_R1", 13) == 0;
  if (ok) {
    std::memcpy(&format, hdr + 16, 4);
    std::memcpy(&bytes, hdr + 20, 4);
    std::memcpy(&nv, hdr + 24, 8);
    std::memcpy(&np, hdr + 32, 8);
    std::memcpy(&nvert, hdr + 40, 8);
    std::memcpy(&iter, hdr + 48, 8);
    std::memcpy(&prevMeanNut, hdr + 56, 8);
    std::memcpy(&hash, hdr + 136, 8);
    ok = format == 1 && bytes == 8;
  }
  if (!ok) {
    std::fclose(f);
    throw Fail("Stage31 checkpoint header/format mismatch: " + path);
  }
  if (nv != (uint64_t)m.nvel || np != (uint64_t)c.nc || nvert != (uint64_t)m.nvert) {
    std::fclose(f);
    throw Fail("Stage31 checkpoint DOF dimensions mismatch: refusing restart (" + path + ")");
  }
  if (!fingerprint.empty() && std::string(reinterpret_cast<const char*>(hdr + 64), 64) != fingerprint) {
    std::fclose(f);
    throw Fail("Stage31 checkpoint mesh/inlet/physical-parameter SHA256 mismatch: " + path);
  }
  if (iter < 1 || iter > 100000000 || !std::isfinite(prevMeanNut)) {
    std::fclose(f);
    throw Fail("Stage31 checkpoint iteration/state invalid: " + path);
  }
  std::vector<double> U[3], p((std::size_t)np), K((std::size_t)nv), W((std::size_t)nv);
  for (int d = 0; d < 3; ++d) U[d].resize((std::size_t)nv);
  for (auto* v : {&U[0], &U[1], &U[2], &p, &K, &W}) ok = ok && std::fread(v->data(), 8, v->size(), f) == v->size();
  char extra = 0;
  const bool trailing = std::fread(&extra, 1, 1, f) == 1;
  std::fclose(f);
  require(ok && !trailing, "Stage31 checkpoint truncated or trailing data: " + path);
  std::uint64_t h = 14695981039346656037ull;
  for (const auto* v : {&U[0], &U[1], &U[2], &p, &K, &W}) h = fnv1a(h, *v);
  require(h == hash, "Stage31 checkpoint payload checksum mismatch: " + path);
  for (const auto* v : {&U[0], &U[1], &U[2], &p, &K, &W})
    for (double x : *v) require(std::isfinite(x), "Stage31 checkpoint contains nonfinite values: " + path);
  std::vector<double> uf[3], ks((std::size_t)c.ns, 0.0), ws((std::size_t)c.ns, 0.0);
  for (int d = 0; d < 3; ++d) uf[d] = getD(c.h, ("u" + std::to_string(d)).c_str(), c.nv);
  for (std::size_t q = 0; q < m.file.size(); ++q) {
    const std::size_t fd = (std::size_t)m.file[q], e = (std::size_t)m.ent[q];
    const int32_t r = m.ref[q];
    if (r >= 0)
      for (int d = 0; d < 3; ++d) uf[d][(std::size_t)r] = U[d][fd];
    ks[e] = K[fd];
    ws[e] = W[fd];
  }
  for (int d = 0; d < 3; ++d) check(nodals_h8_set(c.h, ("u" + std::to_string(d)).c_str(), uf[d].data(), c.nv));
  check(nodals_h8_set(c.h, "p", p.data(), c.nc));
  check(nodals_h8_set(c.h, "k", ks.data(), c.ns));
  check(nodals_h8_set(c.h, "omega", ws.data(), c.ns));
  std::printf("STAGE31_RESTART_LOAD=PASS path=%s iteration=%llu next=%llu mode=WARM_RESTART "
              "fields=FULL_BF3_U_P_K_OMEGA checksum=PASS\n",
              path.c_str(), (unsigned long long)iter, (unsigned long long)iter + 1);
  std::fflush(stdout);
  return (int)iter;
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
  // omega wall closure (the reference binary's default is sample_log)
  const std::string s41 = envOr("STAGE41_OMEGA_MODE", "sample_log");
  require(s41 == "sample_log" || s41 == "of_auto" || s41 == "trace_log" || s41 == "couette_point3",
          "STAGE41_OMEGA_MODE must be sample_log, of_auto or couette_point3");
  require(s41 != "couette_point3", "STAGE41_OMEGA_MODE=couette_point3 is not supported by this build");
  const std::string traceGamma = envOr("STAGE41_TRACE_GAMMA", "50");
  tu.insert(tu.end(), {"--sst-omega-wall-target", s41, "--sst-omega-trace-gamma", traceGamma});
  std::vector<const char*> tv;
  for (const auto& s : tu) tv.push_back(s.c_str());
  check(nodals_h8_turb_init(c.h, (int)tv.size(), tv.empty() ? nullptr : tv.data()));
  if (s41 == "trace_log")
    std::printf("NODALS_OMEGA_TRACE_LOG enabled=1 target=PURE_LOG_AT_INTERIOR_POINT3 "
                "placement=PHYSICAL_TET_WALL_TRACE operator=SYMMETRIC_NITSCHE mu=POINT3_NU_PLUS_SIGMAW_NUT "
                "h=THREE_TET_VOLUME_OVER_WALL_AREA gamma=%.12e kWall=NATURAL_ZERO_GRADIENT production=0 "
                "spaldingFit=SHARED_FULL_VECTOR status=EXPERIMENTAL\n",
                std::stod(traceGamma));
  std::printf("STAGE41_SAMPLE_LOG mode=%s omegaTarget=%s constraint=%s fullVectorSpalding=ON robin=OFF "
              "traceNitsche=%s SSTvolume=UNCHANGED deviceConstants=SET status=PASS\n",
              s41.c_str(), s41 == "of_auto" ? "OF_AUTO_BLEND" : "PURE_LOG",
              s41 == "trace_log" ? "PHYSICAL_TRACE_SYMMETRIC_NITSCHE" : "ORIGINAL_INTERIOR_POINT_RANK1",
              s41 == "trace_log" ? "ON" : "OFF");
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
