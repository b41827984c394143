// Test-only C wrapper around the reference's host setup headers.
//
// This file is compiled against the reference sources where they lie
// (gpu/serial_cuda/g4_host.hpp, g5_sa_host.hpp and what they include) by
// oracle/Makefile into oracle/_ref/libnodals_ref.so.  Nothing of the
// reference is copied here: the wrapper only calls the reference functions
// and exposes their outputs through a flat C interface so that the tests can
// compare the product's setup (DOF numbering, momentum CSR, coloring, cell
// plans, SA hierarchy) and the oracle against the reference itself.
#include "g4_host.hpp"
#include "g5_sa_host.hpp"

#include <cstdint>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <string>
#include <vector>

using namespace nodals_gpu;

namespace {

struct RefCase {
  SerialTetMesh mesh;
  G4SetupHost setup;
  std::vector<double> initRau;
  std::vector<double> initDelta;
  SAHierarchyHost sa;
  bool haveSa = false;
  std::map<std::string, std::vector<char>> blobs;
};

thread_local std::string g_err;

template <class T>
void put(RefCase& R, const std::string& name, const std::vector<T>& v) {
  std::vector<char> b(v.size() * sizeof(T));
  if (!v.empty()) std::memcpy(b.data(), v.data(), b.size());
  R.blobs[name] = std::move(b);
}

void publish_setup(RefCase& R) {
  const auto& S = R.setup;
  put(R, "mom_row", S.topo.row);
  put(R, "mom_col", S.topo.col);
  put(R, "mom_val", S.topo.val);
  put(R, "color", S.coloring.color);
  put(R, "color_offset", S.coloring.offset);
  put(R, "color_rows", S.coloring.rows);
  put(R, "g2free", S.momMask.g2free);
  std::vector<std::int32_t> ref;
  std::vector<std::uint8_t> slot;
  std::vector<double> det, invJ, h2, bcoef;
  for (const auto& cp : S.cells) {
    for (int a = 0; a < 8; ++a) ref.push_back(cp.ref[a]);
    for (int k = 0; k < 64; ++k) slot.push_back(cp.rowSlot[k]);
    det.push_back(cp.det);
    h2.push_back(cp.h2);
    for (int k = 0; k < 9; ++k) invJ.push_back(cp.invJ[k]);
    for (int a = 0; a < 8; ++a)
      for (int d = 0; d < 3; ++d) bcoef.push_back(g4_effective_bcoeff_host(cp, a, d));
  }
  put(R, "cell_ref", ref);
  put(R, "cell_slot", slot);
  put(R, "cell_det", det);
  put(R, "cell_h2", h2);
  put(R, "cell_invJ", invJ);
  put(R, "bcoef", bcoef);
  std::vector<double> srhs;
  for (int d = 0; d < 3; ++d) srhs.insert(srhs.end(), S.staticRhs[d].begin(), S.staticRhs[d].end());
  put(R, "static_rhs", srhs);
  put(R, "fixed_div", S.fixedDiv);
  std::vector<double> fv;
  for (const auto& v : S.fixedValue) fv.insert(fv.end(), v.begin(), v.end());
  put(R, "fixed_value", fv);
  put(R, "init_rau", R.initRau);
  put(R, "init_delta", R.initDelta);
  std::vector<double> pipe = {S.pipe.nu, S.pipe.hpDrop, S.pipe.profileScale, S.pipe.R, S.pipe.L};
  put(R, "pipe", pipe);
}

void publish_sa(RefCase& R) {
  const auto& H = R.sa;
  std::vector<std::int32_t> nl = {(int)H.csr.size(), (int)H.P.size(), H.terminal_n};
  put(R, "sa_dims", nl);
  put(R, "sa_fine_diag", H.fineDiag);
  std::vector<double> lam = {H.fineLambda};
  lam.insert(lam.end(), H.levelLambda.begin(), H.levelLambda.end());
  put(R, "sa_lambda", lam);
  put(R, "sa_terminal_inv", H.terminal_inv);
  for (std::size_t l = 0; l < H.agg.size(); ++l) put(R, "sa_agg" + std::to_string(l), H.agg[l].id);
  for (std::size_t l = 0; l < H.P.size(); ++l) {
    put(R, "sa_P" + std::to_string(l) + "_row", H.P[l].row);
    put(R, "sa_P" + std::to_string(l) + "_col", H.P[l].col);
    put(R, "sa_P" + std::to_string(l) + "_val", H.P[l].val);
  }
  for (std::size_t l = 0; l < H.csr.size(); ++l) {
    put(R, "sa_A" + std::to_string(l) + "_row", H.csr[l].row);
    put(R, "sa_A" + std::to_string(l) + "_col", H.csr[l].col);
    put(R, "sa_A" + std::to_string(l) + "_val", H.csr[l].val);
  }
}

}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

// Loads the polyMesh, runs build_g4_setup (g4_host.hpp:289) and the
// setup-rAU assembly at U=0 (h8_main_base.cu:1917-1922).
void* ref_open(const char* mesh_dir, double re, double bulk, const char* wall,
               const char* inlet, const char* outlet, double alpha_u) {
  try {
    auto R = std::make_unique<RefCase>();
    R->mesh = load_foam_tet_mesh(mesh_dir);
    R->setup = build_g4_setup(R->mesh, re, bulk, wall, inlet, outlet, false, false);
    std::array<std::vector<double>, 3> U0;
    for (auto& u : U0) u.assign((std::size_t)R->setup.topo.n, 0.0);
    std::vector<double> hostA;
    std::array<std::vector<double>, 3> conv;
    host_assemble_central_g4(R->setup, U0, hostA, conv);
    R->initRau = host_finalize_relax_g4(R->setup, hostA, alpha_u, &R->initDelta);
    R->setup.pressure.rAU = R->initRau;
    publish_setup(*R);
    return R.release();
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

// The same setup with the DG weak inlet (build_g4_setup's dgInlet): inlet
// DOFs free, effective B and prescribed-flux fixed divergence; the setup
// assembly starts from the plug on every free vertex (h8_main_base.cu:1918).
void* ref_open_dg(const char* mesh_dir, double re, double bulk, const char* wall, const char* inlet,
                  const char* outlet, double alpha_u) {
  try {
    auto R = std::make_unique<RefCase>();
    R->mesh = load_foam_tet_mesh(mesh_dir);
    R->setup = build_g4_setup(R->mesh, re, bulk, wall, inlet, outlet, true, false);
    std::array<std::vector<double>, 3> U0;
    for (auto& u : U0) u.assign((std::size_t)R->setup.topo.n, 0.0);
    for (std::size_t v = 0; v < R->mesh.points.size(); ++v) {
      const int g = R->setup.momMask.g2free[v];
      if (g < 0) continue;
      for (int d = 0; d < 3; ++d) U0[(std::size_t)d][(std::size_t)g] = R->setup.pipe.inletVelocity[d];
    }
    std::vector<double> hostA;
    std::array<std::vector<double>, 3> conv;
    host_assemble_central_g4(R->setup, U0, hostA, conv);
    R->initRau = host_finalize_relax_g4(R->setup, hostA, alpha_u, &R->initDelta);
    R->setup.pressure.rAU = R->initRau;
    publish_setup(*R);
    return R.release();
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void ref_close(void* h) { delete static_cast<RefCase*>(h); }

// Empty handle for the mesh-free entry points (ref_sa_from_csr).
void* ref_new() { return new RefCase(); }

// Byte size of a published array, or -1.
long long ref_size(void* h, const char* name) {
  auto* R = static_cast<RefCase*>(h);
  auto it = R->blobs.find(name);
  return it == R->blobs.end() ? -1 : (long long)it->second.size();
}

int ref_get(void* h, const char* name, void* out) {
  auto* R = static_cast<RefCase*>(h);
  auto it = R->blobs.find(name);
  if (it == R->blobs.end()) return 1;
  if (!it->second.empty()) std::memcpy(out, it->second.data(), it->second.size());
  return 0;
}

// build_sa_hierarchy (g5_sa_host.hpp:1545) on the setup-rAU pressure operator.
int ref_build_sa(void* h, int terminal, int power_its, double safety, double low_frac,
                 double damping) {
  try {
    auto* R = static_cast<RefCase*>(h);
    R->sa = build_sa_hierarchy(R->mesh, R->setup.pressure, 16, 6, 18, terminal, 8, power_its,
                               safety, low_frac, damping);
    R->haveSa = true;
    publish_sa(*R);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// cpu_schur_apply (g2_host.hpp:77): y = B diag(rAU) B^T x with the setup rAU.
int ref_schur_apply(void* h, const double* x, double* y) {
  try {
    auto* R = static_cast<RefCase*>(h);
    std::vector<double> xv(x, x + R->setup.pressure.cells.size()), yv;
    cpu_schur_apply(R->setup.pressure, xv, yv);
    std::memcpy(y, yv.data(), yv.size() * sizeof(double));
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// host_assemble_central_g4 + host_finalize_relax_g4 (g4_host.hpp:317-422)
// at a caller-supplied free-velocity state.
int ref_assemble(void* h, const double* u0, const double* u1, const double* u2, double alpha_u,
                 double* aval, double* conv, double* rau, double* delta) {
  try {
    auto* R = static_cast<RefCase*>(h);
    const std::size_t n = (std::size_t)R->setup.topo.n;
    std::array<std::vector<double>, 3> U;
    U[0].assign(u0, u0 + n);
    U[1].assign(u1, u1 + n);
    U[2].assign(u2, u2 + n);
    std::vector<double> A;
    std::array<std::vector<double>, 3> C;
    host_assemble_central_g4(R->setup, U, A, C);
    std::vector<double> d;
    auto r = host_finalize_relax_g4(R->setup, A, alpha_u, &d);
    std::memcpy(aval, A.data(), A.size() * sizeof(double));
    for (int k = 0; k < 3; ++k) std::memcpy(conv + k * n, C[k].data(), n * sizeof(double));
    std::memcpy(rau, r.data(), n * sizeof(double));
    std::memcpy(delta, d.data(), n * sizeof(double));
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// SA hierarchy of an explicit CSR operator built from the reference's
// coarse-level functions, composed as the level loop of build_sa_hierarchy
// (g5_sa_host.hpp:1570-1577) applied from level 0: sa_power_csr,
// aggregate_graph(csr_graph), build_sa_explicit_transfer_reference,
// build_sa_ptap_reference, dense_inverse_cholesky.  Results are published
// as csr_agg<l>, csr_P<l>_*, csr_A<l>_*, csr_terminal_inv, csr_dims.
int ref_sa_from_csr(void* h, int n, const long long* row, const int* col, const double* val, int terminal,
                    int power_its, double safety, double damping) {
  try {
    auto* R = static_cast<RefCase*>(h);
    CSRHost A;
    A.n = n;
    A.row.assign(row, row + n + 1);
    A.col.assign(col, col + row[n]);
    A.val.assign(val, val + row[n]);
    A.diag.assign((std::size_t)n, 0.0);
    for (int i = 0; i < n; ++i)
      for (long long k = row[i]; k < row[i + 1]; ++k)
        if (col[k] == i) A.diag[(std::size_t)i] = val[k];
    std::vector<CSRHost> levels;
    std::vector<SATransferHost> P;
    std::vector<AggResult> agg;
    const CSRHost* cur = &A;
    int lev = 0;
    do {
      const double lam = sa_power_csr(*cur, power_its, safety);
      agg.push_back(aggregate_graph(csr_graph(*cur), 16, 6, 18));
      P.push_back(build_sa_explicit_transfer_reference(*cur, agg.back(), lam, damping, 8, lev));
      levels.push_back(build_sa_ptap_reference(*cur, P.back(), lev + 1));
      cur = &levels.back();
      ++lev;
    } while (levels.back().n > terminal);
    for (std::size_t l = 0; l < levels.size(); ++l) {
      put(*R, "csr_agg" + std::to_string(l), agg[l].id);
      put(*R, "csr_P" + std::to_string(l) + "_row", P[l].row);
      put(*R, "csr_P" + std::to_string(l) + "_col", P[l].col);
      put(*R, "csr_P" + std::to_string(l) + "_val", P[l].val);
      put(*R, "csr_A" + std::to_string(l) + "_row", levels[l].row);
      put(*R, "csr_A" + std::to_string(l) + "_col", levels[l].col);
      put(*R, "csr_A" + std::to_string(l) + "_val", levels[l].val);
    }
    put(*R, "csr_terminal_inv", dense_inverse_cholesky(levels.back()));
    std::vector<std::int32_t> dims = {(int)levels.size(), levels.back().n};
    put(*R, "csr_dims", dims);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// pressure_drop_fit_g4 (g4_host.hpp:424) on a cell pressure field.
double ref_pressure_drop(void* h, const double* p) {
  auto* R = static_cast<RefCase*>(h);
  std::vector<double> pv(p, p + R->mesh.tets.size());
  return pressure_drop_fit_g4(R->mesh, pv);
}

}  // extern "C"
