// extern "C" entry points of include/nodals_hip.h.  Every call catches all
// exceptions and maps them to a status code plus a thread-local message.
#include "../../../include/nodals_hip.h"

#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "h8_solver.hpp"
#include "gnn.hpp"
#include "sst_kernels.hpp"
#include "sys_solver.hpp"
#include "../host/ray_trace.hpp"
#include "wall_state.hpp"

struct nodals_h8 {
  std::unique_ptr<nodals::H8Solver> s;
};
struct nodals_sys {
  std::unique_ptr<nodals::SysSolver> s;
};
struct nodals_gnn {
  std::unique_ptr<nodals::GnnModel> g;
};

namespace {

thread_local std::string g_error;

template <class F>
int guarded(F&& f) {
  try {
    g_error.clear();
    f();
    return nodals::kOk;
  } catch (const nodals::Error& e) {
    g_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_error = "host allocation failed";
    return nodals::kErrInternal;
  } catch (const std::exception& e) {
    g_error = e.what();
    return nodals::kErrInternal;
  } catch (...) {
    g_error = "unknown exception";
    return nodals::kErrInternal;
  }
}

nodals::H8Solver& solver(nodals_h8* h) {
  nodals::require(h != nullptr && h->s != nullptr, "null H8 handle");
  return *h->s;
}
nodals::SysSolver& sys(nodals_sys* h) {
  nodals::require(h != nullptr && h->s != nullptr, "null system handle");
  return *h->s;
}

template <class T>
void copyOut(const std::vector<T>& v, void* host, int64_t count) {
  nodals::require(host != nullptr, "null host pointer");
  nodals::require(count == (int64_t)v.size(),
                  "count mismatch: expected " + std::to_string(v.size()) + " got " + std::to_string(count));
  if (!v.empty()) std::memcpy(host, v.data(), v.size() * sizeof(T));
}

void fillResult(const nodals::dev::KrylovResult& r, nodals_solve_result* res) {
  if (!res) return;
  res->its = r.its;
  res->ok = r.ok ? 1 : 0;
  res->rel = r.rel;
  res->r0 = r.r0;
}

}  // namespace

namespace nodals {
namespace dev {
void hexElementBlocks(long long n, const double* coords, const double* u, double nu, double* diff, double* conv,
                      int* badOut);
}  // namespace dev
}  // namespace nodals

extern "C" {

const char* nodals_last_error(void) { return g_error.c_str(); }
const char* nodals_version(void) { return "nodals_amd 0.1 (gfx950)"; }

int nodals_h8_create(int argc, const char* const* argv, nodals_h8** out) {
  return guarded([&]() {
    nodals::require(out != nullptr, "null output handle");
    *out = nullptr;
    const nodals::H8Options o = nodals::parseH8Args(argc, argv);
    auto h = std::make_unique<nodals_h8>();
    h->s = std::make_unique<nodals::H8Solver>(o);
    *out = h.release();
  });
}

int nodals_h8_create_host(int argc, const char* const* argv, nodals_h8** out) {
  return guarded([&]() {
    nodals::require(out != nullptr, "null output handle");
    *out = nullptr;
    const nodals::H8Options o = nodals::parseH8Args(argc, argv);
    auto h = std::make_unique<nodals_h8>();
    h->s = std::make_unique<nodals::H8Solver>(o, false);
    *out = h.release();
  });
}

void nodals_h8_destroy(nodals_h8* h) {
  try {
    delete h;
  } catch (...) {
  }
}

int nodals_h8_info(nodals_h8* h, int64_t info[8]) {
  return guarded([&]() {
    auto& s = solver(h);
    nodals::require(info != nullptr, "null info");
    info[0] = s.setup().nc;
    info[1] = s.setup().nv;
    info[2] = (int64_t)s.setup().mom.col.size();
    info[3] = s.ncolors();
    info[4] = (int64_t)s.fineTopo().col.size();
    info[5] = (int64_t)s.hierarchy().A.size();
    info[6] = s.hierarchy().terminalN;
    info[7] = s.setup().nFixed;
  });
}

int nodals_h8_assemble(nodals_h8* h, double alpha_u, double rau_scale) {
  return guarded([&]() { solver(h).assemble(alpha_u, rau_scale); });
}

int nodals_h8_momentum_work(nodals_h8* h, const char* mode, double omega, int32_t outer) {
  return guarded([&]() {
    nodals::require(mode != nullptr, "null mode");
    solver(h).momentumWork(mode, omega, outer);
  });
}

int nodals_h8_momentum_audit(nodals_h8* h, double rel[3]) {
  return guarded([&]() {
    nodals::require(rel != nullptr, "null audit output");
    const auto q = solver(h).momentumAudit();
    for (int d = 0; d < 3; ++d) rel[d] = q[(std::size_t)d];
  });
}

int nodals_h8_continuity(nodals_h8* h, double* norm) {
  return guarded([&]() {
    const double n = solver(h).continuity();
    if (norm) *norm = n;
  });
}

int nodals_h8_pressure_refresh(nodals_h8* h) { return guarded([&]() { solver(h).refresh(); }); }

int nodals_h8_pressure_solve(nodals_h8* h, const char* solverName, double rtol, double atol, int32_t maxit,
                             nodals_solve_result* res) {
  return guarded([&]() {
    nodals::require(solverName != nullptr, "null solver name");
    const auto r = solver(h).pressureSolve(solverName, rtol, atol, maxit);
    fillResult(r, res);
  });
}

int nodals_h8_momentum_solve(nodals_h8* h, const char* mode, int32_t sweeps, double omega, nodals_solve_result* res) {
  return guarded([&]() {
    nodals::require(mode != nullptr, "null momentum mode");
    fillResult(solver(h).momentumSolve(mode, sweeps, omega), res);
  });
}

int nodals_h8_pcg_amg(nodals_h8* h, const double* rhs, double* x, double rtol, double atol, int32_t maxit,
                      nodals_solve_result* res) {
  return guarded([&]() { fillResult(solver(h).pcgAmg(rhs, x, rtol, atol, maxit), res); });
}

int nodals_h8_scalar_solve(nodals_h8* h, int32_t equation, nodals_solve_result* res) {
  return guarded([&]() { fillResult(solver(h).scalarSolve(equation), res); });
}

int nodals_h8_pressure_update(nodals_h8* h, double alpha_p) {
  return guarded([&]() { solver(h).pressureUpdate(alpha_p); });
}

namespace {
void fillRunResult(const nodals::H8RunResult& r, nodals_h8_run_result* res) {
  if (!res) return;
  res->outer = r.outer;
  res->converged = r.converged;
  res->pass = r.pass;
  res->reserved = 0;
  res->final_rel_cont = r.finalRelCont;
  res->avg_pressure_its = r.avgPressureIts;
  res->pressure_drop = r.pressureDrop;
  res->exact_pressure_drop = r.exactPressureDrop;
  res->loop_ms = r.loopMs;
  res->assembly_ms = r.assemblyMs;
  res->momentum_ms = r.momentumMs;
  res->continuity_ms = r.continuityMs;
  res->pressure_ms = r.pressureMs;
  res->pupdate_ms = r.pupdateMs;
  res->setup_ms = r.setupMs;
  res->pressure_its = r.pressureIts;
  res->warm_starts = r.warmStarts;
  res->fallbacks = r.fallbacks;
  res->recorded = r.recorded;
  res->reserved2 = 0;
  res->turb_ms = r.turbMs;
  res->final_du = r.finalDu;
  res->final_dk = r.finalDk;
  res->final_dw = r.finalDw;
  res->final_cont = r.finalCont;
}
}  // namespace

int nodals_h8_run(nodals_h8* h, nodals_h8_run_result* res) {
  return guarded([&]() { fillRunResult(solver(h).run(), res); });
}

int nodals_h8_run_window(nodals_h8* h, int32_t first_it, int32_t count, nodals_h8_run_result* res) {
  return guarded([&]() { fillRunResult(solver(h).runWindow(first_it, count), res); });
}

int nodals_h8_write_vtu(nodals_h8* h, const char* path) {
  return guarded([&]() {
    nodals::require(path != nullptr, "null VTU path");
    solver(h).writeVtu(path);
  });
}

// CF/PMIS split data: cf_state<l> (uint8, 1 = C, 2 = F), cf_cid<l> (int32
// coarse id or -1) and cf_rounds (int32, one per split).
static void copyCf(const nodals::AmgHierarchy& H, const std::string& n, void* host, int64_t count) {
  if (n == "cf_rounds") {
    copyOut(std::vector<int32_t>(H.cfRounds.begin(), H.cfRounds.end()), host, count);
    return;
  }
  const bool st = n.rfind("cf_state", 0) == 0;
  nodals::require(st || n.rfind("cf_cid", 0) == 0, "unknown field " + n);
  const std::size_t l = (std::size_t)std::stoi(n.substr(st ? 8 : 6));
  nodals::require(l < H.cfState.size(), "CF split index out of range (or not a CF hierarchy)");
  if (st) copyOut(H.cfState[l], host, count);
  else copyOut(H.cfCoarseId[l], host, count);
}

int nodals_h8_get(nodals_h8* h, const char* name, void* host, int64_t count) {
  return guarded([&]() {
    auto& s = solver(h);
    nodals::require(name != nullptr, "null name");
    const std::string n = name;
    if (n == "u0" || n == "u1" || n == "u2") copyOut(s.getVelocity(n[1] - '0'), host, count);
    else if (n == "b0" || n == "b1" || n == "b2") copyOut(s.getBrhs(n[1] - '0'), host, count);
    else if (n == "rau" || n == "delta" || n == "diag") copyOut(s.getRowField(n), host, count);
    else if (n == "p" || n == "cont" || n == "dp" || n == "fine_diag" || n == "fine_l1") copyOut(s.getCellField(n), host, count);
    else if (n == "mom_val") copyOut(s.getMomentumValues(), host, count);
    else if (n == "fine_val") copyOut(s.getFineValues(), host, count);
    else if (n.rfind("wall_", 0) == 0 && n != "wall_distance") copyOut(s.wallField(n), host, count);
    else if (n == "mom_row") copyOut(s.setup().mom.row, host, count);
    else if (n == "mom_col") copyOut(s.setup().mom.col, host, count);
    else if (n == "mom_diff") copyOut(s.setup().mom.val, host, count);
    else if (n == "color") copyOut(s.setup().coloring.color, host, count);
    else if (n == "g2free") copyOut(s.setup().g2free, host, count);
    else if (n == "fine_row") copyOut(s.fineTopo().row, host, count);
    else if (n == "fine_col") copyOut(s.fineTopo().col, host, count);
    else if (n == "init_rau") copyOut(s.setup().initRau, host, count);
    else if (n == "static_rhs") {
      std::vector<double> v;
      for (int d = 0; d < 3; ++d) v.insert(v.end(), s.setup().staticRhs[(std::size_t)d].begin(), s.setup().staticRhs[(std::size_t)d].end());
      copyOut(v, host, count);
    } else if (n == "fixed_div") copyOut(s.setup().fixedDiv, host, count);
    else if (n == "bcoef") copyOut(s.setup().bcoef, host, count);
    else if (n == "cell_det" || n == "cell_invJ") {
      std::vector<double> v;
      for (const auto& cp : s.setup().cells) {
        if (n == "cell_det") v.push_back(cp.det);
        else v.insert(v.end(), cp.invJ, cp.invJ + 9);
      }
      copyOut(v, host, count);
    } else if (n == "fixed_value") {
      std::vector<double> v;
      for (const auto& f : s.setup().fixedValue) v.insert(v.end(), f.begin(), f.end());
      copyOut(v, host, count);
    } else if (n == "pipe") {
      const auto& P = s.setup().pipe;
      copyOut(std::vector<double>{P.nu, P.hpDrop, P.profileScale, P.R, P.L}, host, count);
    } else if (n == "pipe_geometry") {
      const auto& P = s.setup().pipe;
      copyOut(std::vector<double>{P.cx, P.cy, P.R, P.D, P.bulk, P.nu, P.inletVelocity[0], P.inletVelocity[1],
                                  P.inletVelocity[2], (double)P.inlet, (double)P.wall, (double)P.outlet},
              host, count);
    } else if (n == "mesh_sizes") {
      const auto& M = s.mesh();
      copyOut(std::vector<int64_t>{M.nPoints(), M.nFaces(), M.nInternal(), M.nCells()}, host, count);
    } else if (n == "mesh_points") {
      std::vector<double> v;
      for (const auto& X : s.mesh().points) v.insert(v.end(), X.begin(), X.end());
      copyOut(v, host, count);
    } else if (n == "mesh_tets" || n == "mesh_opp" || n == "mesh_faces" || n == "mesh_owner" ||
               n == "mesh_neighbour" || n == "mesh_face_patch") {
      const auto& M = s.mesh();
      std::vector<int32_t> v;
      if (n == "mesh_tets")
        for (const auto& t : M.tets) v.insert(v.end(), t.begin(), t.end());
      else if (n == "mesh_opp")
        for (const auto& t : M.oppFace) v.insert(v.end(), t.begin(), t.end());
      else if (n == "mesh_faces")
        for (const auto& f : M.faces) v.insert(v.end(), f.begin(), f.end());
      else if (n == "mesh_owner") v = M.owner;
      else if (n == "mesh_neighbour") v = M.neighbour;
      else v = M.facePatch;
      copyOut(v, host, count);
    } else if (n == "wall_distance") {
      copyOut(s.setup().wallDistance, host, count);
    } else if (n == "fixed_slot") {
      copyOut(s.setup().fixedSlot, host, count);
    } else if (n == "k" || n == "omega" || n.rfind("turb_", 0) == 0) {
      copyOut(s.turbField(n), host, count);
    }
    else if (n == "rans_stats") copyOut(s.ransStats(), host, count);
    else if (n == "dg_inlet_count") copyOut(std::vector<int64_t>{(int64_t)s.setup().dgFaces.size()}, host, count);
    else if (n == "dg_inlet_faces") {
      std::vector<int32_t> v;
      for (const auto& F : s.setup().dgFaces) v.insert(v.end(), {F.cell, F.lf});
      copyOut(v, host, count);
    } else if (n == "plug") copyOut(s.setup().plug, host, count);
    else if (n == "cell_points") {
      const auto& M = s.mesh();
      std::vector<double> v;
      for (const auto& t : M.tets)
        for (int a = 0; a < 4; ++a) {
          const auto& X = M.points[(std::size_t)t[(std::size_t)a]];
          v.insert(v.end(), X.begin(), X.end());
        }
      copyOut(v, host, count);
    } else if (n == "mesh_dof") {
      std::vector<int32_t> v;
      for (const auto& d : s.mesh().fileDof) v.insert(v.end(), d.begin(), d.end());
      nodals::require(!v.empty(), "the mesh has no file DOF numbering");
      copyOut(v, host, count);
    } else if (n == "mesh_dof_counts") {
      copyOut(std::vector<int64_t>{s.mesh().fileNvel, s.mesh().fileNvert}, host, count);
    } else if (n == "cell_ref") {
      std::vector<int32_t> v;
      for (const auto& cp : s.setup().cells) v.insert(v.end(), cp.ref, cp.ref + 8);
      copyOut(v, host, count);
    } else if (n == "cell_slot") {
      std::vector<uint8_t> v;
      for (const auto& cp : s.setup().cells) v.insert(v.end(), cp.rowSlot, cp.rowSlot + 64);
      copyOut(v, host, count);
    } else if (n.rfind("agg", 0) == 0) {
      const std::size_t l = (std::size_t)std::stoi(n.substr(3));
      nodals::require(l < s.hierarchy().agg.size(), "aggregate level out of range");
      copyOut(s.hierarchy().agg[l].id, host, count);
    } else if (n.rfind("cf_", 0) == 0) {
      copyCf(s.hierarchy(), n, host, count);
    } else if (n == "cell_xy") {
      // (x, y) of each cell's four vertices relative to the pipe axis
      const auto& M = s.mesh();
      const auto& P = s.setup().pipe;
      std::vector<double> v;
      v.reserve((std::size_t)s.setup().nc * 8);
      for (int c = 0; c < s.setup().nc; ++c)
        for (int a = 0; a < 4; ++a) {
          const auto& X = M.points[(std::size_t)M.tets[(std::size_t)c][(std::size_t)a]];
          v.push_back(X[0] - P.cx);
          v.push_back(X[1] - P.cy);
        }
      copyOut(v, host, count);
    } else if (n == "fine_lanes") {
      copyOut(std::vector<int32_t>{s.fineLanes()}, host, count);
    } else if (n == "pin_cell") {
      copyOut(std::vector<int32_t>{s.setup().pinCell}, host, count);
    } else if (n == "cell_nbr_off" || n == "cell_nbr") {
      const nodals::CellGraph g = s.cellGraph();
      copyOut(n == "cell_nbr" ? g.nbr : g.off, host, count);
    } else if (n == "exact_fine_val") {
      copyOut(nodals::exactFineCsr(s.setup(), s.fineTopo(), s.setup().initRau).val, host, count);
    } else if (n.rfind("P", 0) == 0 || n.rfind("A", 0) == 0) {
      // P<l>_row|col|val and A<l>_row|col|val of the host hierarchy
      const auto u = n.find('_');
      nodals::require(u != std::string::npos, "unknown field " + n);
      const std::size_t l = (std::size_t)std::stoi(n.substr(1, u - 1));
      const std::string part = n.substr(u + 1);
      const auto& H = s.hierarchy();
      if (n[0] == 'P') {
        nodals::require(l < H.P.size(), "transfer level out of range");
        if (part == "row") copyOut(H.P[l].row, host, count);
        else if (part == "col") copyOut(H.P[l].col, host, count);
        else copyOut(H.P[l].val, host, count);
      } else {
        nodals::require(l < H.A.size(), "operator level out of range");
        if (part == "row") copyOut(H.A[l].row, host, count);
        else if (part == "col") copyOut(H.A[l].col, host, count);
        else copyOut(H.A[l].val, host, count);
      }
    } else if (n == "fine_lambda") {
      copyOut(std::vector<double>{s.hierarchy().fineLambda}, host, count);
    } else if (n == "terminal_inv") {
      copyOut(s.hierarchy().terminalInv, host, count);
    } else if (n == "level_lambda") {
      copyOut(s.hierarchy().levelLambda, host, count);
    } else if (n == "cell_z") {
      copyOut(s.setup().cellZ, host, count);
    } else {
      nodals::fail("unknown field " + n);
    }
  });
}

int nodals_h8_set(nodals_h8* h, const char* name, const void* host, int64_t count) {
  return guarded([&]() {
    auto& s = solver(h);
    nodals::require(name != nullptr && host != nullptr, "null argument");
    const std::string n = name;
    const double* d = static_cast<const double*>(host);
    std::vector<double> v(d, d + count);
    if (n == "u0" || n == "u1" || n == "u2") s.setVelocity(n[1] - '0', v);
    else if (n == "k" || n == "omega") s.setTurbField(n, v);
    else s.setCellField(n, v);
  });
}

int nodals_h8_fine_spmv(nodals_h8* h, const double* x, double* y) {
  return guarded([&]() {
    auto& s = solver(h);
    const int nc = s.setup().nc;
    NODALS_HIP(hipMemcpyAsync(s.devScratchA(), x, sizeof(double) * nc, hipMemcpyHostToDevice, s.stream()));
    s.fineSpmv(s.devScratchA(), s.devScratchB());
    NODALS_HIP(hipMemcpyAsync(y, s.devScratchB(), sizeof(double) * nc, hipMemcpyDeviceToHost, s.stream()));
    s.sync();
  });
}

int nodals_h8_precond(nodals_h8* h, const double* r, double* z) {
  return guarded([&]() {
    auto& s = solver(h);
    nodals::require(r != nullptr && z != nullptr, "null argument");
    const int nc = s.setup().nc;
    NODALS_HIP(hipMemcpyAsync(s.devScratchA(), r, sizeof(double) * nc, hipMemcpyHostToDevice, s.stream()));
    s.precond(s.devScratchA(), s.devScratchB());
    NODALS_HIP(hipMemcpyAsync(z, s.devScratchB(), sizeof(double) * nc, hipMemcpyDeviceToHost, s.stream()));
    s.sync();
  });
}

int nodals_h8_time_kernel(nodals_h8* h, const char* kernel, int32_t reps, double* ms_avg, double* bytes) {
  return guarded([&]() {
    nodals::require(kernel != nullptr, "null kernel name");
    const double ms = solver(h).timeKernel(kernel, reps, bytes);
    if (ms_avg) *ms_avg = ms;
  });
}

int nodals_h8_probe(nodals_h8* h, int32_t enable, int64_t* launches, double* total_ms) {
  return guarded([&]() {
    if (enable) {
      solver(h).probeStart(enable);
      return;
    }
    long long n = 0;
    double ms = 0;
    solver(h).probeStop(&n, &ms);
    if (launches) *launches = n;
    if (total_ms) *total_ms = ms;
  });
}

int nodals_sys_create(int32_t n, const int64_t* row, const int32_t* col, const double* val, const char* options,
                      nodals_sys** out) {
  return guarded([&]() {
    nodals::require(out != nullptr, "null output handle");
    *out = nullptr;
    auto h = std::make_unique<nodals_sys>();
    h->s = std::make_unique<nodals::SysSolver>(n, row, col, val, options ? options : "");
    *out = h.release();
  });
}

int nodals_sys_create_host(int32_t n, const int64_t* row, const int32_t* col, const double* val,
                           const char* options, nodals_sys** out) {
  return guarded([&]() {
    nodals::require(out != nullptr, "null output handle");
    *out = nullptr;
    auto h = std::make_unique<nodals_sys>();
    h->s = std::make_unique<nodals::SysSolver>(n, row, col, val, options ? options : "", false);
    *out = h.release();
  });
}

void nodals_sys_destroy(nodals_sys* s) {
  try {
    delete s;
  } catch (...) {
  }
}

int nodals_sys_info(nodals_sys* h, int64_t info[16]) {
  return guarded([&]() {
    auto& s = sys(h);
    for (int k = 0; k < 16; ++k) info[k] = 0;
    info[0] = s.n();
    info[1] = s.nnz();
    info[2] = s.levels();
    info[3] = s.hierarchy().terminalN;
    for (int l = 0; l < s.levels() && l < 12; ++l) info[4 + l] = s.hierarchy().A[(std::size_t)l].n;
  });
}

int nodals_sys_get(nodals_sys* h, const char* name, void* host, int64_t count) {
  return guarded([&]() {
    auto& s = sys(h);
    nodals::require(name != nullptr, "null name");
    const std::string n = name;
    const auto& H = s.hierarchy();
    if (n.rfind("agg", 0) == 0) {
      const std::size_t l = (std::size_t)std::stoi(n.substr(3));
      nodals::require(l < H.agg.size(), "aggregate level out of range");
      copyOut(H.agg[l].id, host, count);
    } else if (n.rfind("cf_", 0) == 0) {
      copyCf(H, n, host, count);
    } else if (n == "terminal_inv") {
      copyOut(H.terminalInv, host, count);
    } else if (n == "level_lambda") {
      copyOut(H.levelLambda, host, count);
    } else if (n == "fine_lambda") {
      copyOut(std::vector<double>{H.fineLambda}, host, count);
    } else if (n[0] == 'P' || n[0] == 'A') {
      const auto u = n.find('_');
      nodals::require(u != std::string::npos, "unknown field " + n);
      const std::size_t l = (std::size_t)std::stoi(n.substr(1, u - 1));
      const std::string part = n.substr(u + 1);
      if (n[0] == 'P') {
        nodals::require(l < H.P.size(), "transfer level out of range");
        if (part == "row") copyOut(H.P[l].row, host, count);
        else if (part == "col") copyOut(H.P[l].col, host, count);
        else copyOut(H.P[l].val, host, count);
      } else {
        nodals::require(l < H.A.size(), "operator level out of range");
        if (part == "row") copyOut(H.A[l].row, host, count);
        else if (part == "col") copyOut(H.A[l].col, host, count);
        else copyOut(H.A[l].val, host, count);
      }
    } else {
      nodals::fail("unknown field " + n);
    }
  });
}

int nodals_sys_spmv(nodals_sys* h, const double* x, double* y) {
  return guarded([&]() {
    auto& s = sys(h);
    const auto r = s.spmvHost(std::vector<double>(x, x + s.n()));
    std::memcpy(y, r.data(), sizeof(double) * r.size());
  });
}

int nodals_sys_precond(nodals_sys* h, const double* r, double* z) {
  return guarded([&]() {
    auto& s = sys(h);
    const auto v = s.precondHost(std::vector<double>(r, r + s.n()));
    std::memcpy(z, v.data(), sizeof(double) * v.size());
  });
}

int nodals_sys_solve(nodals_sys* h, const double* b, double* x, const char* solverName, double rtol, int32_t restart,
                     int32_t maxit, nodals_solve_result* res) {
  return guarded([&]() {
    auto& s = sys(h);
    nodals::require(b && x && solverName, "null argument");
    std::vector<double> xv;
    const auto r = s.solve(std::vector<double>(b, b + s.n()), xv, solverName, rtol, restart, maxit);
    std::memcpy(x, xv.data(), sizeof(double) * xv.size());
    fillResult(r, res);
  });
}

int nodals_sys_set_rhs(nodals_sys* h, const double* b) {
  return guarded([&]() {
    auto& s = sys(h);
    s.setRhs(std::vector<double>(b, b + s.n()));
  });
}

int nodals_sys_solve_timed(nodals_sys* h, const char* solverName, double rtol, int32_t restart, int32_t maxit,
                           int32_t reps, double* ms_avg, nodals_solve_result* res) {
  return guarded([&]() {
    nodals::dev::KrylovResult r;
    const double ms = sys(h).timedSolve(solverName, rtol, restart, maxit, reps, &r);
    if (ms_avg) *ms_avg = ms;
    fillResult(r, res);
  });
}

int nodals_sys_time_spmv(nodals_sys* h, int32_t level, int32_t reps, double* ms_avg, double* bytes) {
  return guarded([&]() {
    const double ms = sys(h).timeSpmv(level, reps, bytes);
    if (ms_avg) *ms_avg = ms;
  });
}

int nodals_sst_local(int64_t n, const double* dist, const double* pipe_xy, double radius, double nu,
                     const double* gu, const double* k, const double* omega, const double* grad_k,
                     const double* grad_omega, double k_floor, double omega_floor, int32_t deviatoric,
                     double* out, int32_t* floored) {
  return guarded([&]() {
    nodals::sstLocalBatch(n, dist, pipe_xy, radius, nu, gu, k, omega, grad_k, grad_omega, k_floor, omega_floor,
                          deviatoric, out, floored);
  });
}

int nodals_wall_spalding(int64_t n, const double* slip, const double* y, double nu, double* out, int32_t* ok) {
  return guarded([&]() { nodals::spaldingBatch(n, slip, y, nu, out, ok); });
}

int nodals_wall_profile5(int64_t n, const double* U, const double* U_signed, const double* y, double nu,
                         double* out, int32_t* valid) {
  return guarded([&]() { nodals::profile5Batch(n, U, U_signed, y, nu, out, valid); });
}

int nodals_wall_profile5_noalloc(int64_t n, const double* U, const double* U_signed, const double* y, double nu,
                                 double* out, int32_t* valid) {
  return guarded([&]() { nodals::profile5BatchNoAlloc(n, U, U_signed, y, nu, out, valid); });
}

int nodals_wall_fit_tangential(int64_t n, const double* samples, const double* y, const double* nrm,
                               const double* pic, double nu, int32_t count, int32_t zero_limit, double* out,
                               int32_t* valid) {
  return guarded([&]() { nodals::fitTangentialBatch(n, samples, y, nrm, pic, nu, count, zero_limit, out, valid); });
}

int nodals_hex_element_blocks(int64_t n, const double* coords, const double* u, double nu, double* diff,
                              double* conv, int32_t* bad) {
  return guarded([&]() {
    int b = 0;
    nodals::dev::hexElementBlocks(n, coords, u, nu, diff, conv, &b);
    if (bad) *bad = b;
  });
}

int nodals_ray_trace(int64_t np, const double* points, int64_t nt, const int32_t* tets, int64_t nrays,
                     const int32_t* start, const double* origin, const double* dir, int32_t ndist, const double* dist,
                     int32_t max_crossings, double tol, int32_t* cell, double* bary, int32_t* crossings,
                     int32_t* status, double* certified) {
  return guarded([&]() {
    nodals::require(np > 0 && nt > 0 && nrays >= 0 && ndist >= 1, "invalid ray-trace sizes");
    nodals::require(points && tets && (nrays == 0 || (start && origin && dir && dist && cell && bary && crossings &&
                                                     status)),
                    "null arrays");
    std::vector<std::array<double, 3>> P((std::size_t)np);
    for (int64_t i = 0; i < np; ++i) P[(std::size_t)i] = {points[3 * i], points[3 * i + 1], points[3 * i + 2]};
    std::vector<std::array<int32_t, 4>> T((std::size_t)nt);
    for (int64_t c = 0; c < nt; ++c) T[(std::size_t)c] = {tets[4 * c], tets[4 * c + 1], tets[4 * c + 2], tets[4 * c + 3]};
    const nodals::TetRayTracer tracer(P, T);
    for (int64_t r = 0; r < nrays; ++r) {
      const std::array<double, 3> o = {origin[3 * r], origin[3 * r + 1], origin[3 * r + 2]};
      const std::array<double, 3> d = {dir[3 * r], dir[3 * r + 1], dir[3 * r + 2]};
      const std::vector<double> ds(dist + r * ndist, dist + (r + 1) * ndist);
      nodals::require(start[r] >= 0 && start[r] < nt, "ray start cell out of range");
      try {
        const auto hits = tracer.trace(start[r], o, d, ds, max_crossings, tol);
        for (int m = 0; m < ndist; ++m) {
          const auto& h = hits[(std::size_t)m];
          cell[r * ndist + m] = h.cell;
          crossings[r * ndist + m] = h.crossings;
          for (int j = 0; j < 4; ++j) bary[(r * ndist + m) * 4 + j] = h.l[j];
        }
        status[r] = 0;
        if (certified) certified[r] = ds.back();
      } catch (const nodals::Error&) {
        for (int m = 0; m < ndist; ++m) {
          cell[r * ndist + m] = -1;
          crossings[r * ndist + m] = 0;
          for (int j = 0; j < 4; ++j) bary[(r * ndist + m) * 4 + j] = 0.0;
        }
        status[r] = 1;
        if (certified) certified[r] = tracer.certifiedDistance(start[r], o, d, ds[0], max_crossings, tol);
      }
    }
  });
}

int nodals_wall_omega_target(int64_t n, const double* in, const int32_t* shared, double* out) {
  return guarded([&]() { nodals::omegaWallBatch(n, in, shared, out); });
}

int nodals_h8_sst_volume(nodals_h8* h, int32_t equation, const double* k_free, const double* k_fixed,
                         const double* omega_free, const double* omega_fixed, double k_floor, double omega_floor,
                         int32_t deviatoric, double* mat, double* rhs) {
  return guarded([&]() {
    auto& s = solver(h);
    const std::size_t nv = (std::size_t)s.setup().nv, nf = (std::size_t)s.setup().nFixed;
    nodals::require(k_free && omega_free && mat && rhs && (nf == 0 || (k_fixed && omega_fixed)), "null arrays");
    std::vector<double> m, r;
    s.sstVolume(equation, std::vector<double>(k_free, k_free + nv), std::vector<double>(k_fixed, k_fixed + nf),
                std::vector<double>(omega_free, omega_free + nv), std::vector<double>(omega_fixed, omega_fixed + nf),
                k_floor, omega_floor, deviatoric, m, r);
    std::memcpy(mat, m.data(), m.size() * sizeof(double));
    std::memcpy(rhs, r.data(), r.size() * sizeof(double));
  });
}

static nodals::GnnModel& gnnOf(nodals_gnn* g) {
  nodals::require(g != nullptr && g->g != nullptr, "null GNN handle");
  return *g->g;
}

int nodals_gnn_create_for_h8(nodals_h8* h, uint32_t seed, nodals_gnn** out) {
  return guarded([&]() {
    nodals::require(out != nullptr, "null output handle");
    *out = nullptr;
    auto& s = solver(h);
    nodals::require(s.hasDevice(), "H8 context was created without a device (host setup only)");
    auto g = std::make_unique<nodals_gnn>();
    g->g = std::make_unique<nodals::GnnModel>(s.cellGraph(), seed);
    *out = g.release();
  });
}

int nodals_gnn_destroy(nodals_gnn* g) {
  return guarded([&]() { delete g; });
}

int nodals_gnn_info(nodals_gnn* g, int64_t info[4]) {
  return guarded([&]() {
    auto& m = gnnOf(g);
    nodals::require(info != nullptr, "null info");
    info[0] = m.n();
    info[1] = m.paddedRows();
    info[2] = m.paramCount();
    info[3] = m.samples();
  });
}

int nodals_gnn_get_params(nodals_gnn* g, float* host, int64_t count) {
  return guarded([&]() {
    auto& m = gnnOf(g);
    nodals::require(host && count == m.paramCount(), "parameter count mismatch");
    m.getParams(host);
  });
}

int nodals_gnn_set_params(nodals_gnn* g, const float* host, int64_t count) {
  return guarded([&]() {
    auto& m = gnnOf(g);
    nodals::require(host && count == m.paramCount(), "parameter count mismatch");
    m.setParams(host);
  });
}

int nodals_gnn_reserve(nodals_gnn* g, int32_t capacity) {
  return guarded([&]() { gnnOf(g).reserveSamples(capacity); });
}

int nodals_gnn_sample(nodals_gnn* g, int32_t i, float* x, float* t) {
  return guarded([&]() {
    auto& m = gnnOf(g);
    nodals::require(i >= 0 && i < m.samples() && x && t, "sample index out of range");
    NODALS_HIP(hipStreamSynchronize(m.stream()));
    NODALS_HIP(hipMemcpy(x, m.sampleX(i), sizeof(float) * (std::size_t)m.n() * nodals::GnnModel::kIn,
                         hipMemcpyDeviceToHost));
    NODALS_HIP(hipMemcpy(t, m.sampleT(i), sizeof(float) * (std::size_t)m.n(), hipMemcpyDeviceToHost));
  });
}

namespace {
// Host rows (n x 16) into a padded device copy.
nodals::DBuf<float> gnnRows(nodals::GnnModel& m, const float* x, int width) {
  nodals::DBuf<float> d;
  d.alloc((std::size_t)m.paddedRows() * width);
  nodals::dev::fill<float>(m.stream(), (long long)m.paddedRows() * width, d.data(), 0.0f);
  NODALS_HIP(hipMemcpyAsync(d.data(), x, sizeof(float) * (std::size_t)m.n() * width, hipMemcpyHostToDevice,
                            m.stream()));
  return d;
}
}  // namespace

int nodals_gnn_forward(nodals_gnn* g, const float* x, float* y) {
  return guarded([&]() {
    auto& m = gnnOf(g);
    nodals::require(x && y, "null arrays");
    nodals::DBuf<float> X = gnnRows(m, x, nodals::GnnModel::kIn);
    m.forward(m.stream(), X.data(), nullptr);
    NODALS_HIP(hipMemcpyAsync(y, m.outputDevice(), sizeof(float) * (std::size_t)m.n(), hipMemcpyDeviceToHost,
                              m.stream()));
    NODALS_HIP(hipStreamSynchronize(m.stream()));
  });
}

int nodals_gnn_gradient(nodals_gnn* g, const float* x, const float* t, float* grad, double* loss) {
  return guarded([&]() {
    auto& m = gnnOf(g);
    nodals::require(x && t && grad, "null arrays");
    nodals::DBuf<float> X = gnnRows(m, x, nodals::GnnModel::kIn);
    nodals::DBuf<float> T = gnnRows(m, t, 1);
    nodals::DBuf<float> L;
    L.alloc(1);
    m.forward(m.stream(), X.data(), nullptr);
    m.backward(m.stream(), X.data(), T.data(), L.data());
    float lh = 0;
    NODALS_HIP(hipMemcpyAsync(grad, m.gradDevice(), sizeof(float) * (std::size_t)m.paramCount(),
                              hipMemcpyDeviceToHost, m.stream()));
    NODALS_HIP(hipMemcpyAsync(&lh, L.data(), sizeof(float), hipMemcpyDeviceToHost, m.stream()));
    NODALS_HIP(hipStreamSynchronize(m.stream()));
    if (loss) *loss = (double)lh / m.n();
  });
}

int nodals_gnn_train(nodals_gnn* g, int32_t epochs, float lr, uint32_t seed, double* last_loss) {
  return guarded([&]() {
    const double l = gnnOf(g).train(epochs, lr, seed, nullptr);
    if (last_loss) *last_loss = l;
  });
}

int nodals_gnn_evaluate(nodals_gnn* g, double* loss) {
  return guarded([&]() {
    const double l = gnnOf(g).evaluate();
    if (loss) *loss = l;
  });
}

int nodals_h8_turb_init(nodals_h8* h, int argc, const char* const* argv) {
  return guarded([&]() {
    auto& s = solver(h);
    nodals::require(argc == 0 || argv != nullptr, "null argv");
    s.turbInit(nodals::parseTurbArgs(argc, argv));
  });
}

int nodals_h8_turb_step(nodals_h8* h, nodals_turb_step_result* res) {
  return guarded([&]() {
    auto& s = solver(h);
    nodals::require(res != nullptr, "null result");
    const nodals::TurbStepResult r = s.turbStep();
    res->k_clips = r.kClips;
    res->omega_clips = r.omegaClips;
    res->wall_invalid = r.wallInvalid;
    res->reserved = 0;
    res->ms = r.ms;
  });
}

int nodals_h8_turb_couple(nodals_h8* h, int32_t on) {
  return guarded([&]() { solver(h).turbCouple(on != 0); });
}

int nodals_h8_wall_init(nodals_h8* h, int argc, const char* const* argv) {
  return guarded([&]() {
    auto& s = solver(h);
    nodals::require(argc == 0 || argv != nullptr, "null argv");
    s.wallInit(nodals::parseWallArgs(argc, argv));
  });
}

int nodals_h8_wall_assemble(nodals_h8* h) {
  return guarded([&]() {
    auto& s = solver(h);
    s.wallAssemble();
    s.sync();
  });
}

int nodals_h8_wall_momentum(nodals_h8* h) {
  return guarded([&]() {
    auto& s = solver(h);
    s.wallMomentum();
    s.sync();
  });
}

int nodals_h8_wall_schur(nodals_h8* h, const double* x, double* y) {
  return guarded([&]() {
    auto& s = solver(h);
    nodals::require(x != nullptr && y != nullptr, "null argument");
    nodals::require(s.hasWall(), "wall-function path not initialised (nodals_h8_wall_init)");
    const int nc = s.setup().nc;
    NODALS_HIP(hipMemcpyAsync(s.devScratchA(), x, sizeof(double) * nc, hipMemcpyHostToDevice, s.stream()));
    s.wallSchur(s.stream(), s.devScratchA(), s.devScratchB());
    NODALS_HIP(hipMemcpyAsync(y, s.devScratchB(), sizeof(double) * nc, hipMemcpyDeviceToHost, s.stream()));
    s.sync();
  });
}

int nodals_h8_wall_correct(nodals_h8* h, const double* dp) {
  return guarded([&]() {
    auto& s = solver(h);
    nodals::require(dp != nullptr, "null argument");
    nodals::require(s.hasWall(), "wall-function path not initialised (nodals_h8_wall_init)");
    const int nc = s.setup().nc;
    NODALS_HIP(hipMemcpyAsync(s.devScratchA(), dp, sizeof(double) * nc, hipMemcpyHostToDevice, s.stream()));
    s.wallCorrect(s.devScratchA());
    s.sync();
  });
}

int nodals_h8_attach_gnn(nodals_h8* h, nodals_gnn* g, int32_t warm_start, int32_t record_every) {
  return guarded([&]() {
    auto& s = solver(h);
    if (g && warm_start)
      nodals::require(s.options().pressureSolver == "fgmres", "the learned warm start needs --pressure-solver fgmres");
    s.attachGnn(g ? g->g.get() : nullptr, warm_start, record_every);
  });
}

}  // extern "C"
