// FP64 C++/OpenMP restatement of the H8 SIMPLE loop on the host: the CPU
// baseline that bench.py times beside the GPU.
//
// TEST INFRASTRUCTURE ONLY: built into oracle/_cpu/libnodals_cpu.so by
// oracle/Makefile and loaded only by tests/ and bench.py's cpu_baseline
// (through oracle/cpubind.py); the product never links it.
//
// It runs the same iterations as oracle/h8_oracle.py's `simple` (FGMRES
// pressure solve with the SA V-cycle and damped-Jacobi smoothing), which
// restates gpu/serial_cuda/h8_main_base.cu:2046-2119:
//   assembly  host_assemble_central_g4 (g4_host.hpp:317-320): diffusion +
//             Picard central convection C_ab = det sum_{m,j} ur[m][j]
//             T[a][m][b][j], Dirichlet columns to the right-hand side;
//   relax     host_finalize_relax_g4 (g4_host.hpp:422): row-L1 relaxation;
//   rhs       s + c + B^T p + delta u (g5e_fp32_kernels.cuh:97-104);
//   momentum  one forward multicolour Gauss-Seidel sweep (fgs1,
//             h8_main_base.cu:1678-1688) on the three components;
//   continuity fixedDiv + B u (h8_precomputed_b.cuh:140-163);
//   Schur     B diag(rAU) B^T (h2b_fine_csr.cuh:465-501);
//   pressure  right-preconditioned FGMRES(m) with the V-cycle
//             (h8_main_base.cu:544-564) and p += alpha_p dp.
// Setup arrays (cells, CSR patterns, colours, the SA hierarchy) are the
// product's host setup, as for the numpy oracle.  Loops are row- or
// cell-parallel with OpenMP; every scatter is turned into a gather, so each
// output is summed by one thread, but the
// summation order differs from the numpy oracle's, so the two agree to
// rounding, not bitwise; dot products are OpenMP reductions, so results
// also move at rounding level with the thread count.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" {

struct CpuCase {
  int32_t nc, nv, ncolors, pin;
  const int64_t* cell_ref;    // [nc][8]: >= 0 free DOF, < 0 -(fixed slot + 1)
  const uint8_t* cell_slot;   // [nc][64]
  const double* cell_det;     // [nc]
  const double* cell_invJ;    // [nc][3][3]
  const double* bcoef;        // [nc][8][3]
  const double* fixed_value;  // [nfixed][3]
  const int64_t* mom_row;     // [nv+1]
  const int32_t* mom_col;
  const double* mom_diff;
  const int32_t* color;       // [nv]
  const double* static_rhs;   // [3][nv]
  const double* fixed_div;    // [nc]
  const double* tcent;        // [8][8][8][3]
  // fine P0 Schur topology (numpy oracle's fine_topology)
  const int64_t* f_row;       // [nc+1]
  const int64_t* f_col;
  const int64_t* f_diag;      // [nc] positions of the diagonal
  const int64_t* inc_off;     // [nv+1] incidence of free DOFs
  const int64_t* inc_c;
  const int64_t* inc_a;
  const int64_t* contrib_off; // [nc+1] directed contributions per row
  const int64_t* slot;        // per contribution: position within the row
};

struct CpuLevel {             // transfer into level l and level l's operator
  int32_t nfine, ncoarse;
  const int64_t* p_row;       // P: nfine x ncoarse
  const int32_t* p_col;
  const double* p_val;
  const int64_t* a_row;       // A_l: ncoarse x ncoarse
  const int32_t* a_col;
  const double* a_val;
};

struct CpuHier {
  int32_t nlevels, nterm;
  const CpuLevel* levels;
  const double* tinv;         // [nterm][nterm]
};

}  // extern "C"

namespace {

struct Csr {
  int n = 0;
  const int64_t* row = nullptr;
  const int32_t* col32 = nullptr;
  const int64_t* col64 = nullptr;
  const double* val = nullptr;
  int64_t c(int64_t k) const { return col32 ? col32[k] : col64[k]; }
  void mv(const double* x, double* y) const {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      double s = 0;
      for (int64_t k = row[i]; k < row[i + 1]; ++k) s += val[k] * x[c(k)];
      y[i] = s;
    }
  }
  void resid(const double* b, const double* x, double* r) const {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      double s = b[i];
      for (int64_t k = row[i]; k < row[i + 1]; ++k) s -= val[k] * x[c(k)];
      r[i] = s;
    }
  }
};

double dot(int n, const double* a, const double* b) {
  double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

struct Solver {
  const CpuCase& C;
  const CpuHier& H;
  double jacOmega;
  int nnz;
  std::vector<double> aval, delta, diag, rau, conv, bt, b, cont, fval, fdiag, dp;
  std::vector<double> cell;                  // [nc][64] convection blocks
  std::vector<int64_t> termOff, term;        // per nonzero: cell*64 + a*8 + b
  std::vector<int64_t> fixOff, fixTerm;      // per row: (cell*64 + a*8 + b) * 2^20 + fixed slot
  std::vector<int64_t> diagPos;
  std::vector<std::vector<int>> colourRows;
  std::vector<std::vector<double>> lvlDiag, lx, lb, lr, lt;
  std::vector<std::vector<int64_t>> ptRow;   // P_l^T, built once
  std::vector<std::vector<int32_t>> ptCol;
  std::vector<std::vector<double>> ptVal;
  std::vector<double> fr, ft, gr, gw;        // V-cycle and FGMRES scratch
  std::vector<std::vector<double>> V, Z;
  Csr fine;

  Solver(const CpuCase& c, const CpuHier& h, double om) : C(c), H(h), jacOmega(om) {
    const int nv = C.nv, nc = C.nc;
    nnz = (int)C.mom_row[nv];
    aval.resize(nnz);
    delta.resize(nv);
    diag.resize(nv);
    rau.resize(nv);
    conv.resize(3 * (size_t)nv);
    bt.resize(3 * (size_t)nv);
    b.resize(3 * (size_t)nv);
    cont.resize(nc);
    cell.resize((size_t)nc * 64);
    fval.resize(C.f_row[nc]);
    fdiag.resize(nc);
    dp.resize(nc);
    // gather lists for the convection scatter
    std::vector<int64_t> cnt(nnz + 1, 0), fcnt(nv + 1, 0);
    for (int e = 0; e < nc; ++e)
      for (int a = 0; a < 8; ++a) {
        const int64_t ra = C.cell_ref[8 * e + a];
        if (ra < 0) continue;
        for (int bb = 0; bb < 8; ++bb) {
          const int64_t rb = C.cell_ref[8 * e + bb];
          if (rb >= 0) ++cnt[C.mom_row[ra] + C.cell_slot[64 * e + 8 * a + bb] + 1];
          else ++fcnt[ra + 1];
        }
      }
    for (int k = 0; k < nnz; ++k) cnt[k + 1] += cnt[k];
    for (int i = 0; i < nv; ++i) fcnt[i + 1] += fcnt[i];
    termOff = cnt;
    fixOff = fcnt;
    term.resize(cnt[nnz]);
    fixTerm.resize(2 * (size_t)fcnt[nv]);
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1), ffill(fcnt.begin(), fcnt.end() - 1);
    for (int e = 0; e < nc; ++e)
      for (int a = 0; a < 8; ++a) {
        const int64_t ra = C.cell_ref[8 * e + a];
        if (ra < 0) continue;
        for (int bb = 0; bb < 8; ++bb) {
          const int64_t rb = C.cell_ref[8 * e + bb];
          const int64_t t = 64ll * e + 8 * a + bb;
          if (rb >= 0) term[fill[C.mom_row[ra] + C.cell_slot[64 * e + 8 * a + bb]]++] = t;
          else {
            const int64_t q = ffill[ra]++;
            fixTerm[2 * q] = t;
            fixTerm[2 * q + 1] = -rb - 1;
          }
        }
      }
    diagPos.resize(nv);
    for (int i = 0; i < nv; ++i)
      for (int64_t k = C.mom_row[i]; k < C.mom_row[i + 1]; ++k)
        if (C.mom_col[k] == i) diagPos[i] = k;
    colourRows.resize(C.ncolors);
    for (int i = 0; i < nv; ++i) colourRows[C.color[i]].push_back(i);
    fine.n = nc;
    fine.row = C.f_row;
    fine.col64 = C.f_col;
    fine.val = fval.data();
    for (int l = 0; l < H.nlevels; ++l) {
      const CpuLevel& L = H.levels[l];
      std::vector<double> d(L.ncoarse, 1.0);
      for (int i = 0; i < L.ncoarse; ++i)
        for (int64_t k = L.a_row[i]; k < L.a_row[i + 1]; ++k)
          if (L.a_col[k] == i) d[i] = L.a_val[k];
      lvlDiag.push_back(d);
      lx.emplace_back(L.ncoarse);
      lb.emplace_back(L.ncoarse);
      lr.emplace_back(L.ncoarse);
      lt.emplace_back(L.ncoarse);
      std::vector<int64_t> tr(L.ncoarse + 1, 0);
      const int64_t pn = L.p_row[L.nfine];
      for (int64_t k = 0; k < pn; ++k) ++tr[L.p_col[k] + 1];
      for (int j = 0; j < L.ncoarse; ++j) tr[j + 1] += tr[j];
      std::vector<int32_t> tc(pn);
      std::vector<double> tv(pn);
      std::vector<int64_t> at(tr.begin(), tr.end() - 1);
      for (int i = 0; i < L.nfine; ++i)  // ascending fine rows within a coarse row
        for (int64_t k = L.p_row[i]; k < L.p_row[i + 1]; ++k) {
          const int64_t q = at[L.p_col[k]]++;
          tc[q] = i;
          tv[q] = L.p_val[k];
        }
      ptRow.push_back(std::move(tr));
      ptCol.push_back(std::move(tc));
      ptVal.push_back(std::move(tv));
    }
    fr.resize(nc);
    ft.resize(nc);
    gr.resize(nc);
    gw.resize(nc);
  }

  double uval(const double* U, int64_t r, int d) const {
    return r >= 0 ? U[(size_t)d * C.nv + r] : C.fixed_value[3 * (-r - 1) + d];
  }

  void assemble(const double* U, const double* p, double alphaU) {
    const int nv = C.nv, nc = C.nc;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < nc; ++e) {
      double ur[8][3];
      const double* I = C.cell_invJ + 9 * (size_t)e;
      for (int m = 0; m < 8; ++m) {
        const int64_t r = C.cell_ref[8 * e + m];
        for (int j = 0; j < 3; ++j) {
          double s = 0;
          for (int d = 0; d < 3; ++d) s += uval(U, r, d) * I[3 * j + d];
          ur[m][j] = s;
        }
      }
      for (int a = 0; a < 8; ++a)
        for (int bb = 0; bb < 8; ++bb) {
          double s = 0;
          for (int m = 0; m < 8; ++m)
            for (int j = 0; j < 3; ++j) s += ur[m][j] * C.tcent[((a * 8 + m) * 8 + bb) * 3 + j];
          cell[64 * (size_t)e + 8 * a + bb] = s * C.cell_det[e];
        }
    }
    const double fac = 1.0 / alphaU - 1.0;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nv; ++i) {
      double l1 = 0;
      for (int64_t k = C.mom_row[i]; k < C.mom_row[i + 1]; ++k) {
        double v = C.mom_diff[k];
        for (int64_t t = termOff[k]; t < termOff[k + 1]; ++t) v += cell[term[t]];
        aval[k] = v;
        l1 += std::fabs(v);
      }
      delta[i] = fac * l1;
      aval[diagPos[i]] += delta[i];
      diag[i] = aval[diagPos[i]];
      rau[i] = 1.0 / diag[i];
      double c3[3] = {0, 0, 0};
      for (int64_t q = fixOff[i]; q < fixOff[i + 1]; ++q) {
        const double cv = cell[fixTerm[2 * q]];
        for (int d = 0; d < 3; ++d) c3[d] -= cv * C.fixed_value[3 * fixTerm[2 * q + 1] + d];
      }
      double t3[3] = {0, 0, 0};
      for (int64_t k = C.inc_off[i]; k < C.inc_off[i + 1]; ++k) {
        const int64_t e = C.inc_c[k], a = C.inc_a[k];
        for (int d = 0; d < 3; ++d) t3[d] += C.bcoef[24 * e + 3 * a + d] * p[e];
      }
      for (int d = 0; d < 3; ++d)
        b[(size_t)d * nv + i] = C.static_rhs[(size_t)d * nv + i] + c3[d] + t3[d] + delta[i] * U[(size_t)d * nv + i];
    }
  }

  void sweep(double* U, double omega) {
    for (const auto& rows : colourRows) {
      const int n = (int)rows.size();
#pragma omp parallel for schedule(static)
      for (int q = 0; q < n; ++q) {
        const int i = rows[q];
        double s[3] = {0, 0, 0};
        for (int64_t k = C.mom_row[i]; k < C.mom_row[i + 1]; ++k) {
          if (k == diagPos[i]) continue;
          const int j = C.mom_col[k];
          for (int d = 0; d < 3; ++d) s[d] += aval[k] * U[(size_t)d * C.nv + j];
        }
        for (int d = 0; d < 3; ++d) {
          double& u = U[(size_t)d * C.nv + i];
          const double z = (b[(size_t)d * C.nv + i] - s[d]) / diag[i];
          u = (1.0 - omega) * u + omega * z;
        }
      }
    }
  }

  double continuity(const double* U) {
    const int nc = C.nc;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < nc; ++e) {
      double s = C.fixed_div[e];
      for (int a = 0; a < 8; ++a) {
        const int64_t r = C.cell_ref[8 * e + a];
        if (r < 0) continue;
        for (int d = 0; d < 3; ++d) s += C.bcoef[24 * e + 3 * a + d] * U[(size_t)d * C.nv + r];
      }
      cont[e] = s;
    }
    return std::sqrt(dot(nc, cont.data(), cont.data()));
  }

  void refresh() {
    const int nc = C.nc;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < nc; ++e) {
      double* v = fval.data() + C.f_row[e];
      std::fill(v, v + (C.f_row[e + 1] - C.f_row[e]), 0.0);
      int64_t t = C.contrib_off[e];
      for (int a = 0; a < 8; ++a) {
        const int64_t g = C.cell_ref[8 * e + a];
        if (g < 0) continue;
        const double* Ba = C.bcoef + 24 * (size_t)e + 3 * a;
        for (int64_t k = C.inc_off[g]; k < C.inc_off[g + 1]; ++k, ++t) {
          const double* Bb = C.bcoef + 24 * C.inc_c[k] + 3 * C.inc_a[k];
          v[C.slot[t]] += rau[g] * (Ba[0] * Bb[0] + Ba[1] * Bb[1] + Ba[2] * Bb[2]);
        }
      }
      if (C.pin >= 0)
        for (int64_t k = C.f_row[e]; k < C.f_row[e + 1]; ++k)
          if (C.f_col[k] != e && (e == C.pin || C.f_col[k] == C.pin)) fval[k] = 0.0;
    }
    for (int e = 0; e < nc; ++e) fdiag[e] = fval[C.f_diag[e]];
  }

  // damped Jacobi from zero: x = w b / d
  void jacobi(int n, const double* d, const double* bb, double* x) const {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) x[i] = jacOmega * bb[i] / d[i];
  }

  // out = P_l^T r through the transposed copy: one thread per coarse row
  void restrictTo(int l, const double* r, double* out) const {
    const int n = H.levels[l].ncoarse;
    const int64_t* tr = ptRow[l].data();
    const int32_t* tc = ptCol[l].data();
    const double* tv = ptVal[l].data();
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j) {
      double s = 0;
      for (int64_t k = tr[j]; k < tr[j + 1]; ++k) s += tv[k] * r[tc[k]];
      out[j] = s;
    }
  }
  void prolongAdd(const CpuLevel& L, const double* xc, double* x) const {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < L.nfine; ++i) {
      double s = 0;
      for (int64_t k = L.p_row[i]; k < L.p_row[i + 1]; ++k) s += L.p_val[k] * xc[L.p_col[k]];
      x[i] += s;
    }
  }

  void level(int l, const double* bb, double* x) {
    const CpuLevel& L = H.levels[l];
    const int n = L.ncoarse;
    if (l == H.nlevels - 1) {
#pragma omp parallel for schedule(static)
      for (int i = 0; i < n; ++i) {
        double s = 0;
        for (int j = 0; j < n; ++j) s += H.tinv[(size_t)i * n + j] * bb[j];
        x[i] = s;
      }
      return;
    }
    Csr A{n, L.a_row, L.a_col, nullptr, L.a_val};
    std::vector<double>& r = lr[l];
    std::vector<double>& t = lt[l];
    jacobi(n, lvlDiag[l].data(), bb, x);
    A.resid(bb, x, r.data());
    const CpuLevel& N = H.levels[l + 1];
    std::vector<double>& cb = lb[l + 1];
    std::vector<double>& cx = lx[l + 1];
    restrictTo(l + 1, r.data(), cb.data());
    level(l + 1, cb.data(), cx.data());
    prolongAdd(N, cx.data(), x);
    A.resid(bb, x, r.data());
    jacobi(n, lvlDiag[l].data(), r.data(), t.data());
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) x[i] += t[i];
  }

  void precond(const double* bb, double* z) {
    const int n = C.nc;
    double* r = fr.data();
    double* t = ft.data();
    jacobi(n, fdiag.data(), bb, z);
    fine.resid(bb, z, r);
    restrictTo(0, r, lb[0].data());
    level(0, lb[0].data(), lx[0].data());
    prolongAdd(H.levels[0], lx[0].data(), z);
    fine.resid(bb, z, r);
    jacobi(n, fdiag.data(), r, t);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) z[i] += t[i];
  }

  // right-preconditioned FGMRES(m), classical Gram-Schmidt twice, Givens
  // rotations on the Hessenberg column (textbook FGMRES, Saad 1993)
  int fgmres(const double* rhs, double* x, double rtol, double atol, int maxit, int m, double* relOut) {
    const int n = C.nc;
    double* r = gr.data();
    double* w = gw.data();
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      r[i] = rhs[i];
      x[i] = 0.0;
    }
    double beta = std::sqrt(dot(n, r, r));
    const double r0 = beta;
    if (beta <= atol) {
      *relOut = 0.0;
      return 0;
    }
    const double target = std::max(atol, rtol * r0);
    if ((int)V.size() < m + 1) {
      V.assign(m + 1, std::vector<double>(n));
      Z.assign(m, std::vector<double>(n));
    }
    std::vector<double> Hm((m + 1) * m), cs(m), sn(m), g(m + 1), h(m + 1);
    int its = 0;
    for (;;) {
      {
        double* v0 = V[0].data();
        const double ib = 1.0 / beta;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) v0[i] = r[i] * ib;
      }
      std::fill(g.begin(), g.end(), 0.0);
      g[0] = beta;
      int k = 0;
      bool done = false;
      for (int j = 0; j < m && its < maxit; ++j) {
        precond(V[j].data(), Z[j].data());
        fine.mv(Z[j].data(), w);
        for (int i = 0; i <= j; ++i) Hm[i * m + j] = 0.0;
        for (int pass = 0; pass < 2; ++pass) {
          for (int i = 0; i <= j; ++i) h[i] = dot(n, V[i].data(), w);
#pragma omp parallel for schedule(static)
          for (int q = 0; q < n; ++q) {
            double s = w[q];
            for (int i = 0; i <= j; ++i) s -= h[i] * V[i][q];
            w[q] = s;
          }
          for (int i = 0; i <= j; ++i) Hm[i * m + j] += h[i];
        }
        const double hn = std::sqrt(dot(n, w, w));
        ++its;
        k = j + 1;
        for (int i = 0; i < j; ++i) {
          const double a0 = Hm[i * m + j], a1 = Hm[(i + 1) * m + j];
          Hm[i * m + j] = cs[i] * a0 + sn[i] * a1;
          Hm[(i + 1) * m + j] = -sn[i] * a0 + cs[i] * a1;
        }
        const double a = Hm[j * m + j];
        const double den = std::sqrt(a * a + hn * hn);
        cs[j] = den > 0 ? a / den : 1.0;
        sn[j] = den > 0 ? hn / den : 0.0;
        Hm[j * m + j] = den;
        g[j + 1] = -sn[j] * g[j];
        g[j] = cs[j] * g[j];
        if (std::fabs(g[j + 1]) <= target || hn == 0.0) {
          done = true;
          break;
        }
        double* vn = V[j + 1].data();
        const double ih = 1.0 / hn;
#pragma omp parallel for schedule(static)
        for (int q = 0; q < n; ++q) vn[q] = w[q] * ih;
      }
      std::vector<double> y(k);
      for (int i = k - 1; i >= 0; --i) {
        double s = g[i];
        for (int q = i + 1; q < k; ++q) s -= Hm[i * m + q] * y[q];
        y[i] = s / Hm[i * m + i];
      }
#pragma omp parallel for schedule(static)
      for (int q = 0; q < n; ++q) {
        double s = x[q];
        for (int i = 0; i < k; ++i) s += y[i] * Z[i][q];
        x[q] = s;
      }
      if (done && std::fabs(g[k]) <= target) {
        *relOut = std::fabs(g[k]) / r0;
        return its;
      }
      fine.resid(rhs, x, r);
      beta = std::sqrt(dot(n, r, r));
      *relOut = beta / r0;
      if (beta <= target || its >= maxit) return its;
    }
  }
};

}  // namespace

extern "C" {

// Builds the gather lists, the transposed prolongations and the scratch of
// one case; the arrays behind `c` and `h` must outlive the handle.
void* cpu_h8_open(const CpuCase* c, const CpuHier* h, double jacobi_omega) {
  return new Solver(*c, *h, jacobi_omega);
}

void cpu_h8_close(void* s) { delete static_cast<Solver*>(s); }

// `iters` SIMPLE iterations from U [3][nv] and p [nc] (in place); stats
// [iters][3] = continuity norm before the pressure solve, pressure
// iterations, pressure relative residual.  threads <= 0: OpenMP default.
int cpu_h8_simple(void* handle, int iters, double alpha_u, double alpha_p, double p_rtol, double p_atol, int p_max,
                  int restart, int threads, double* U, double* p, double* stats) {
  if (threads > 0) omp_set_num_threads(threads);
  Solver& S = *static_cast<Solver*>(handle);
  const CpuCase* c = &S.C;
  const int nc = c->nc;
  std::vector<double> rhs(nc);
  for (int it = 0; it < iters; ++it) {
    S.assemble(U, p, alpha_u);
    S.sweep(U, 1.0);
    const double cn = S.continuity(U);
    S.refresh();
    for (int e = 0; e < nc; ++e) rhs[e] = -S.cont[e];
    if (c->pin >= 0) rhs[c->pin] = 0.0;
    double rel = 0;
    const int pits = S.fgmres(rhs.data(), S.dp.data(), p_rtol, p_atol, p_max, restart, &rel);
    for (int e = 0; e < nc; ++e) p[e] += alpha_p * S.dp[e];
    stats[3 * it] = cn;
    stats[3 * it + 1] = pits;
    stats[3 * it + 2] = rel;
  }
  return 0;
}

int cpu_h8_threads() { return omp_get_max_threads(); }

}  // extern "C"
