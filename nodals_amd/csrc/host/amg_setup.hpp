// Smoothed-aggregation AMG hierarchy for the P0 pressure Schur operator and
// for explicit CSR operators.
//
// Reference: gpu/serial_cuda/g2_host.hpp:79-139 (face graph, greedy BFS
// aggregation, dense Cholesky terminal inverse) and g5_sa_host.hpp
// (power iteration, smoothed prolongator with normalise/prune, Galerkin
// coarse operators).  Aggregates and all CSR patterns are bit-exact with the
// reference; the arithmetic follows the reference operation order (compile
// with -ffp-contract=off), which keeps prune decisions identical.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "h8_setup.hpp"

namespace nodals {

struct AggResult {
  std::vector<std::int32_t> id;
  int nagg = 0;
};

struct Transfer {
  int nFine = 0, nCoarse = 0;
  std::vector<std::int64_t> row;
  std::vector<std::int32_t> col;
  std::vector<double> val;
};

struct AmgOptions {
  int target = 16, minSize = 6, softMax = 18;
  int terminal = 1000;
  int maxPnnz = 8;
  int powerIts = 16;
  double safety = 1.5;
  double lowFraction = 0.05;
  double damping = 4.0 / 3.0;
};

struct AmgHierarchy {
  std::vector<AggResult> agg;
  std::vector<Transfer> P;     // P[0]: fine -> A[0]; P[l]: A[l-1] -> A[l]
  std::vector<CsrD> A;         // explicit coarse operators; A.back() is terminal
  std::vector<double> fineDiag;
  double fineLambda = 0.0;
  std::vector<double> levelLambda;  // per explicit level, 0 for the terminal
  std::vector<double> terminalInv;  // dense row-major inverse of A.back()
  int terminalN = 0;
  // CF/PMIS only: the C/F state (1 = C, 2 = F) and coarse id of every split
  // made, in construction order (an aggressive first stage adds two).
  std::vector<std::vector<std::uint8_t>> cfState;
  std::vector<std::vector<std::int32_t>> cfCoarseId;
  std::vector<int> cfRounds;
};

struct CfOptions {
  double theta = 0.25;
  int pmax = 4;
  std::string interp = "direct";  // direct | exti
  bool aggressiveFirst = false;
  int terminal = 1000;
  int powerIts = 0;
  double safety = 1.5;
};

// PMIS C/F split of A under the classical negative-coupling strength test.
struct CfSplit {
  int n = 0, nC = 0, rounds = 0;
  std::vector<std::uint8_t> state;  // 1 = C, 2 = F
  std::vector<std::int32_t> coarseId;
  std::vector<std::vector<std::int32_t>> strong, graph;
};
std::uint64_t cfHash64(std::uint64_t x);
CfSplit pmisSplit(const CsrD& A, double theta);
Transfer cfTransfer(const CsrD& A, const CfSplit& S, int pmax, const std::string& interp);
Transfer composeTransfers(const Transfer& P01, const Transfer& P12);
CsrD galerkinProduct(const CsrD& A, const Transfer& P);

// Exact assembled fine Schur CSR B diag(rau) B^T in the reference summation
// order (g8_cf_hierarchy_host.hpp:116-166).
CsrD exactFineCsr(const H8Setup& S, const FineTopo& F, const std::vector<double>& rau);

// Zero the off-diagonal entries of row and column `cell` (pressure pin of a
// closed domain); the pattern and the diagonal are kept.
void pinCsr(CsrD& A, int cell);

// CF/PMIS hierarchy of the fine Schur operator (g8_cf_hierarchy_host.hpp:
// 1584-1783); of an explicit CSR operator with buildCfHierarchyCsr.
AmgHierarchy buildCfHierarchy(const H8Setup& S, const FineTopo& F, const std::vector<double>& rau,
                              const CfOptions& o);
AmgHierarchy buildCfHierarchyCsr(const CsrD& A0, const CfOptions& o);

using Graph = std::vector<std::vector<std::int32_t>>;

Graph meshFaceGraph(const TetMesh& M);
Graph csrGraph(const CsrD& A);
AggResult aggregateGraph(const Graph& g, int target, int minSize, int softMax);

// SA hierarchy of B diag(rAU) B^T given by the H8 setup and the fine
// topology (g5_sa_host.hpp:1545-1580).
AmgHierarchy buildSaHierarchy(const TetMesh& M, const H8Setup& S, const FineTopo& F,
                              const std::vector<double>& rau, const AmgOptions& o);

// SA hierarchy of an explicit SPD CSR operator: the coarse-level steps of the
// same construction applied from level 0.
AmgHierarchy buildSaHierarchyCsr(const CsrD& A0, const AmgOptions& o);

double powerCsr(const CsrD& A, int its, double safety);
std::vector<double> denseInverseCholesky(const CsrD& A);

}  // namespace nodals
