/*
 * nodals_hip.h -- C ABI of the MI355X NodalS SIMPLE pipeline (libnodals_hip.so).
 *
 * Plain C types only: opaque handles, host pointers, sizes, int status.
 * Every entry point returns 0 on success or a nonzero status:
 *   1 invalid argument / unsupported option / malformed input
 *   2 HIP runtime error            3 non-finite value
 *   4 Krylov breakdown             5 not converged
 *   6 internal error
 * and leaves a message for nodals_last_error().  No C++ exception crosses the
 * ABI.  A context owns all of its device memory and its HIP stream; calls on
 * one context are not reentrant.
 *
 * Reference interfaces each entry point replaces are cited as
 * path:line in the reference tree (JD63021/NodalS).
 */
#ifndef NODALS_HIP_H
#define NODALS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nodals_h8 nodals_h8;   /* H8 laminar SIMPLE solver context  */
typedef struct nodals_sys nodals_sys; /* explicit SPD system + AMG + Krylov */

/* G4PCGResult (gpu/serial_cuda/h8_main_base.cu:568) plus the initial norm. */
typedef struct {
  int32_t its;
  int32_t ok;
  double rel;
  double r0;
} nodals_solve_result;

/* Summary of one SIMPLE run: the fields of the NODALS_GPU_H8_FULLCONV /
 * NODALS_GPU_H8_FIXED10 markers (gpu/serial_cuda/h8_main_base.cu:2190-2193). */
typedef struct {
  int32_t outer;
  int32_t converged;
  int32_t pass;
  int32_t reserved;
  double final_rel_cont;
  double avg_pressure_its;
  double pressure_drop;
  double exact_pressure_drop;
  double loop_ms;
  double assembly_ms, momentum_ms, continuity_ms, pressure_ms, pupdate_ms;
  double setup_ms;
  /* Extensions (no reference counterpart): total pressure Krylov
   * iterations, learned warm starts tried, dropped (fallback), samples
   * recorded; device time of the k-omega steps of a coupled run. */
  int64_t pressure_its;
  int32_t warm_starts, fallbacks, recorded, reserved2;
  double turb_ms;
  /* Wall-function path: the last iteration's relative changes of u, k and
   * omega and its continuity residual after the correction. */
  double final_du, final_dk, final_dw, final_cont;
} nodals_h8_run_result;

/* Last error message of the calling thread ("" if none). */
const char* nodals_last_error(void);
const char* nodals_version(void);

/* ---------------------------------------------------------------- H8 ----
 * Create from the H8 command line: the same options and validation as the
 * reference binary (gpu/serial_cuda/h8_main_base.cu:1787-1911), argv[0]
 * ignored.  Runs mesh input (load_foam_tet_mesh, foam_mesh_g2.hpp:55), host
 * setup (build_g4_setup, g4_host.hpp:289; build_h2_fine_csr_host,
 * h2b_fine_csr.cuh:445; build_sa_hierarchy, g5_sa_host.hpp:1545) and the
 * device upload (upload_all, h8_main_base.cu:761-788). */
int nodals_h8_create(int argc, const char* const* argv, nodals_h8** out);
/* Host setup only (no HIP call): the same setup as nodals_h8_create without
 * the device upload.  Only nodals_h8_info and the setup/topology fields of
 * nodals_h8_get are valid on such a handle; other calls return status 1. */
int nodals_h8_create_host(int argc, const char* const* argv, nodals_h8** out);
void nodals_h8_destroy(nodals_h8* h);

/* info[0..7] = cells, free velocity DOFs, momentum nnz, colors,
 * fine pressure nnz, AMG explicit levels, terminal rows, fixed slots. */
int nodals_h8_info(nodals_h8* h, int64_t info[8]);

/* assemble_live (h8_main_base.cu:1137-1155) fused with bt_state + the
 * momentum RHS kernel (h8_main_base.cu:2061-2062): convection on the current
 * velocity, row-L1 relaxation, rAU, b = s + c + B^T p + delta u. */
int nodals_h8_assemble(nodals_h8* h, double alpha_u, double rau_scale);

/* fixed_momentum_work (h8_main_base.cu:1700-1715): mode "fgs1", "altgs1"
 * or "sgs1"; multicolour Gauss-Seidel on all three components. */
int nodals_h8_momentum_work(nodals_h8* h, const char* mode, double omega, int32_t outer);

/* momentum_initial_rel_audit (h8_main_base.cu:1716-1726): ||b-Au||/||b||. */
int nodals_h8_momentum_audit(nodals_h8* h, double rel[3]);

/* Continuity residual fixedDiv + B u (h8_continuity_kernel,
 * h8_precomputed_b.cuh:140); stores its negation as the pressure RHS and
 * returns the FP64 2-norm (h8_main_base.cu:2070-2081). */
int nodals_h8_continuity(nodals_h8* h, double* norm);

/* Fine P0 Schur numeric refresh B diag(rAU) B^T with the current rAU
 * (H2FineCSRDevice::refresh, h8_main_base.cu:322-328). */
int nodals_h8_pressure_refresh(nodals_h8* h);

/* pressure_pcg / pressure_richardson / pressure_cheb_poly
 * (h8_main_base.cu:569-726) and FGMRES: solver = "pcg", "richardson",
 * "cheb" or "fgmres", preconditioned by the AMG V-cycle. */
int nodals_h8_pressure_solve(nodals_h8* h, const char* solver, double rtol, double atol, int32_t maxit,
                             nodals_solve_result* res);

/* p += alpha_p dp (h8_main_base.cu:2102). */
int nodals_h8_pressure_update(nodals_h8* h, double alpha_p);

/* hxt5e::momentum_solve (turbulence/sst_gpu_v2/source/
 * puretet_profile5_eq5_hxt5b/hxt5e_fgmres.cuh:480-527) in its fixed-sweep
 * modes on the assembled momentum system: "m3-fgs" / "m3-sgs" (the
 * wall-function path's 3x3 block sweeps, nodals_h8_wall_assemble first) or
 * "fgs1" / "sgs1" / "altgs1" (the H8 component sweeps, nodals_h8_assemble
 * first), `sweeps` times with damping omega.  res: its = sweeps, r0 / rel =
 * ||b - A u|| / ||b|| before / after (largest component for H8). */
int nodals_h8_momentum_solve(nodals_h8* h, const char* mode, int32_t sweeps, double omega, nodals_solve_result* res);
/* hxt4c::pcg_amg (hybrid_amg.cuh:642-698): PCG preconditioned by the AMG
 * V-cycle on the live pressure operator (the fine Schur matrix, or the
 * wall path's constrained Schur operator) for a host right-hand side rhs
 * [cells]; the solution x [cells] is written to host memory.  The AMG fine
 * level is the one of the last refresh.  No snapshot POD. */
int nodals_h8_pcg_amg(nodals_h8* h, const double* rhs, double* x, double rtol, double atol, int32_t maxit,
                      nodals_solve_result* res);
/* hxt5c::scalar_solve (modules/transport/scalar_linear_solver.inc:3-136)
 * for equation 0 (k) or 1 (omega) at the current velocity: the equation's
 * masked-P1 assembly and its fixed multicolour sweeps (fgs-fixed), as in
 * one nodals_h8_turb_step.  res: its = sweeps, r0 / rel = relative
 * residual before / after the sweeps. */
int nodals_h8_scalar_solve(nodals_h8* h, int32_t equation, nodals_solve_result* res);

/* The SIMPLE loop of h8_main_base.cu:2046-2119 with its marker lines. */
int nodals_h8_run(nodals_h8* h, nodals_h8_run_result* res);
/* Wall-function path: `count` SIMPLE iterations numbered first_it ..
 * first_it + count - 1, continuing from the current u, p, k and omega, so a
 * run can be cut into windows around checkpoints (the SST application's
 * Stage31 checkpoint cadence, sst_application.inc, STAGE31_CHECKPOINT_EVERY);
 * res->outer is the last iteration's number. */
int nodals_h8_run_window(nodals_h8* h, int32_t first_it, int32_t count, nodals_h8_run_result* res);
/* ASCII VTK unstructured grid (.vtu) of the current fields: U, k, omega and
 * the wall distance at the vertices, p per cell (the SST application's
 * flow_sst.vtu, --vtu; SST_G2_VTU_EVERY). */
int nodals_h8_write_vtu(nodals_h8* h, const char* path);

/* Host copies in reference numbering.  Names (double unless noted):
 *   u0 u1 u2 (free DOFs), b0 b1 b2, rau delta diag (free DOFs),
 *   p cont dp fine_diag fine_l1 (cells), mom_val (momentum nnz, reference
 *   CSR order), fine_val (fine nnz), and int32 topology: mom_row (int64),
 *   mom_col, color, fine_row, fine_col, agg0.., cell_ref (8 per cell). */
int nodals_h8_get(nodals_h8* h, const char* name, void* host, int64_t count);
int nodals_h8_set(nodals_h8* h, const char* name, const void* host, int64_t count);

/* Fine Schur SpMV y = A x on device-resident copies of host vectors, and a
 * timed loop of `reps` launches of a named hot kernel ("fine_spmv",
 * "refresh", "mom_gs", "conv_local", "mom_rows", "continuity"); ms_avg gets
 * the mean launch time from HIP events on the context stream and
 * bytes the algorithmic bytes of one launch. */
int nodals_h8_fine_spmv(nodals_h8* h, const double* x, double* y);
/* One AMG preconditioner application z = M^-1 r (G4AMG::apply,
 * h8_main_base.cu:557-564) on the current fine operator. */
int nodals_h8_precond(nodals_h8* h, const double* r, double* z);
int nodals_h8_time_kernel(nodals_h8* h, const char* kernel, int32_t reps, double* ms_avg, double* bytes);
/* Launch probe inside the solver's own calls: enable=1 starts bracketing
 * every fine Schur SpMV launch (the Krylov operator of
 * h8_main_base.cu:567-726), enable=2 every momentum Gauss-Seidel colour
 * launch (g4_mcgs_color_kernel, g5e_fp32_kernels.cuh:106-116; the wall
 * path's M3 block colour, stage22d4d_coupled_matrixfree.inc:843-889) with
 * HIP events on the context stream; enable=0 stops and returns the launch
 * count and the summed launch durations (ms).  No
 * reference counterpart (measurement hook, like the reference's
 * G5StageEvents, h8_main_base.cu:2048-2105). */
int nodals_h8_probe(nodals_h8* h, int32_t enable, int64_t* launches, double* total_ms);

/* ---------------------------------------------------------------- SYS ---
 * Explicit SPD system (FP64 CSR, int64 row pointers) with an SA hierarchy
 * built by the coarse-level steps of build_sa_hierarchy
 * (g5_sa_host.hpp:1570-1577) applied from level 0.  options: space separated
 * key=value: precision=fp64|fp32|bf16 smoother=jacobi|l1jacobi|cheb2|mcgs
 * sweeps=N omega=W order=symmetric|forward|backward terminal=N power_its=N. */
int nodals_sys_create(int32_t n, const int64_t* row, const int32_t* col, const double* val, const char* options,
                      nodals_sys** out);
/* Host hierarchy only (no HIP call); only nodals_sys_info/nodals_sys_get. */
int nodals_sys_create_host(int32_t n, const int64_t* row, const int32_t* col, const double* val,
                           const char* options, nodals_sys** out);
void nodals_sys_destroy(nodals_sys* s);
/* info[0..3] = rows, nnz, AMG explicit levels, terminal rows;
 * info[4+l] = rows of explicit level l (l < 12). */
int nodals_sys_info(nodals_sys* s, int64_t info[16]);
/* Host hierarchy arrays: agg<l>, P<l>_row|col|val, A<l>_row|col|val,
 * terminal_inv, level_lambda, fine_lambda. */
int nodals_sys_get(nodals_sys* s, const char* name, void* host, int64_t count);
int nodals_sys_spmv(nodals_sys* s, const double* x, double* y);
/* One preconditioner application z = M^-1 r. */
int nodals_sys_precond(nodals_sys* s, const double* r, double* z);
/* solver = "fgmres" | "pcg"; x0 = 0. */
int nodals_sys_solve(nodals_sys* s, const double* b, double* x, const char* solver, double rtol, int32_t restart,
                     int32_t maxit, nodals_solve_result* res);
/* Solve with the RHS already resident, timed with HIP events around the
 * whole Krylov solve on the context stream (ms). */
int nodals_sys_solve_timed(nodals_sys* s, const char* solver, double rtol, int32_t restart, int32_t maxit,
                           int32_t reps, double* ms_avg, nodals_solve_result* res);
int nodals_sys_set_rhs(nodals_sys* s, const double* b);
/* Mean launch time of `reps` SpMV launches on level l (-1: finest FP64
 * operator; 0..: explicit AMG levels in the hierarchy precision), with the
 * algorithmic bytes of one launch. */
int nodals_sys_time_spmv(nodals_sys* s, int32_t level, int32_t reps, double* ms_avg, double* bytes);

/* k-omega SST volume terms of every cell of an H8 context's mesh (the
 * reference kernel sst_tet_volume, turbulence/sst_gpu_v2/source/
 * puretet_profile5_eq5_hxt5b/modules/transport/tet_volume_kernel.inc:1-30,
 * with volume_coefficients, modules/transport/portable/transport_core.hpp:
 * 14-25): Keast 24-point rule, SST-2003 local state at each point with the
 * pipe wall distance, from the context's current velocity and the given
 * k, omega (k_free/omega_free in reference free-DOF order, k_fixed/
 * omega_fixed per fixed slot).  equation 0: k, 1: omega.  mat[cells][64]
 * (local a, b), rhs[cells][8]. */
int nodals_h8_sst_volume(nodals_h8* h, int32_t equation, const double* k_free, const double* k_fixed,
                         const double* omega_free, const double* omega_fixed, double k_floor, double omega_floor,
                         int32_t deviatoric, double* mat, double* rhs);

/* k-omega SST transport step of an H8 context (masked-P1 state space: one
 * scalar per mesh entity, vertices first then faces, face-bubble values
 * zero).  Replaces step_turbulence (turbulence/sst_gpu_v2/source/
 * puretet_profile5_eq5_hxt5b/modules/coupling/g2_step_initialization.inc,
 * g2_step_convection.inc, g2_step_k.inc, g2_step_omega.inc) for tetrahedral
 * pipes: convection at the context's current velocity, then k and omega in
 * turn: volume terms, inlet flux (tet_inlet_flux.inc), omega wall penalty at
 * the third of five wall-normal samples with the OF-auto target
 * (puretet_of_auto_point3.cuh), row-L1 relaxation (scalar_relaxation.inc:
 * 8-33), fixed coloured Gauss-Seidel sweeps (scalar_linear_solver.inc:
 * 33-67 "fgs-fixed") and the floor.
 * nodals_h8_turb_init takes the reference's SST flags (--sst-intensity,
 * --sst-length-ratio, --sst-alpha-k, --sst-alpha-omega, --sst-k-fgs-sweeps,
 * --sst-omega-fgs-sweeps, --sst-gs-omega, --sst-k-floor, --sst-omega-floor,
 * --sst-deviatoric-production, --sst-omega-wall-mode of_auto|none,
 * --sst-wall-penalty-gamma, --sst-profile-fit-y1..y5,
 * --sst-wall-sample-scale, --sst-of-kappa, --sst-of-E, --sst-of-beta1,
 * --sst-of-re-blend, --sst-of-production-limit; argv holds flags only) and
 * sets k, omega to their inlet values.  k and omega are read and written
 * with nodals_h8_get / nodals_h8_set ("k", "omega"). */
typedef struct nodals_turb_step_result {
  int32_t k_clips, omega_clips; /* vertex values raised to the floor */
  int32_t wall_invalid;         /* wall points without a valid fit or target */
  int32_t reserved;
  double ms;                    /* device time of the step */
} nodals_turb_step_result;
int nodals_h8_turb_init(nodals_h8* h, int argc, const char* const* argv);
int nodals_h8_turb_step(nodals_h8* h, nodals_turb_step_result* res);
/* Coupled k-omega SST mode of nodals_h8_run (on = 1; needs
 * nodals_h8_turb_init): the momentum assembly adds the eddy-viscosity
 * diffusion (nu_t of the SST local state at each Keast point, the isotropic
 * part of sst_mom_tet, turbulence/sst_gpu_v2/source/
 * puretet_profile5_eq5_hxt5b/hxt5d_sst_g2.cuh:1220) and every SIMPLE
 * iteration ends with one k-omega step (the G2 order: momentum, pressure,
 * turbulence, sst_application.inc:2818-2978). */
int nodals_h8_turb_couple(nodals_h8* h, int32_t on);

/* Wall-function momentum path (the reference's spiral-case momentum:
 * turbulence/sst_gpu_v2/source/puretet_profile5_eq5_hxt5b/modules/coupling/
 * simple_momentum_assembly.inc, simple_pressure_predictor.inc,
 * simple_pressure_corrector.inc).  Needs a duct context created with
 * --duct 1 --wall-model spalding and nodals_h8_turb_init.  argv (flags
 * only): --gamma G (Nitsche normal penalty, SST22D4_GAMMA_N),
 * --wall-active-points N (SST24B_ACTIVE_POINTS), --wall-zero-limit 0|1,
 * --convection-form skew_symmetric|advective, --stress-form
 * deviatoric|vector_laplacian, --m3-mode fgs|sgs, --m3-sweeps N,
 * --m3-omega W, --initial-velocity potential|plug (default potential: the
 * potential-flow start of modules/initialization/
 * stage27_potential_host.hpp:44-167), --potential-tol T (its Laplace
 * solve, default 1e-10).  After it nodals_h8_run runs the wall-function SIMPLE loop
 * and nodals_h8_pressure_solve uses the constrained Schur operator.
 * nodals_h8_get names: "wall_K" [nnz][9] (reference entry order),
 * "wall_rhs", "wall_rau3", "wall_r3", "wall_delta3" [nv][3], "wall_fit"
 * [nf][3] (beta, u_tau, y+), "wall_shear" [nf][4] (integrated traction
 * x, y, z and |traction|), "wall_nuq" [nf][12], and the wall topology
 * ("wall_normal", "wall_area", "wall_h", "wall_bw", "wall_bp", "wall_qw",
 * "wall_ql", "wall_cell", "wall_lf", "wall_bub", "wall_vert"), the initial
 * velocity "wall_u0" [nv][3] and "wall_start" [7] (Laplace iterations,
 * final relative residual, inlet flux, outlet flux, relative imbalance,
 * largest wall face-mean normal velocity, largest vertex speed). */
int nodals_h8_wall_init(nodals_h8* h, int argc, const char* const* argv);
/* Face-averaged Spalding fit (f2_face_averaged_spalding,
 * modules/wall/stage22d4b_nitsche_live.inc:25), Nitsche face blocks
 * (s22d4_nitsche_tet_component, :99), component rows and row-L1 relaxation
 * (s22c_finalize_scalar, modules/wall/segregated_vector_tet_22c.inc:81),
 * the gamma-free rAU (s22d4i3a_refinv, modules/coupling/
 * stage22d4i3a_schur.cuh:11), the 3x3 block operator K = T^T M T
 * (m3_build, modules/wall/stage22d4d_coupled_matrixfree.inc:934) and
 * T^T on the right-hand side, at the current u, p, k, omega. */
int nodals_h8_wall_assemble(nodals_h8* h);
/* The fixed block Gauss-Seidel momentum sweeps on K (m3_fixed_coupled_csr,
 * stage22d4d_coupled_matrixfree.inc:967; the momentum_solve call of
 * hxt5e_fgmres.cuh:480-527 in its M3 mode). */
int nodals_h8_wall_momentum(nodals_h8* h);
/* y = B T Q R Q T^T B^T x for host vectors of length cells
 * (S22D4I3ASchur::apply, stage22d4i3a_schur.cuh:70-74). */
int nodals_h8_wall_schur(nodals_h8* h, const double* x, double* y);
/* u += T Q R Q T^T B^T dp for a host vector dp (S22D4I3ASchur::correct,
 * stage22d4i3a_schur.cuh:61-69, with this build's sign convention
 * p += alpha_p dp). */
int nodals_h8_wall_correct(nodals_h8* h, const double* dp);

/* ---------------------------------------------------------------- GNN ---
 * Learned pressure-correction warm start (no reference counterpart; the
 * reference's only data-driven pressure aid is the snapshot POD of
 * turbulence/sst_gpu_v2/source/puretet_profile5_eq5_hxt5b/pod_balanced.inc).
 * A 4-layer message-passing network, width 64, on the graph of face-adjacent
 * cells, with BF16-MFMA forward and backward kernels (FP32 accumulation) and
 * Adam; it predicts each SIMPLE iteration's pressure correction (from the
 * second on) from the continuity residual, the two previous corrections and
 * the wall distance, in units of the previous correction's RMS.
 * The FGMRES pressure solve then starts from the prediction and stops at the
 * same tolerance relative to ||b||; a prediction that does not lower the
 * residual is dropped.  Training samples are recorded from the solver's own
 * SIMPLE iterations (features, converged correction). */
typedef struct nodals_gnn nodals_gnn;
/* A network on the cell graph of an H8 context, weights from seed. */
int nodals_gnn_create_for_h8(nodals_h8* h, uint32_t seed, nodals_gnn** out);
int nodals_gnn_destroy(nodals_gnn* g);
/* info[4] = cells, padded rows, parameter count, recorded samples. */
int nodals_gnn_info(nodals_gnn* g, int64_t info[4]);
int nodals_gnn_get_params(nodals_gnn* g, float* host, int64_t count);
int nodals_gnn_set_params(nodals_gnn* g, const float* host, int64_t count);
/* Clears the training set and sizes it for `capacity` samples. */
int nodals_gnn_reserve(nodals_gnn* g, int32_t capacity);
/* Reads recorded sample i: x[cells][16], t[cells]. */
int nodals_gnn_sample(nodals_gnn* g, int32_t i, float* x, float* t);
/* y = network(x) for x[cells][16] (host arrays). */
int nodals_gnn_forward(nodals_gnn* g, const float* x, float* y);
/* Gradient of mean((y - t)^2) over the cells w.r.t. every parameter
 * (grad[params]) and the loss, without updating. */
int nodals_gnn_gradient(nodals_gnn* g, const float* x, const float* t, float* grad, double* loss);
/* Adam epochs over the recorded samples in a seeded shuffled order; mean
 * loss of the last epoch. */
int nodals_gnn_train(nodals_gnn* g, int32_t epochs, float lr, uint32_t seed, double* last_loss);
int nodals_gnn_evaluate(nodals_gnn* g, double* loss);
/* Attach g to an H8 context (NULL detaches).  warm_start 1: every pressure
 * solve (from the second SIMPLE iteration) starts from the prediction at its
 * residual-minimising multiple (never above the zero start's residual);
 * 2: the prediction is added as the first FGMRES direction; 0: off.  Needs
 * --pressure-solver fgmres.  record_every > 0 records a sample every that
 * many SIMPLE iterations until the set is full.  g must stay alive while
 * attached. */
int nodals_h8_attach_gnn(nodals_h8* h, nodals_gnn* g, int32_t warm_start, int32_t record_every);

/* ---------------------------------------------------------------- SST ---
 * Pointwise k-omega SST-2003 and wall-law evaluations on the GPU, batched
 * over n points (host arrays in and out, copied through device memory).
 *
 * nodals_sst_local replaces nodals::sst2003m::evaluate_host /
 * evaluate_at_distance (turbulence/sst_gpu_v2/source/puretet_profile5_eq5_hxt5b/
 * modules/physics/sst2003m_host.hpp:5-10, sst2003m_distance_core.hpp:9-22):
 * wall distance from dist[n] (floored), or when pipe_xy[2n] is given, the
 * pipe distance radius - |(x, y)|.  gu[n][9] row-major du_i/dx_j, k[n],
 * omega[n], grad_k[n][3], grad_omega[n][3].  out[n][14] = nut F1 F2 sigmaK
 * sigmaW beta gamma strain pk srcOmega reactK reactOmega omegaProduction
 * omegaCross; floored[n][2] = k, omega floored. */
int nodals_sst_local(int64_t n, const double* dist, const double* pipe_xy, double radius, double nu,
                     const double* gu, const double* k, const double* omega, const double* grad_k,
                     const double* grad_omega, double k_floor, double omega_floor, int32_t deviatoric,
                     double* out, int32_t* floored);
/* Spalding friction velocity per sample (nodals::wall::spalding,
 * modules/wall/portable/spalding_core.hpp:86-230): out[n][3] = u_tau y+ beta;
 * ok[n]. */
int nodals_wall_spalding(int64_t n, const double* slip, const double* y, double nu, double* out, int32_t* ok);
/* Five-point least-squares Spalding fit per wall face (fit_profile5_host,
 * modules/wall/portable/wall_host.hpp:11-12): U, U_signed, y are [n][5];
 * out[n][6] = u_tau y+ beta u_mid fit_rms point_u_tau_rel_span; valid[n]. */
int nodals_wall_profile5(int64_t n, const double* U, const double* U_signed, const double* y, double nu,
                         double* out, int32_t* valid);
/* Tangential Spalding fit of a stationary wall per face
 * (nodals::wall::vector3::fit_active, modules/wall/portable/
 * spalding_active.hpp:129-154, with active::fit_values :48-128):
 * samples[n][5][3] velocities at the five wall-normal slots, y[n][5] their
 * heights, nrm[n][3] the unit wall normal, pic[n][3] the Picard
 * velocity; count = active points 1..5 (slots {2}, {0,4}, {0,2,4},
 * {0,1,3,4}, all); zero_limit keeps the fit's molecular nu/y coefficient
 * when a count < 5 fit returns u_tau = 0 (STAGE27_SPALDING_ZERO_LIMIT).
 * out[n][3] = beta u_tau y+; valid[n]. */
int nodals_wall_fit_tangential(int64_t n, const double* samples, const double* y, const double* nrm,
                               const double* pic, double nu, int32_t count, int32_t zero_limit, double* out,
                               int32_t* valid);
/* The five-point fit of nodals_wall_profile5 with the same arguments and
 * results, allocating nothing: no hipMalloc, hipFree or allocating library
 * call.  No counterpart in the reference; it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid it).  Inputs and outputs are staged through a fixed
 * device array of the library, 11.3 MB on each GPU it runs on, held until the
 * process ends, so n is at most 65536 (larger n is rejected with an error
 * before any device work).  Runs on the calling thread's current device and
 * returns when the results are in out/valid.  Calls that overlap on one GPU
 * share that array and race. */
int nodals_wall_profile5_noalloc(int64_t n, const double* U, const double* U_signed, const double* y, double nu,
                                 double* out, int32_t* valid);
/* Q1+BF2 hexahedral element blocks (the SST application's 14-DOF hex:
 * modules/foundation/hxt2_foundation.inc:166-225 hex_q1_ref, hex_basis_ref,
 * hex_metric, phys_grad14, hex_volume_kernel; hxt4b_foundation.inc:246-345
 * hex_conv_assemble_cached), 3x3x3 Gauss: coords[n][8][3] (vertex order of
 * HexConn), u[n][14][3] (8 vertices, then the face bubbles z-, z+, y-, x+,
 * y+, x-); diff[n][14][14] = nu * sum_q det w grad(phi_a).grad(phi_b),
 * conv[n][14][14] = sum_q det w phi_a (u_q . grad(phi_b)); bad = number of
 * Gauss points with a non-positive Jacobian (their terms are left out). */
int nodals_hex_element_blocks(int64_t n, const double* coords, const double* u, double nu, double* diff,
                              double* conv, int32_t* bad);
/* Host only (no GPU): wall-normal segments traced through a tetrahedral mesh
 * by face-connected traversal (the SST application's ray-traced wall
 * sampling, turbulence/sst_gpu_v2/source/puretet_profile5_eq5_hxt5b/modules/
 * wall/portable/tet_ray_trace.hpp TetRayTrace::trace, used by
 * modules/wall/tet_wall_ray_cache.inc:3-205).  points[np][3], tets[nt][4];
 * ray r starts at origin[r] inside cell start[r], runs along dir[r] and is
 * sampled at the ndist increasing distances dist[r][ndist].  Per (ray,
 * sample): cell (-1 for a failed ray), bary[4] (barycentric in that cell),
 * crossings (cells left before the sample).  status[r] = 0, or 1 when the
 * segment leaves the fluid or cannot be followed.  certified (may be null):
 * for a failed ray the distance, to dist[r][0] * 2^-28, up to which its
 * first sample could be traced (the local shortening's bisection); for a
 * traced ray its last distance.  Returns non-zero on invalid input (bad
 * indices, degenerate or non-manifold tetrahedra). */
int nodals_ray_trace(int64_t np, const double* points, int64_t nt, const int32_t* tets, int64_t nrays,
                     const int32_t* start, const double* origin, const double* dir, int32_t ndist, const double* dist,
                     int32_t max_crossings, double tol, int32_t* cell, double* bary, int32_t* crossings,
                     int32_t* status, double* certified);
/* OpenFOAM-style omega wall target (evaluate_of_auto_host,
 * modules/wall/portable/wall_host.hpp:5-9): in[n][15] = y nu Ut uTau k omega
 * kFloor omegaFloor kappa E beta1 reBlend prodLimit pk nut; shared[n];
 * out[n][8] = omegaTarget omegaVis omegaLog lamFrac rey yPlusK pkAuto
 * nutSpalding. */
int nodals_wall_omega_target(int64_t n, const double* in, const int32_t* shared, double* out);

#ifdef __cplusplus
}
#endif

#endif /* NODALS_HIP_H */
