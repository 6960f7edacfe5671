/*
 * mpsd_gs.h -- C-ABI of whole ground-state solves on the device (the C++ Contractor of
 * include/mpsdyn/mpsdyn.h behind plain C).  One call = build the model MPO, draw a seeded
 * random MPS, and run Contractor::findGroundState.  Mirrors the reference's ground-state
 * drivers: src/programs/ising/testIsing.cpp:71-91, src/programs/heisenberg/testHeisenberg.cpp:84-110,
 * src/programs/banulsm/schwinger/schwingerPD.cpp:51,93-102, calling
 * Contractor::findGroundState (src/core/cpp/Contractor.cpp:1352-1533).
 */
#ifndef MPSD_GS_H
#define MPSD_GS_H

#include <stdint.h>

#include "mpsd.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MPSD_MODEL_ISING 0      /* p = {J, g, h, K}: sum J ZZ + g X + h Z (IsingHamiltonian.h:19-25) */
#define MPSD_MODEL_HEISENBERG 1 /* p = {Jx, Jy, Jz, h}: S = sigma/2 (HeisenbergHamiltonian.cpp:266-331) */
#define MPSD_MODEL_SCHWINGER 2  /* p = {mu, x, l0, alpha} (SchwingerHamiltonian.cpp:13-21) */

typedef struct mpsd_gs_config {
  int model;
  int L;              /* chain length */
  int D;              /* MPS bond dimension */
  double p[4];        /* model parameters, see MPSD_MODEL_* */
  double offset;      /* findGroundState offset argument */
  double convtol;     /* Contractor::setConvTol (1e-12 for the L=20 reference value, 1e-10 elsewhere) */
  double eigtol;      /* residual tolerance of each local eigensolve (0: 1e-10) */
  int max_halfsweeps; /* <= 0: until converged */
  int precision;      /* MPSD_PREC_* of the matvecs */
  uint64_t seed;      /* random initial MPS */
  int solver;         /* MPSD_SOLVER_*: local eigensolver (Contractor::setEigenSolver) */
  int want_residual;  /* 1: also compute residual_norm = ||H psi - E psi|| of the result (built-in
                         models only); 0: not computed (residual_norm = -1) */
  int breakdown;      /* 1: time the sweep's phases (t_env, t_eig, t_gauge, t_matvec of the result);
                         the stream is drained at each phase boundary */
  int tol_schedule;   /* 0: every local problem solved to eigtol (the reference's behaviour);
                         1: FP64 local solves of a half-sweep to max(eigtol, min(1e-3, 1e-2 |dE|/|E| of
                         the previous half-sweep, 1e-2 times the previous half-sweep's tolerance)), loose
                         while the sweep is far from converged (the precision forms always schedule) */
} mpsd_gs_config;

#define MPSD_SOLVER_LANCZOS 0 /* device Lanczos on the matvec (in place of primme / arpack) */
#define MPSD_SOLVER_FULLEIG 1 /* fulleig: Hermitian multiplier + dense eigensolver, FP64 */

#define MPSD_GS_MAX_HISTORY 256

typedef struct mpsd_gs_result {
  double energy;          /* ground-state energy (findGroundState *lambda) */
  double initial_energy;  /* <psi0|H|psi0> of the start state */
  int halfsweeps;
  int64_t matvecs;
  int64_t env_updates;
  double seconds;         /* wall time of findGroundState */
  double residual;        /* last local eigen-residual */
  int64_t device_bytes;   /* HBM held by the context after the solve */
  int64_t matvecs_lowp;   /* matvecs done in the low-precision form (precision != FP64) */
  int refine_from;        /* first half-sweep refined in FP64 (-1: FP64 throughout) */
  int n_history;
  double history[MPSD_GS_MAX_HISTORY]; /* energy at the end of each half-sweep */
  double residual_norm;   /* ||H psi - E psi|| of the final (unit) state when cfg->want_residual, computed
                             as sqrt(<psi|(H - E)^2|psi>) on the device over the MPO of (H - E)^2, with
                             -E spread over the sites' local terms; else -1 */
  double t_env, t_eig, t_gauge; /* with cfg->breakdown: seconds in environment updates, local eigensolves
                                   and gauge steps */
  double t_matvec;        /* with cfg->breakdown: seconds of matvec kernels inside the eigensolves */
  double matvec_flops;    /* algorithmic flops of all matvecs (8 per complex MAC, per site shape) */
  double env_flops;       /* algorithmic flops of all environment updates */
  double residual_floor;  /* with cfg->want_residual: the rounding level of residual_norm (the square root
                             of the rounding of <(H - E)^2>); residuals below it are not resolved */
  /* per half-sweep (the first n_history): wall seconds, matvecs, local eigensolver tolerance */
  double hs_seconds[MPSD_GS_MAX_HISTORY];
  double hs_tol[MPSD_GS_MAX_HISTORY];
  int64_t hs_matvecs[MPSD_GS_MAX_HISTORY];
} mpsd_gs_result;

/* Run one solve on ctx.  Returns MPSD_OK or an MPSD_ERR_* code (message: mpsd_last_error). */
int mpsd_gs_solve(mpsd_ctx* ctx, const mpsd_gs_config* cfg, mpsd_gs_result* out);

/* Same, and also write the final MPS with MPS::exportMPS's binary layout
 * (src/core/cpp/MPS.cpp:440-459) to `export_path` (NULL: no export). */
int mpsd_gs_solve_export(mpsd_ctx* ctx, const mpsd_gs_config* cfg, const char* export_path,
                         mpsd_gs_result* out);

/* findGroundState from a given start state: the MPS in file mps_in (exportMPS layout), for the
 * Hamiltonian of cfg's model or of the MPO file at mpo_path (format MPSD_MPO_*) when that is not
 * NULL.  As the reference's findGroundState does (src/core/cpp/Contractor.cpp:1360-1370), the
 * state is first fitted to the MPO's physical dimensions (MPS::adjustPhysDimensions,
 * MPS.cpp:722-756) and then, for cfg->D > 0, grown by zero padding to bond cfg->D
 * (MPS::increaseBondDimension, MPS.cpp:650-719) or cut to it by the truncating copy
 * MPS(init, D) (MPS.cpp:761-773); cfg->D = 0 keeps its bond.  cfg->seed is not used.
 * export_path as in mpsd_gs_solve_export. */
int mpsd_gs_solve_from(mpsd_ctx* ctx, const mpsd_gs_config* cfg, const char* mpo_path, int format, const char* mps_in,
                       const char* export_path, mpsd_gs_result* out);

/* <psi|H|psi> of an MPS file in exportMPS layout for the model of cfg (Contractor::contract,
 * src/core/cpp/Contractor.cpp:528), e.g. to check a checkpoint. */
int mpsd_gs_expectation(mpsd_ctx* ctx, const mpsd_gs_config* cfg, const char* mps_path, double* re,
                        double* im);

/* MPO files.  MPSD_MPO_BINARY is MPO::saveMPO / loadMPO (src/core/cpp/MPO.cpp:214-232,
 * 293-316): int length, then per operator mwArray::save (int rank, int dims[rank], 2*n doubles
 * of interleaved complex128, column-major [du, Dl, dd, Dr]).  MPSD_MPO_TEXT is
 * MPO::exportMPOtext / importMPOtext (MPO.cpp:234-271): "length " then per operator
 * mwArray::savetext, "rank dims... re im ..." separated by spaces (written with 17 significant
 * digits; files with fewer digits read the same way). */
#define MPSD_MPO_BINARY 0
#define MPSD_MPO_TEXT 1

/* Write the Hamiltonian MPO of cfg's model (cfg->model, L, p) to `path`.  Host-only: ctx may
 * be NULL (then no error message is kept). */
int mpsd_mpo_save(mpsd_ctx* ctx, const mpsd_gs_config* cfg, const char* path, int format);

/* Ground state of the chain Hamiltonian read from an MPO file (the reference's programs that
 * importMPOtext / loadMPO their Hamiltonian, e.g. src/programs/banulsm/bosehub/bosehubFile.cpp,
 * then call Contractor::findGroundState).  The file fixes L and the physical dimension of each
 * site (any d >= 1, du == dd); cfg supplies D, offset, tolerances, precision, seed and solver,
 * and its model, L and p are not read.  Rejects files whose operators are not rank 4, not square
 * in the physical index or whose bonds do not chain (MPSD_ERR_SHAPE).  export_path as in
 * mpsd_gs_solve_export. */
int mpsd_gs_solve_mpo(mpsd_ctx* ctx, const mpsd_gs_config* cfg, const char* mpo_path, int format,
                      const char* export_path, mpsd_gs_result* out);

/* <psi|H|psi> of an MPS file (exportMPS layout) for the MPO in an MPO file. */
int mpsd_mpo_expectation(mpsd_ctx* ctx, const char* mpo_path, int format, const char* mps_path, double* re,
                         double* im);

/* The lowest nlevels eigenstates of the Hamiltonian of cfg's model, or of the MPO file at
 * mpo_path (format MPSD_MPO_*) when it is not NULL: level 0 by Contractor::findGroundState,
 * level k > 0 by Contractor::findNextExcitedState against levels 0..k-1
 * (src/core/cpp/Contractor.cpp:2676-2727), as the reference's excited-state programs chain
 * them.  Each level starts from its own random state (seed cfg->seed + 7919 k) with bond
 * cfg->D; excited levels are solved in FP64 with the Hermitian projector multiplier
 * (TensorMultiplierProjMulti).  exc_orth_tol > 0 enables Contractor::setExcOrthTol.
 * energies: nlevels values.  overlaps (may be NULL): nlevels x nlevels, column-major,
 * |<level_i|level_j>|.  results (may be NULL): per-level statistics. */
int mpsd_gs_solve_levels(mpsd_ctx* ctx, const mpsd_gs_config* cfg, const char* mpo_path, int format, int nlevels,
                         double exc_orth_tol, double* energies, double* overlaps, mpsd_gs_result* results);

/* One partial sweep of the MPS in file mps_in (exportMPS layout) over sites pos0 .. lim
 * (excluded; moving right when lim >= pos0): Contractor::sweepPart (Contractor.cpp:1260-1350)
 * when nproj = 0, else Contractor::sweepPartWithProjectorPenalty (Contractor.cpp:2575-2666) with
 * the states in proj_paths and the multiplier weights `weights` (H - sum_k weights[k] |k><k|).
 * The Hamiltonian is cfg's model or the MPO file at mpo_path.  energy (may be NULL) receives
 * <psi|H|psi> / <psi|psi> of the result; mps_out (may be NULL) receives the result. */
int mpsd_gs_sweep_part(mpsd_ctx* ctx, const mpsd_gs_config* cfg, const char* mpo_path, int format, const char* mps_in,
                       int nproj, const char* const* proj_paths, const double* weights, int pos0, int lim,
                       const char* mps_out, double* energy);

/* MPS(const MPS&, D) (src/core/cpp/MPS.cpp:761-773) on an exported state: the copy is cut to
 * bond D by gaugeCond('L', false, D), truncated SVDs from the right with the norm kept in the
 * norm factor; written to mps_out in exportMPS layout. */
int mpsd_mps_truncate(mpsd_ctx* ctx, const char* mps_in, int D, const char* mps_out);

/* MPS::gaugeCond(pos, dir, next) (src/core/cpp/MPS.cpp:366-419) on an exported state: site pos
 * (0-based) becomes an isometry by QR (dir 'R') or LQ (dir 'L'); the factor goes into the
 * neighbour when next is set or the bond changed; with next set at the chain's end (gaugeL at
 * site 0, gaugeR at the last site) the 1x1 factor's magnitude multiplies the norm factor
 * (MPS.cpp:380-387, 402-409).  The result is written to mps_out. */
int mpsd_mps_gauge_site(mpsd_ctx* ctx, const char* mps_in, int pos, char dir, int next, const char* mps_out);

/* Lowest eigenpair of an Heff by the device FP64 Lanczos solver.  Replaces
 * Contractor::solveEigs(multi, 1, "SR", ...) (src/core/cpp/Contractor.cpp:3067-3138) and the
 * eigensolver wrappers behind it (src/wrapper/eigs.cpp:71-422).  x_dev: device vector of
 * mpsd_heff_size() complex numbers, the start vector on entry (warm start), the unit
 * eigenvector on exit.  tol: residual tolerance relative to max(1, |value|). */
int mpsd_heff_eigs_lowest(mpsd_heff* h, double* x_dev, double tol, int max_krylov, double* value,
                          double* residual, int* matvecs);

#ifdef __cplusplus
}
#endif

#endif /* MPSD_GS_H */
