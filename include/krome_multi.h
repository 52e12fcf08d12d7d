/* krome_multi.h -- several networks in one persistent launch
 * (libkrome_multi.so: primordial and primordial3; libkrome_multi_c3.so:
 * primordial, primordialZ and cloud, the uneven batch).
 *
 * This build's extension: no counterpart in the reference.  The reference
 * generates one library per network (runtime/interfaces/C/krome.h: NSPEC,
 * NREA and the tables are compile-time constants), so a launch mixing
 * networks is a new entry point beside the per-network ones.  A multi
 * library is generated for a list of networks (krome-x_amd/build.py
 * MULTI); each batch is solved as krome_solve_cells_device with the
 * persistent kernel (device_kernel 1) would solve it, in a single launch
 * shared by all the batches: networks of up to 16 equations in the
 * lockstep persistent kernel, networks above 64 species (Tgas fixed) a
 * wave per cell, the others in the lane queue (each lane solves whole
 * cells); every wave and lane takes its next cells from the network's
 * device work queue.
 */
#ifndef KROME_MULTI_MI355X_H
#define KROME_MULTI_MI355X_H

#ifdef __cplusplus
extern "C" {
#endif

/* one network's cells, every array in device memory: n species-major
 * (n[i * ncell + c]), Tgas, dt and flags [ncell], par [ncell][npar] (the
 * network's parameter rows, krome_params_default order).  The solver
 * options of the per-network ABI that these fields name keep their
 * meaning; the others take their defaults (clip_negative 1, conserve 0,
 * linsol auto, max_err_fails 7, max_conv_fails 10, max_restarts 5) */
typedef struct {
  int network;       /* the library's index (krome_multi_network_name) */
  int ncell;
  double* n;
  double* Tgas;
  const double* dt;
  const double* par;
  int* flags;
  int evolve_T;      /* as the network is built (krome_multi_evolve_T) */
  double rtol, atol_n, atol_T;
  long max_steps;
  int jac;           /* 0 auto, 1 analytic, 2 difference quotients */
  long* stats;       /* NULL, or [ncell][9] counters (as
                        krome_solve_cells_device_stats) */
} krome_multi_batch;

int krome_multi_nnet(void);
const char* krome_multi_network_name(int i);  /* NULL out of range */
int krome_multi_nspec(int i);                 /* -1 out of range */
int krome_multi_npar(int i);
int krome_multi_evolve_T(int i);   /* 1: Tgas evolved, 0 fixed, -1 */
int krome_multi_lockstep(int i);   /* 1: the lockstep kernel, 0: not,
                                      -1 out of range */
int krome_multi_mode(int i);       /* how network i's cells run: 1 the
                                      lockstep kernel, 2 a wave per cell
                                      (above 64 species, Tgas fixed), 0
                                      the lane queue; -1 out of range */

/* the launch shape of the following calls: waves per CU (0: the
 * default, one wave per SIMD) and the lockstep machine's cells per queue
 * grab (0: as many as a wave's idle lanes); -1 out of range */
int krome_multi_set_launch(int waves_per_cu, int queue_chunk);

/* solve nbatch batches of distinct networks in one persistent launch;
 * mixed_precision as the option of that name (the FP32 Newton matrix
 * where Tgas is evolved with difference quotients; the lane queue is
 * FP64); kernel_ms (may be NULL) receives the launch's duration from HIP
 * events.  Synchronizes before and after.  Returns 0, the lowest failing
 * cell's flag of the first batch in launch order with a failure, -105
 * when a REQUIRED parameter is unset (that batch's flags set, nothing
 * integrated), -104 when evolve_T does not match the network, -111 when
 * the workspaces do not fit, or -1 on bad arguments. */
int krome_multi_solve_device(const krome_multi_batch* b, int nbatch,
                             int mixed_precision, double* kernel_ms);

/* release the library's device buffers */
int krome_multi_finalize(void);

#ifdef __cplusplus
}
#endif

#endif
