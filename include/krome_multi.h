/* krome_multi.h -- several networks in one persistent launch
 * (libkrome_multi.so).
 *
 * This build's extension: no counterpart in the reference.  The reference
 * generates one library per network (runtime/interfaces/C/krome.h: NSPEC,
 * NREA and the tables are compile-time constants), so a launch mixing
 * networks is a new entry point beside the per-network ones.  The library
 * holds the networks primordial (0, Tgas fixed) and primordial3 (1, Tgas
 * evolved); each batch is solved as krome_solve_cells_device with the
 * persistent kernel would solve it (device_kernel 1), bitwise, in a single
 * launch shared by both batches.
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
 * meaning; the others take their defaults (jac auto, clip_negative 1,
 * conserve 0, linsol auto, max_err_fails 7, max_conv_fails 10,
 * max_restarts 5) */
typedef struct {
  int network;       /* 0 primordial, 1 primordial3 (krome_multi_network_name) */
  int ncell;
  double* n;
  double* Tgas;
  const double* dt;
  const double* par;
  int* flags;
  int evolve_T;      /* 0 for primordial, 1 for primordial3 */
  double rtol, atol_n, atol_T;
  long max_steps;
} krome_multi_batch;

int krome_multi_nnet(void);
const char* krome_multi_network_name(int i);  /* NULL out of range */
int krome_multi_nspec(int i);                 /* -1 out of range */
int krome_multi_npar(int i);

/* solve nbatch (1 or 2) batches of distinct networks in one persistent
 * launch; mixed_precision as the option of that name (primordial's
 * Tgas is fixed: its matrix stays FP64); kernel_ms (may be
 * NULL) receives the launch's duration from HIP events.  Synchronizes
 * before and after.  Returns 0, the lowest failing cell's flag of the
 * first batch (primordial3 first) with a failure, -105 when a REQUIRED
 * parameter is unset (that batch's flags set, nothing integrated), -104
 * when evolve_T does not match the network, or -1 on bad arguments. */
int krome_multi_solve_device(const krome_multi_batch* b, int nbatch,
                             int mixed_precision, double* kernel_ms);

#ifdef __cplusplus
}
#endif

#endif
