/* krome.h -- the C ABI of a network library built by this project
 * (libkrome_<network>.so): the batched per-cell chemistry + Tgas solve on
 * an AMD MI355X (gfx950).  The signatures, flag values and option keys are
 * those of KROME-X's generated C interface, so a caller of the reference's
 * libkrome links against this one unchanged; each declaration cites the
 * reference declaration it replaces (runtime/interfaces/C/krome.h) and the
 * implementation (runtime/interfaces/C/krome.cpp).
 *
 * Differences of this build (not of the signatures):
 *  - every solve runs on the GPU: option path 0 (auto) and 2 (device) take
 *    the device for any number of cells, path 1 (per-cell host loop)
 *    returns KROME_FLAG_UNSUPPORTED; a cell that fails on the device keeps
 *    its input and its flag (no host re-solve);
 *  - linsol 2 (LAPACK) returns KROME_FLAG_UNSUPPORTED; linsol 3 (the
 *    static-pivot sparse LU) runs in the plain kernel and the lane queue;
 *    the lockstep persistent kernel returns KROME_FLAG_UNSUPPORTED for it;
 *  - photon bins and fields as the reference's (no broadening, dust
 *    opacity, shielding or co-evolved field); no dust bins;
 *  - host-array calls (solves, element totals, conservation) copy their
 *    rows to the device, compute there and copy back;
 *  - extra option keys select this build's kernels: device_kernel (0 the
 *    plain FP64 kernel; 1 the persistent batched kernel: the lockstep
 *    kernel up to 16 equations, the lane queue above; 2 the lane queue
 *    for any network: each lane solves whole cells and takes the next
 *    from a device work queue), mixed_precision (lockstep kernel: 1 =
 *    FP32 Jacobian, LU and triangular solves with the FP64 Newton
 *    residual, state and error test, with Tgas evolved and the
 *    difference-quotient Jacobian; with Tgas fixed the matrix stays FP64),
 *    queue_chunk (lockstep kernel: each wave takes this many cells of the
 *    queue per atomic and hands them to its idle lanes; 0, or below 64,
 *    takes exactly the cells its idle lanes need), waves_per_cu (resident
 *    waves per CU; 0 = automatic) and lane_block (the lane queue's lanes
 *    per workgroup, 64 or 128).  None changes a cell's result, only which
 *    lane integrates it (the lane queue computes the plain kernel's
 *    arithmetic).
 */
#ifndef KROME_MI355X_H
#define KROME_MI355X_H

/* the integrator's failures (CVODE's values; krome.h:62-75) */
#define KROME_FLAG_TOO_MUCH_WORK (-1)
#define KROME_FLAG_TOO_MUCH_ACC (-2)
#define KROME_FLAG_ERR_FAILURE (-3)
#define KROME_FLAG_CONV_FAILURE (-4)
#define KROME_FLAG_LSETUP_FAIL (-6)
#define KROME_FLAG_LSOLVE_FAIL (-7)
#define KROME_FLAG_RHSFUNC_FAIL (-8)
#define KROME_FLAG_FIRST_RHSFUNC_ERR (-9)
#define KROME_FLAG_REPTD_RHSFUNC_ERR (-10)
#define KROME_FLAG_UNREC_RHSFUNC_ERR (-11)
#define KROME_FLAG_NLS_FAIL (-16)
#define KROME_FLAG_ILL_INPUT (-22)
#define KROME_FLAG_TOO_CLOSE (-27)
/* KROME's flags (krome.h:76-89) */
#define KROME_FLAG_OK 0
#define KROME_FLAG_NO_DUST (-102)
#define KROME_FLAG_NO_FIELD (-103)
#define KROME_FLAG_UNSUPPORTED (-104)
#define KROME_FLAG_PARAM_UNSET (-105)
#define KROME_FLAG_CONSERVE_GAS_TOTAL (-106)
#define KROME_FLAG_SURFACE_BINS (-107)
#define KROME_FLAG_NO_LAPACK (-108)
#define KROME_FLAG_NO_KOKKOS (-109)
#define KROME_FLAG_NONPOSITIVE_T (-110)
#define KROME_FLAG_DEVICE_MEMORY (-111)
#define KROME_FLAG_CONSERVE_UNREACHED (-112)
#define KROME_FLAG_CONSERVE_LARGE 2
#define KROME_FLAG_HOST_RECOVERED 1

/* species-array layouts of the device-resident calls (krome.h:326-327) */
#define KROME_LAYOUT_SPECIES_MAJOR 0
#define KROME_LAYOUT_CELL_MAJOR 1

#ifdef __cplusplus
extern "C" {
#endif

/* a one-line description of a flag, never NULL (krome.h:95;
 * krome.cpp krome_flag_message) */
const char* krome_flag_message(int flag);

/* network description (krome.h:98-113) */
const char* krome_version(void);
const char* krome_network_name(void);
int krome_nspec(void);
int krome_nrea(void);
int krome_npar(void);
const char* krome_species_name(int i);
int krome_species_index(const char* name);
const char* krome_param_name(int i);
int krome_param_index(const char* name);
void krome_params_default(double* par);
const char* krome_param_required_by(int i);

/* solver options, an opaque handle set by key (krome.h:116-173): the keys
 * of krome.h plus this build's device_kernel, mixed_precision,
 * queue_chunk, waves_per_cu, lane_block; -1 for an unknown key or a bad
 * value */
typedef struct krome_opts krome_opts;
krome_opts* krome_opts_new(void);
void krome_opts_free(krome_opts* h);
int krome_opts_set(krome_opts* h, const char* key, double value);
int krome_opts_get(const krome_opts* h, const char* key, double* value);

/* per-species tolerances (krome.h:175-185): atol[NSPEC] [cm^-3] /
 * rtol[NSPEC] replacing atol_n / rtol for each species; a negative entry
 * takes the scalar, NULL the scalar for every species.  Copied into the
 * handle.  Set returns -1 for a NaN, or a zero atol.  Get writes the
 * tolerance each species gets (its entry, or the scalar).  With some
 * species' rtol different from the scalar the error weights are
 * 1 / fma(rtol_i, |y_i|, atol_i) (the reference's per-slot form) in every
 * kernel */
int krome_opts_set_atol_species(krome_opts* h, const double* atol);
int krome_opts_set_rtol_species(krome_opts* h, const double* rtol);
int krome_opts_get_atol_species(const krome_opts* h, double* atol);
int krome_opts_get_rtol_species(const krome_opts* h, double* rtol);

/* the device backend (krome.h:193-201): compiled in (1); initialize
 * selects the HIP device (device_id < 0: the current one) */
int krome_has_device(void);
int krome_device_initialize(void);
int krome_device_finalize(void);
int krome_device_initialize_id(int device_id);

/* photon bins (krome.h:203-228, krome.cpp:313-377): nbins bins with edges
 * 0 < Elo[b] < Ehi[b] [eV] (-1 otherwise); the bin-averaged cross sections
 * of the photo reactions are computed once, on the host, when the bins or
 * the in-bin spectrum (alpha: dN/dE ~ E^-alpha, default 1) are set.  A call
 * with bins set needs the cells' fields Phi[ncell][nbins] [photons cm^-2
 * s^-1] (krome_cells.Phi, or the queries' Phi), else -103.  Doppler
 * broadening and a dust opacity of the bins are not built: only their
 * "none" (0, 0 and NULL) is accepted */
int krome_opts_set_photo(krome_opts* h, int nbins, const double* Elo,
                         const double* Ehi);
int krome_opts_nbins(const krome_opts* h);
int krome_opts_set_photo_alpha(krome_opts* h, double alpha);
int krome_opts_set_photo_broadening(krome_opts* h, double Tgas,
                                    double vturb);
int krome_opts_set_photo_dust_opacity(krome_opts* h, const double* kappa);

/* photo reactions (krome.h:598-617, krome.cpp:1121-1131): count, reaction
 * and reactant index (-1 out of range), photons per event (1), the cross
 * section sigma_r(E) [cm^2] (0 outside its window, NaN for a bad r), and
 * the bin-averaged cross sections sigma[nphoto][nbins] of h's bins (-1
 * without bins) */
int krome_nphoto(void);
int krome_photo_reaction(int r);
int krome_photo_reactant(int r);
double krome_photo_pdiss(int r);
double krome_photo_xsec(int r, double E);
int krome_opts_photo_sigma(const krome_opts* h, double* sigma);
/* nbins linear / logarithmic bins in [lower, upper] eV, lower > 0
 * (krome.h:547-551) */
int krome_bins_lin(int nbins, double lower, double upper, double* Elo,
                   double* Ehi);
int krome_bins_log(int nbins, double lower, double upper, double* Elo,
                   double* Ehi);

/* the optional per-cell arrays of a multi-cell call (krome.h:275-279):
 * Phi[ncell][nbins] the photon fields when bins are set (host arrays for
 * krome_solve_cells(_x), device arrays for the _device calls); no dust
 * bins here (Td, adust unused) */
typedef struct {
  double* Phi;
  double* Td;
  double* adust;
} krome_cells;

/* integrate ncell independent cells over their own dt[c] [s]: n[ncell]
 * [nspec] and Tgas (with evolve_T) updated in place, per-cell parameter
 * rows par (par_stride npar, or 0: one row for every cell; NULL: the
 * defaults), flags[ncell] optional.  Returns 0 or the flag of the lowest
 * failing cell; -105 (every cell flagged) when a REQUIRED parameter is
 * unset, -1 on bad arguments (krome.h:291-294, krome.cpp:396-436) */
int krome_solve_cells(int ncell, double* n, double* Tgas, const double* dt,
                      const double* par, int par_stride,
                      const krome_cells* extra, int* flags,
                      const krome_opts* h);

/* the same with mass fractions x[ncell][nspec] and rho[ncell] [g cm^-3]
 * (krome.h:302-305, krome.cpp:438-470) */
int krome_solve_cells_x(int ncell, double* x, const double* rho,
                        double* Tgas, const double* dt, const double* par,
                        int par_stride, const krome_cells* extra,
                        int* flags, const krome_opts* h);

/* device-resident arrays (krome.h:328-350, krome.cpp:1237-1293): n in
 * the species-major (n[i ld + c], ld >= ncell, 0: ncell) or cell-major
 * layout, Tgas, dt, par and flags in device memory; the device is
 * synchronized before and after.  summary (host, may be NULL) counts the
 * cells solved and failed.  The mass-fraction variant takes x in either
 * layout and rho[ncell] on the device (krome.cpp:1295-1304); the cells
 * that succeed with dt > 0 are converted back, failed cells keep x. */
typedef struct {
  long device;
  long host;
  long failed;
} krome_device_summary;
int krome_solve_cells_device(int ncell, double* n, int layout, long ld,
                             double* Tgas, const double* dt,
                             const double* par, int par_stride,
                             const krome_cells* extra, int* flags,
                             krome_device_summary* summary,
                             const krome_opts* h);
int krome_solve_cells_x_device(int ncell, double* x, const double* rho,
                               int layout, long ld, double* Tgas,
                               const double* dt, const double* par,
                               int par_stride, const krome_cells* extra,
                               int* flags, krome_device_summary* summary,
                               const krome_opts* h);

/* This build's extension, no counterpart in the reference: the
 * device-resident solve of krome_solve_cells_device for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results, less the photon field; it calls no hipMalloc or hipFree: its
 * buffers (the persistent workspace, the queue order, a parameter row and
 * flags when none are given) live in one static device array per GPU,
 * held from the library's load until the process ends (about 140 MB for
 * primordial3: the FP64 persistent workspace of 65536 lanes, capped at
 * 160 MiB, plus 12.5 MB).  At most 2^20 cells per call (-1 beyond);
 * networks with photo reactions, photon bins and conserve 2 return
 * KROME_FLAG_UNSUPPORTED with every cell flagged; the persistent kernel
 * returns it too where its workspace exceeds the cap (above about 12
 * equations), the plain kernel then runs in chunks that fit.  Calls on
 * one GPU share its array: host threads are serialized by a lock, and
 * calls that overlap on one GPU from several processes race. */
int krome_solve_cells_device_noalloc(int ncell, double* n, int layout,
                                     long ld, double* Tgas, const double* dt,
                                     const double* par, int par_stride,
                                     int* flags,
                                     krome_device_summary* summary,
                                     const krome_opts* h);

/* This build's extension, no counterpart in the reference (which has no
 * adjoint): the discrete adjoint of the per-cell BDF solve.  Integrates
 * the cells like krome_solve_cells_device (n species-major [NSPEC][ncell],
 * Tgas, dt, flags [ncell], par [ncell][NPAR] or one row, all in device
 * memory, n and Tgas updated in place) with every rate k_r scaled by
 * exp(lnk[c][r]) (lnk [ncell][NREA] on the device, NULL: the network's
 * rates), in FP64 with its own integrator.  With a cotangent of the
 * final state (gn [NSPEC][ncell] and/or gT [ncell], device, NULL: zero;
 * both NULL: the forward only), it also writes the vector-Jacobian
 * products g_lnk [ncell][NREA] (by lnk), g_n0 [NSPEC][ncell] (by the
 * initial densities) and, with evolve_T, g_T0 [ncell] (by the initial
 * Tgas): the gradient of the integration as the forward run stepped it
 * (its step sizes, orders and accepted attempts held fixed).  The forward
 * is integrated again for the gradient, with a checkpoint every
 * checkpoint_steps accepted steps (option, default 32) in a pool of
 * checkpoint_memory bytes (option, 0: room for about 1536 steps per cell
 * of a launch of up to 65536 cells); a cell that runs out of it gets
 * KROME_FLAG_DEVICE_MEMORY and, like any failed cell, keeps its inputs
 * with the identity's gradient.  mem_bytes (host, may be NULL): the peak
 * device bytes of the call.  conserve != 0, photo reactions, photon bins,
 * linsol 2 and path 1 return KROME_FLAG_UNSUPPORTED with every cell
 * flagged; device_kernel and mixed_precision are ignored.  Returns 0, the
 * lowest failing cell's flag, -105, or -1 on bad arguments. */
int krome_adjoint_cells_device(int ncell, double* n, double* Tgas,
                               const double* dt, const double* par,
                               int par_stride, const double* lnk, int* flags,
                               const double* gn, const double* gT,
                               double* g_lnk, double* g_n0, double* g_T0,
                               double* mem_bytes, const krome_opts* h);

/* This build's extension, no counterpart in the reference's C ABI: the
 * device-resident solve with the per-cell counters of krome::SolveStats
 * (runtime/krome.hpp:257-268; 9 per cell: steps, rhs_evals, jac_evals,
 * rhs_evals_jac, lin_setups, newton_iters, newton_fails, err_fails,
 * restarts) into the device array stats[ncell][9] (may be NULL), for
 * benchmarks and tests.  par is required; the call is asynchronous on the
 * null stream (the caller synchronizes).  It takes no photon field: -103
 * when h has photon bins (the photo rates are 0 without bins). */
int krome_solve_cells_device_stats(int ncell, double* n, int layout, long ld,
                                   double* Tgas, const double* dt,
                                   const double* par, int par_stride,
                                   int* flags, long* stats,
                                   const krome_opts* h);

/* extension (no counterpart in the reference): the duration in ms of the
 * last device solve's (or conservation's) kernels, from HIP events
 * recorded around their launches on their stream; waits for them to
 * finish.  0, or -1 when no such call ran yet. */
int krome_device_last_kernel_ms(double* ms);

/* the name of a flag, e.g. "FLAG_TOO_MUCH_WORK" (the reference's
 * krome::flag_name, runtime/core/solve_cell.hpp:106-137) */
const char* krome_flag_name(int flag);

/* ---- queries: rates, fluxes and the thermal balance of states ------------
 * (krome.h:378-412, krome.cpp:106-133, 621-644, 753-770; query.hpp:37-108)
 * n[ncell][nspec] and Tgas[ncell] on the host; par, par_stride as for
 * krome_solve_cells (NULL: the defaults); Phi[ncell][nbins] the fields of
 * the cells when h has photon bins; h may be NULL (no bins).  Evaluated as the solver sees them:
 * the rates at min(max(Tgas, Tcmb), 1e9), the thermal terms at
 * min(max(Tgas, 2.73), 1e9).  0, -1 on bad arguments, or -105 when a cell's
 * REQUIRED parameter is unset (that cell's outputs untouched).  Computed
 * on the device. */
int krome_get_rates(int ncell, const double* n, const double* Tgas,
                    const double* par, int par_stride, const double* Phi,
                    const krome_opts* h, double* k);
int krome_get_fluxes(int ncell, const double* n, const double* Tgas,
                     const double* par, int par_stride, const double* Phi,
                     const krome_opts* h, double* fluxes);
/* the thermal balance [erg cm^-3 s^-1]: totals heat[ncell], cool[ncell],
 * the active terms heats[ncell][nheat], cools[ncell][ncool] (named by
 * krome_heating_name / krome_cooling_name, the reference's order) and
 * dTdt[ncell] [K/s]; any output may be NULL */
int krome_nheat(void);
int krome_ncool(void);
const char* krome_heating_name(int i);
const char* krome_cooling_name(int i);
int krome_get_thermal(int ncell, const double* n, const double* Tgas,
                      const double* par, int par_stride, const double* Phi,
                      const krome_opts* h, double* heat, double* cool,
                      double* heats, double* cools, double* dTdt);

/* ---- element totals and consistent advection --------------------------- */

/* species and element metadata (krome.h:642-650, krome.cpp:1182-1196):
 * mass [g] (NaN out of range), charge (0 out of range), the number of
 * elements, an element's name (NULL out of range; elements sorted by
 * name), atoms of element e in species i (-1 out of range) */
double krome_species_mass(int i);
int krome_species_charge(int i);
int krome_nelem(void);
const char* krome_element_name(int e);
int krome_species_natoms(int i, int e);

/* element totals tot[ncell][nelem] [atoms cm^-3] of host rows n[ncell]
 * [nspec]; _x: per gram of gas [atoms g^-1] from mass fractions
 * (krome.h:473-477, krome.cpp:929-943).  0, or -1 on bad arguments.
 * Computed on the device. */
int krome_get_element_totals(int ncell, const double* n, double* tot);
int krome_get_element_totals_x(int ncell, const double* x, double* tot);

/* consistent advection (krome.h:478-495, krome.cpp:945-977): each cell of
 * n[ncell][nspec] (or mass fractions x, ref per gram) moved to its
 * reference element totals ref (ref_stride 0: one row for every cell,
 * nelem: a row per cell) and to zero net charge: negatives clipped, the
 * elements by the minimal relative correction, then the electrons from
 * neutrality.  par, par_stride as for krome_solve_cells (only the
 * network's freeze directives read them; these networks have none).
 * flags[ncell] (may be NULL): 0, KROME_FLAG_CONSERVE_UNREACHED or the
 * warning KROME_FLAG_CONSERVE_LARGE.  Returns the lowest failing cell's
 * flag (negative), 0, or -1 on bad arguments.  Host rows are copied to
 * the device, corrected there and copied back. */
int krome_conserve_cells(int ncell, double* n, const double* ref,
                         int ref_stride, const double* par, int par_stride,
                         int* flags);
int krome_conserve_cells_x(int ncell, double* x, const double* ref,
                           int ref_stride, const double* par,
                           int par_stride, int* flags);
/* the same on device arrays (krome.h:351-360, krome.cpp:1306-1343): n or
 * x in either layout (ld as for krome_solve_cells_device, ignored for the
 * cell-major layout), ref, par and flags in device memory; the device is
 * synchronized before the call and when it returns */
int krome_conserve_cells_device(int ncell, double* n, int layout, long ld,
                                const double* ref, int ref_stride,
                                const double* par, int par_stride,
                                int* flags);
int krome_conserve_cells_x_device(int ncell, double* x, int layout, long ld,
                                  const double* ref, int ref_stride,
                                  const double* par, int par_stride,
                                  int* flags);

#ifdef __cplusplus
}
#endif

#endif
