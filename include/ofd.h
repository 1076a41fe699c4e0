/*
 * ofd.h -- C ABI of the MI355X-native PROFESS energy/potential evaluation and
 * truncated-Newton (NTN) density optimizer (libprofess_amd.so).
 *
 * Conventions (all entry points):
 *   - Real-space arrays are FP64, n1*n2*n3 values, x fastest: element
 *     (i1,i2,i3) at i1 + n1*(i2 + n2*i3), exactly the reference's Fortran
 *     rho(n1,n2,n3) layout (CellInfo.f90:27-36).
 *   - Energies are in Hartree; etable9 uses the reference's eTable(9) slots
 *     (CalPotential.f90:610-644), 0-based here:
 *     [0] total [1] kinetic [2] ion-electron [3] Hartree [4] XC [5] ion-ion
 *     [6] Thomas-Fermi [7] von Weizsacker [8] nonlocal (WT or WGC).
 *   - Every call returns 0 on success and a negative status otherwise; the
 *     message is available from ofd_last_error(). A Fortran caller maps a
 *     nonzero status to QUIT(message), which keeps the reference's fail-stop
 *     behaviour (Output.f90:1505-1549).
 *   - Host pointers unless the name says _device. The caller owns all arrays.
 *   - One context per host thread; a context is not reentrant. Calls on one
 *     context are ordered on its own HIP stream and synchronous at return.
 */
#ifndef PROFESS_AMD_OFD_H
#define PROFESS_AMD_OFD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ofd_ctx ofd_ctx;

/* Grid and cell. cell[3*j + i] = Cartesian component i of lattice vector j
 * (bohr): the columns of the reference's cell%cellReal (CellInfo.f90:247-330). */
typedef struct {
  int n1, n2, n3;
  double cell[9];
} ofd_grid;

/* Kinetic functional and parameters (SetupKEDF.f90:217-322). kinetic = 4 (WT)
 * or 5 (WGC, second order, WGCT = 2). Negative alpha/beta/gamma/rho0/rhoS
 * select the reference's defaults, resolved as KEDFRefresh does
 * (SetupKEDF.f90:613-683) from num_electrons / volume. */
typedef struct {
  int kinetic;
  double lambda, mu;
  double alpha, beta, gamma;
  double rho0, rhoS;
  int precision;   /* 0 = FP64 everywhere; 1 = FP32 FFT/convolution passes
                      (FP64 fields, epilogue and reductions) */
} ofd_functional;

/* Optimizer controls (RhoOptimizers.f90:44-101). conv_check: 1 = ENE,
 * 2 = POT, 3 = BOTH (ReadInputFile.f90:591-611). Negative pot_tol/tole/max_iter
 * select the defaults (1.E-5 and 1.E-6 as single-precision literals, 500).
 * mixed_switch: with precision = 1, potential norm below which the optimizer
 * switches from FP32 to FP64 passes (<= 0: stay in the context's precision).
 * fused_dots: 0 computes every inner product of the NTN loop as the reference
 * does (RhoDirNew.f90, RhoOptN.f90); 1 takes the CG residual norms, the
 * Newton step's |p|^2 and the next |dL/dphi|^2, |phi|^2 from sums the
 * evaluation already produces and from the CG recurrences (one host round
 * trip per inner iteration instead of three; direct products whenever the
 * recurrence would lose accuracy).
 * cg_alg: the NCG formula of dir_type 2, as the input keyword DENC
 * (ReadInputFile.f90:517-533): 1 Polak-Ribiere, 2 (or 0, the default)
 * Hager-Zhang (RhoDirCG.f90:26). */
typedef struct {
  int max_iter;
  double pot_tol;
  double tole;
  int conv_check;
  int max_line;
  double mixed_switch;
  int fused_dots;
  int cg_alg;
} ofd_opt_params;

typedef struct {
  int iterations;          /* outer NTN iterations (RhoOptN.f90:553) */
  int num_evals;           /* CalculatePotentialPlus calls */
  int newton_iters_total;  /* inner CG iterations summed */
  int fp32_evals;          /* evaluations done in FP32 passes */
  double final_pot_norm;
  double seconds;
} ofd_opt_stats;

/* ---- context (replaces the module globals listed in CalPotential.f90:41-116:
 *      grid dims, cell%dV/vol/cellRecip, kinetic, lambda, mu, alpha, beta,
 *      gamma, rho0, rhoS, qTable, keKernel, ionPotReal, ionIonEnergy) ---- */
int ofd_create(const ofd_grid* grid, const ofd_functional* func, double num_electrons, ofd_ctx** out);
void ofd_destroy(ofd_ctx* ctx);
const char* ofd_last_error(const ofd_ctx* ctx);
/* resolved functional parameters: alpha, beta, gamma, rho0, rhoS, lambda, mu */
int ofd_get_params(const ofd_ctx* ctx, double* out7);
/* the cell volume (bohr^3) as computed from the lattice (cell%vol,
 * CellInfo.f90:247-330); the uniform density is num_electrons / volume */
int ofd_get_volume(const ofd_ctx* ctx, double* vol);

/* ionPotReal and ionIonEnergy as RefreshIonTerms leaves them
 * (RefreshIons.f90:196,221). */
int ofd_set_ion_terms(ofd_ctx* ctx, const double* vion_real, double e_ionion);

/* Device build of ionPotReal: cardinal B-spline structure factor per species,
 * V_loc(|G|) from .recpot data, b-spline moduli, conjugate, C2R
 * (IonElectronPotRecipSpline, IonElectronSpline.f90:83-216; FillQIonTable
 * :537-638; FillB CBSpline.f90:93-196; ReadPseudo units/spline
 * ReadIonFile.f90:848-976; PspRecipCubicSpline LocalPseudoPot.f90:136-198).
 * Ions are grouped by species: species s owns ions
 * [type_start[s], type_start[s+1]). frac_xyz: 3 fractional coordinates per ion.
 * Per species s the recpot table is psp_values[psp_offset[s] .. +psp_npts[s])
 * in the file's units (eV*A^3) with psp_maxg_invA[s] the file's maxG (1/A). */
int ofd_build_ion_potential(ofd_ctx* ctx, int nspecies, const int* type_start, const double* frac_xyz,
                            const double* psp_maxg_invA, const int* psp_npts, const int* psp_offset,
                            const double* psp_values, int spline_order, double e_ionion);
/* the ionic potential currently held (n1*n2*n3) */
int ofd_get_ion_potential(ofd_ctx* ctx, double* vion_real);

/* WGC 1D kernel table w(eta), w'(eta), w''(eta) on increasing knots eta, as
 * makeKernel2 produces it (KEDF_KernelODE.f90:254-478). The 3D kernels
 * K1..K4 are then filled on the device (FillWGC_ReciprocalSpace,
 * KEDF_WGCkernel.f90:109-277). */
int ofd_set_wgc_table(ofd_ctx* ctx, int nknots, const double* eta, const double* w, const double* w1,
                      const double* w2);
/* Same table integrated by this library (ofd_wgc_kernel_table below) and set
 * as by ofd_set_wgc_table; returns the knot count in *nknots_out if not NULL. */
int ofd_build_wgc_table(ofd_ctx* ctx, int* nknots_out);
/* The 1D WGC table of makeKernel2 (KEDF_KernelODE.f90:254-478) computed on the
 * host, no device and no context: the ODE of KEDF_KernelODE.f90:481-537
 * integrated from eta = 100 toward 0 on makeKernel2's 12,696 knots by an
 * adaptive Dormand-Prince 5(4) pair at rtol 1e-11. The reference integrates
 * with RKSUITE at TOL 5e-5 (KEDF_KernelODE.f90:376), whose table carries that
 * tolerance's error: against it this table differs by up to about 1e-4 of
 * max|w| in w and w' and 5e-3 of max|w''| in w'' near eta = 1, and WGC
 * energies built on it differ from the reference-table energies at the
 * 1e-6 eV/atom level (4.7e-7 eV/atom in the total, 3e-6 relative in the
 * nonlocal term on the test/optDen cell), not at the 1e-10 per-term level. A caller who needs the reference's own
 * numbers passes the reference's table to ofd_set_wgc_table instead.
 * Writes at most nmax knots; returns the knot count (when nmax is smaller or
 * an output is NULL nothing is written), or a negative code on bad input or
 * failure. */
int ofd_wgc_kernel_table(double alpha, double beta, double gamma, int nmax, double* eta, double* w, double* w1,
                         double* w2);

/* CalculatePotentialPlus(rho, optSqrt, potential, eTable) (CalPotential.f90:41).
 * etable9 == NULL means the reference's call without eTable (no energies). */
int ofd_calpot(ofd_ctx* ctx, const double* rho, int opt_sqrt, double* potential, double* etable9);
/* Same with device-resident rho/potential (pointers from hipMalloc). */
int ofd_calpot_device(ofd_ctx* ctx, const double* rho_device, int opt_sqrt, double* potential_device,
                      double* etable9);

/* Asynchronous form for device-resident callers (no reference counterpart:
 * the reference is synchronous): enqueues the evaluation on the context's
 * stream and returns at once. sums9_device (device, 9 doubles, may be NULL)
 * receives the raw energy sums; ofd_sums_to_etable turns them into eTable. */
int ofd_calpot_enqueue(ofd_ctx* ctx, const double* rho_device, int opt_sqrt, double* potential_device,
                       double* sums9_device);
int ofd_sums_to_etable(const ofd_ctx* ctx, const double* sums9, double* etable9);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory during a call (code under HIP stream capture, serving loops that
 * must not allocate). The evaluation of ofd_calpot_enqueue, enqueued on the
 * context's stream and not waited for, with all of its per-call device
 * scratch (descriptor, energy partial sums, the transformed fields' spectra)
 * in one static device array of ofd_noalloc_limit_bytes() = 64 MiB that each
 * device holds until the process ends: no allocation, free, copy, memset or
 * graph. rho_device, potential_device: n1*n2*n3 device arrays; etable9_device:
 * 9 device doubles receiving eTable (Hartree), or NULL. A grid and functional
 * whose scratch exceeds the limit (Wang-Teter with alpha == beta up to
 * 128^3) is refused with a nonzero return and ofd_last_error. All calls on
 * one device share the array: calls that overlap in time on one device race
 * (order them on one stream). */
int ofd_calpot_noalloc(ofd_ctx* ctx, const double* rho_device, int opt_sqrt, double* potential_device,
                       double* etable9_device);
long long ofd_noalloc_limit_bytes(void);
/* the context's HIP stream (hipStream_t) and a blocking wait on it */
void* ofd_get_stream(ofd_ctx* ctx);
int ofd_sync(ofd_ctx* ctx);
/* Batched evaluation of nb contexts of one grid shape and functional in one
 * launch set (device-resident rho/potential per context); etables: nb x 9 on
 * the host, or NULL. */
int ofd_calpot_many(ofd_ctx* const* ctxs, int nb, const double* const* rho_device, double* const* potential_device,
                    int opt_sqrt, double* etables);

/* ProjNormRhoMinimization(energy, dir_type) (RhoOptN.f90:60-850) with
 * dir_type = 1 (NTN, RhoDirNew.f90), 2 (NCG, RhoDirCG.f90) or 3 (NBF,
 * RhoDirBFGS.f90, 5 stored pairs) and numSpin = 1; rhoR in/out (SYS::rhoR),
 * energy9 out,
 * rho_outcome as RhoOptimizers::rhoOutcome (0 converged, 1 failed). */
int ofd_rho_opt(ofd_ctx* ctx, double* rhoR, int dir_type, const ofd_opt_params* params, double* energy9,
                int* rho_outcome, ofd_opt_stats* stats);

/* ---- spin polarization (numSpin = 2). After ofd_set_spin with nspin = 2
 * every density and potential array of ofd_calpot, ofd_calpot_device,
 * ofd_calpot_enqueue, ofd_calpot_noalloc, ofd_calpot_many, ofd_rho_opt and
 * ofd_ion_forces holds both channels, 2*n1*n2*n3 values: spin up, then spin
 * down (the reference's rho(n1,n2,n3,2)). The evaluation is CalculatePotentialPlus
 * for numSpin = 2 (CalPotential.f90:120-160): ion and Hartree terms on the
 * total density, the LSDA exchange-correlation (:262-282), Thomas-Fermi,
 * Wang-Teter and vW on 2 rho_s, halved (KEDF_TF.f90:66-74, CalPotential.f90:
 * 400-407, KEDF_VW.f90:100-140). ofd_rho_opt runs the numSpin = 2 NTN solve
 * (dir_type 1) with SpinThetaSearch (RhoLineSearch.f90:447-657) inside the
 * magnetization loop (RhoOptN.f90:336-480), in FP64; ofd_ion_forces sums the
 * channels (CalForce.f90:180). Wang-Teter (kinetic 4) only; pools and
 * ofd_ion_relax take numSpin = 1 contexts. */
typedef struct {
  int nspin;        /* 1 or 2 (input keyword SPIN, ReadInputFile.f90:729-745) */
  int lsda;         /* 1 Perdew-Zunger (LSDAPotPZ, the reference's default), 2 Perdew-Wang 92
                       (LSDAPotPW92); <= 0: 1 (XC_LDA.f90:44) */
  int fixed_mag;    /* 1: one solve at the initial magnetization (FIXMAG); 0: the
                       magnetization loop (RhoOptimizers.f90:85) */
  int nmag;         /* magnetization loop steps at most; <= 0: 100 (RhoOptimizers.f90:97) */
  double tolm;      /* stop when |dE/dm| < tolm (Hartree); <= 0: 5e-6 as a
                       single-precision literal (RhoOptimizers.f90:100) */
} ofd_spin;
int ofd_set_spin(ofd_ctx* ctx, const ofd_spin* spin);
/* after a spin ofd_rho_opt: out3 = final magnetization, N_up, N_down; the
 * number of solves of the magnetization loop */
int ofd_get_magnetization(const ofd_ctx* ctx, double* out3, int* mag_iters);
/* initial channels from a total density and the input MAGMOM (System.f90:
 * 326-329): rho2[s] = (rho - (-1)^s magmom / volume) / 2, s = 1, 2 */
int ofd_spin_split(const ofd_ctx* ctx, const double* rho, double magmom, double* rho2);

/* Ion-electron forces from the B-spline structure factor
 * (IonElectronForcesSpline, IonElectronSpline.f90:219-355, called with
 * FFT(rho) as in CalForce.f90:180). forces: 3 per ion (Ha/bohr), ion order as
 * given to ofd_build_ion_potential. */
int ofd_ion_forces(ofd_ctx* ctx, const double* rho, double* forces);

/* Ion-ion Ewald energy (Hartree) and, when forces is not NULL, forces
 * (Ha/bohr, 3 per ion) of the ions given to ofd_build_ion_potential, with the
 * species' valence charges from their .recpot data: EwaldSetup (Ewald.f90:
 * 208-460; the particle-mesh grid is the density grid), EwaldEnergy
 * (:465-596) with the spline reciprocal sum EwaldRecipSpaceSpline
 * (:993-1147), EwaldForces (:1196-1310) with EwaldRecipSpaceForcesSpline
 * (:1649-1851), at the spline order given to ofd_build_ion_potential. The
 * energy becomes the context's ion-ion energy, eTable(6) of every later
 * evaluation, as RefreshIonTerms sets ionIonEnergy (RefreshIons.f90:196).
 * info5 (may be NULL): eta, real-space cutoff (bohr), reciprocal cutoff
 * (1/bohr), and the two error bounds (Hartree). */
int ofd_ewald(ofd_ctx* ctx, double* energy, double* forces, double* info5);

/* Ion relaxation by conjugate gradients, ConjugateGradient of IonOptCG2
 * (IonOptCG2.f90:60-520) with its LnSrch line search (MathLNsrch.f90:52-296),
 * called as OptimizeIons does with RhoOptimizer = ProjNormRhoMinimization
 * (dir_type from ion->rho_dir_type) and CalculateForces = Ewald + ion-electron spline forces
 * (CalForce.f90:134-194). Every geometry refreshes the ion terms (ionic
 * potential and Ewald energy, RefreshIons.f90:125-238), optimizes the density
 * from the previous one (trashPreviousRho = .FALSE.) and computes forces.
 * rhoR (n1*n2*n3, in/out) and frac_xyz (3 per ion, in/out, wrapped into
 * [0,1)) are updated to the final geometry; energy9 and forces (3 per ion,
 * Ha/bohr) are those of the final geometry (either may be NULL); energies
 * (may be NULL) receives the total energy of every geometry evaluated, up
 * to max_energies of them. */
typedef struct {
  int cg_type;          /* 1 Polak-Ribiere, 2 Hager-Zhang, 3 Dai-Yuan; <= 0: 2 (IonOptimizers.f90:36) */
  double force_cutoff;  /* stop when max |F| falls below it (Ha/bohr); <= 0: 5e-5 (:29) */
  int max_steps;        /* <= 0: 200 (:41) */
  double nn_dist;       /* nearest-ion distance guard (bohr); < 0: 2.0 (Neighbor.f90:33); 0: off */
  int rho_dir_type;     /* density optimizer at every geometry, as the input's METH selects
                           (Optimizer.f90:336-341): 1 NTN, 2 NCG, 3 NBF; <= 0: 1 */
} ofd_ion_params;
typedef struct {
  int steps;            /* CG iterations */
  int geometries;       /* geometries evaluated (density solves) */
  double max_force;     /* at the final geometry */
  double seconds;
} ofd_ion_stats;
int ofd_ion_relax(ofd_ctx* ctx, double* rhoR, double* frac_xyz, const ofd_opt_params* rho_params,
                  const ofd_ion_params* ion, double* energy9, double* forces, double* energies, int max_energies,
                  ofd_ion_stats* stats);

/* ---- input and density files (host only) ---- */
typedef struct { char s[16]; } ofd_name;
typedef struct { char s[256]; } ofd_path;
/* A geometry as ReadGeometry leaves it (ReadIonFile.f90:80-705): cell in
 * bohr (cell[3*j + i] = component i of lattice vector j, the columns of
 * cellReal; LATTICE_CART in Angstrom or LATTICE_ABC), ions grouped by species
 * in order of first appearance (type_start: nspecies + 1 offsets, as
 * ofd_build_ion_potential takes them), fractional positions wrapped into
 * [0,1) (POSITIONS_FRAC, POSITIONS_CART in Angstrom, POSITIONS_BOHR),
 * file_index[k] = the ion's position in the file, and per species its name
 * and SPECIES_POT file. Arrays are allocated by the library. */
typedef struct {
  double cell[9];
  int nion, nspecies;
  int* type_start;
  double* frac;
  int* file_index;
  ofd_name* species;
  ofd_path* pseudo_file;
} ofd_geometry;
int ofd_read_geometry(const char* path, ofd_geometry* out, char* err, int errlen);
void ofd_free_geometry(ofd_geometry* g);
/* A .recpot table as ReadPseudo reads it (ReadIonFile.f90:708-930): maxG in
 * 1/A and the values in eV*A^3, as ofd_build_ion_potential takes them.
 * values NULL: only *npts is set (the count to allocate). */
int ofd_read_recpot(const char* path, double* maxg_invA, double* values, int capacity, int* npts, char* err,
                    int errlen);
/* Density files in the reference's format (27-byte records: four A14,I12
 * header records, then E26.20 values, x fastest): Print3DArray without
 * rescaling (Output.f90:617-873) and ReadDensity (ReadIonFile.f90:1054-1314),
 * which copies a file of the same grid (optionally repeated repeat_x/y/z
 * times) and otherwise interpolates trilinearly, periodically, at the cell
 * centres of the new grid; with cell (bohr) and num_electrons > 0 the result
 * is scaled to hold num_electrons. One spin channel. The two-digit exponent
 * of E26.20 bounds what can be written: magnitudes below 1e-99 are written as
 * zero, non-finite values and magnitudes of 1e99 or more are an error. */
int ofd_write_density(const char* path, int n1, int n2, int n3, int nspin, const double* rho, char* err, int errlen);
int ofd_read_density(const char* path, int n1, int n2, int n3, const double* cell, double num_electrons,
                     int repeat_x, int repeat_y, int repeat_z, double* rho, char* err, int errlen);

/* ---- batched serving (no reference counterpart: many density
 *      optimizations of one grid shape advanced together) ---- */
typedef struct ofd_pool ofd_pool;
/* A pool of up to `capacity` concurrent requests of grid n1*n2*n3. */
int ofd_pool_create(int n1, int n2, int n3, int capacity, int kinetic, int precision, ofd_pool** out);
void ofd_pool_destroy(ofd_pool* pool);
const char* ofd_pool_last_error(const ofd_pool* pool);
/* Admit a request: takes ownership of a context (same grid, ion terms set)
 * and the starting density (rho_init NULL: the uniform density
 * num_electrons / volume); returns the slot id (>= 0), -2 when the pool is
 * full, another negative status on error. */
int ofd_pool_admit(ofd_pool* pool, ofd_ctx* ctx, const double* rho_init, const ofd_opt_params* params);
/* Advance every active request by one evaluation (one batched launch set).
 * Returns the number of requests still active. */
int ofd_pool_step(ofd_pool* pool);
/* The same for several pools (different grid shapes) at once: their launches
 * go to their own streams and run concurrently; the host waits once per
 * round for all of them. Returns the number of requests still active in all
 * pools (on error every pool's last_error is set). */
int ofd_pools_step(ofd_pool* const* pools, int npool);
/* Slots whose requests have finished and wait for collection: writes up to
 * max_slots of them and returns how many there are. */
int ofd_pool_finished(const ofd_pool* pool, int* slots, int max_slots);
/* Collect a finished request: returns 1 and fills outputs when slot `slot`
 * has finished, 0 while it is running. The slot is freed on collection; the
 * context is handed back through ctx_out, or destroyed when ctx_out is NULL.
 * With rho_out NULL the final density is not formed (the context's device
 * density is then stale) and the call does not wait on the device. */
int ofd_pool_collect(ofd_pool* pool, int slot, double* rho_out, double* energy9, int* rho_outcome,
                     ofd_opt_stats* stats, ofd_ctx** ctx_out);
/* Device bytes one request of this grid needs (admission by HBM footprint). */
long long ofd_request_bytes(int n1, int n2, int n3, int kinetic, int precision);

/* ---- progress server (no reference counterpart): a host thread that drives
 *      pools asynchronously. Each pool keeps one phase (vector updates, the
 *      batched inner products and evaluations its requests wait on) in flight
 *      on its own stream; when it completes, the thread hands the results to
 *      the requests and launches the pool's next phase at once, so pools of
 *      different grids never wait for each other and the caller's admissions
 *      and collections overlap the device work. Per-request results are
 *      identical to ofd_pool_step's (each request runs the same operations in
 *      the same order). ofd_pool_admit, ofd_pool_finished and
 *      ofd_pool_collect stay usable on a driven pool; ofd_pool_step and
 *      ofd_pools_step refuse it. ---- */
typedef struct ofd_server ofd_server;
int ofd_server_create(ofd_server** out);
/* From now on the server drives `pool` (a pool belongs to at most one
 * server; ofd_pool_destroy takes it out of its server first). */
int ofd_server_add_pool(ofd_server* server, ofd_pool* pool);
/* Blocks until a request of any of the server's pools has finished since the
 * previous call, or `timeout_s` has passed. Returns the number of requests
 * finished since the previous call (0 on timeout), negative on a device or
 * launch error (message from ofd_server_last_error; the failing pool stops). */
int ofd_server_wait(ofd_server* server, double timeout_s);
const char* ofd_server_last_error(const ofd_server* server);
/* Stops the thread after completing the phases in flight; the pools remain
 * with the caller and can be stepped or destroyed. */
void ofd_server_destroy(ofd_server* server);

/* ---- autotuner (no reference counterpart: FFTW chooses the reference's
 *      plans inside FFTW_PLAN_DFT_R2C_3D, Fourier_NEW.f90) ---- */
typedef struct ofd_tune_result {
  int rows_x;         /* rows per workgroup of the generic x passes (0 = derived) */
  int kb_y, kb_z;     /* columns per workgroup of the y and z passes (0 = derived) */
  int plane;          /* 1 when the fused x+y forward plane pass is used */
  int configs_tried;
  double us_default;  /* one evaluation with the derived tiles, microseconds */
  double us_best;     /* one evaluation with the chosen tiles */
} ofd_tune_result;
/* Times `reps` evaluations per candidate tile shape on the context's stream
 * (coordinate descent over plane pass, z tiles, y tiles, x rows), keeps the
 * fastest in the context and records it for every context created later
 * with the same grid, functional kind and precision. */
int ofd_autotune(ofd_ctx* ctx, int reps, ofd_tune_result* out);

#ifdef __cplusplus
}
#endif
#endif /* PROFESS_AMD_OFD_H */
