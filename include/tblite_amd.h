/*
 * tblite_amd.h -- C ABI of the MI355X GFN2-xTB singlepoint engine.
 *
 * The engine replaces the reference's device entry point
 *
 *   __device__ void xtb_singlepoint(const structure_type& mol,
 *       const xtb_calculator& calc, wavefunction_type& wfn, const float accuracy,
 *       float energy, float (&gradient)[MAX_NAT][3], float (&sigma)[3][3],
 *       const int verbosity)                                   (src/xtb.cu:188-197)
 *
 * and the data types it takes: structure_type (src/structure.h:5-15),
 * xtb_calculator (src/xtb.h:105-126) and wavefunction_type
 * (src/wavefunction/type.h:81-99).  Those are fixed-size FP32 PODs living on one
 * GPU thread; here they are opaque library-owned handles holding FP64 data of
 * any size, and the call is made from the host.  The wavefunction keeps
 * tblite's array layout (Fortran column-major, first index fastest):
 *   qsh(nsh), qat(nat), dpat(3,nat), qpat(6,nat), emo(nao), focc(nao),
 *   density(nao,nao), coeff(nao,nao).
 * Coordinates are bohr, xyz(3,nat) with the three components of an atom
 * contiguous, as in tblite's mol%xyz.
 *
 * All entry points return TBLITE_AMD_OK (0) or a negative error code, except
 * constructors, which return NULL on failure; the message of the last failure
 * on a context is available from tblite_amd_get_error().  No C++ exception
 * crosses this boundary.  A context owns one HIP stream on one device; handles
 * made from a context must not be used from several host threads at once.
 */
#ifndef TBLITE_AMD_H
#define TBLITE_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- codes */
#define TBLITE_AMD_OK 0
#define TBLITE_AMD_ERR_ARGUMENT -1      /* invalid argument (null, size, element) */
#define TBLITE_AMD_ERR_HIP -2           /* HIP runtime failure / no device / no kernel */
#define TBLITE_AMD_ERR_NOT_CONVERGED -3 /* SCF reached max_iter */
#define TBLITE_AMD_ERR_CHOLESKY -4      /* overlap matrix not positive definite */
#define TBLITE_AMD_ERR_EIGEN -5         /* eigensolver did not converge */
#define TBLITE_AMD_ERR_NAN -6           /* non-finite number in the SCF */
#define TBLITE_AMD_ERR_MIXER -7         /* Broyden update singular */
#define TBLITE_AMD_ERR_TOO_LARGE -8     /* molecule beyond the size the engine supports */

/* per-molecule status bits written by the batched entry points */
#define TBLITE_AMD_STATUS_CONVERGED 0
#define TBLITE_AMD_STATUS_NOT_CONVERGED 1
#define TBLITE_AMD_STATUS_CHOLESKY 2
#define TBLITE_AMD_STATUS_EIGEN 4
#define TBLITE_AMD_STATUS_NAN 8
#define TBLITE_AMD_STATUS_MIXER 16
#define TBLITE_AMD_STATUS_TOO_LARGE 32
#define TBLITE_AMD_STATUS_FAILED 64     /* not run: its batch could not be launched (server) */

/* ------------------------------------------------------- parameter data */
#define TBLITE_AMD_MAX_SHELL 3
#define TBLITE_AMD_MAX_ELEM 16

/* Raw GFN2 element record: the per-element entries of tblite's parameter file
 * (the reference only declares the containers built from them:
 * tb_hamiltonian src/h0.h:8-26, tb_repulsion src/xtb.h:70-90, tb_coulomb and
 * damped_multipole src/coulomb.h:98-223, gfn_ncoord_type src/coulomb.h:90-95). */
typedef struct {
  int z;                                  /* atomic number */
  int nsh;                                /* number of shells */
  int ang[TBLITE_AMD_MAX_SHELL];          /* angular momentum of each shell */
  int pqn[TBLITE_AMD_MAX_SHELL];          /* principal quantum number */
  int ngauss[TBLITE_AMD_MAX_SHELL];       /* STO-nG expansion length */
  double levels[TBLITE_AMD_MAX_SHELL];    /* atomic levels / eV */
  double slater[TBLITE_AMD_MAX_SHELL];    /* Slater exponents / bohr^-1 */
  double refocc[TBLITE_AMD_MAX_SHELL];    /* reference occupations */
  double shpoly[TBLITE_AMD_MAX_SHELL];    /* polynomial enhancement of H0 */
  double kcn[TBLITE_AMD_MAX_SHELL];       /* CN shift of the levels / eV */
  double lgam[TBLITE_AMD_MAX_SHELL];      /* shell scaling of the hardness */
  double gam;                             /* atomic hardness / Eh */
  double gam3;                            /* third-order hubbard derivative / Eh */
  double zeff;                            /* repulsion effective charge */
  double arep;                            /* repulsion exponent */
  double en;                              /* Pauling electronegativity */
  double dkernel;                         /* on-site dipole XC kernel */
  double qkernel;                         /* on-site quadrupole XC kernel */
  double mprad;                           /* multipole damping base radius / bohr */
  double mpvcn;                           /* multipole valence CN */
  double rcov;                            /* covalent radius for the CN / bohr */
  double atomic_rad;                      /* atomic radius for shpoly / bohr */
} tblite_amd_element;

/* ------------------------------------------------- D4 dispersion data
 * Self-consistent D4 dispersion of GFN2 (tblite's d4_dispersion: the hooks
 * dispersion%get_potential / get_energy of next_scf, src/scf/iterators.cu:353-354
 * and :415-417, and dispersion_get_engrad of xtb_singlepoint,
 * src/xtb.cu:274-277, all left as comments in the reference).
 *
 * Reference polarisabilities alpha(i w) are tabulated on the fixed 23-point
 * imaginary-frequency grid of the D4 model (w = 1e-6, 0.05, 0.1, 0.2, ..., 1.0,
 * 1.2, 1.4, 1.6, 1.8, 2.0, 2.5, 3.0, 4.0, 5.0, 7.5, 10.0 Eh), integrated with
 * trapezoidal weights (Casimir-Polder).  Element records follow the order of
 * tblite_amd_param.elem. */
#define TBLITE_AMD_D4_MAXREF 7
#define TBLITE_AMD_D4_NFREQ 23

typedef struct {
  int nref;                                                 /* reference systems (1..7) */
  double cn[TBLITE_AMD_D4_MAXREF];                          /* reference coordination numbers */
  double q[TBLITE_AMD_D4_MAXREF];                           /* reference atomic charges */
  double alpha[TBLITE_AMD_D4_MAXREF][TBLITE_AMD_D4_NFREQ];  /* alpha(i w) / bohr^3 */
  double zeff;                                              /* effective nuclear charge */
  double eta;                                               /* chemical hardness / Eh */
  double r4r2;                                              /* sqrt(0.5 <r^4>/<r^2> sqrt(Z)) */
} tblite_amd_d4_element;

typedef struct {
  int enabled;                  /* 0: no dispersion term */
  double s6, s8, a1, a2;        /* rational (Becke-Johnson) damping of the two-body term */
  double s9, alp;               /* three-body (Axilrod-Teller-Muto) scaling and damping exponent */
  double ga, gc;                /* charge scaling: exp(ga (1 - exp(gc eta (1 - (z+q_ref)/(z+q))))) */
  double wf;                    /* Gaussian weighting of the reference CNs */
  double cn_k4, cn_k5, cn_k6;   /* electronegativity factor of the D4 CN */
  double cn_kn;                 /* steepness of its error-function count */
  double cn_cutoff;             /* real-space cutoffs / bohr: CN, two-body, three-body */
  double disp2_cutoff, disp3_cutoff;
  tblite_amd_d4_element elem[TBLITE_AMD_MAX_ELEM];
} tblite_amd_d4_param;

/* Global GFN2 parameters plus the element table. */
typedef struct {
  double kshell[3][3];       /* H0 shell-pair scaling k_ll' */
  double enscale;            /* electronegativity scaling k_EN */
  double wexp;               /* exponent of the Slater-exponent term */
  double kpol;               /* scaling for polarisation shells */
  double rep_kexp;           /* repulsion distance exponent, heavy pairs */
  double rep_kexp_light;     /* repulsion distance exponent, H/He pairs */
  double rep_rexp;           /* repulsion 1/r exponent */
  double rep_cutoff;         /* repulsion real-space cutoff / bohr */
  double gexp;               /* ES2 kernel exponent */
  double thirdorder_shell[3];/* ES3 shell scaling for s, p, d */
  double dmp3, dmp5;         /* AES2 damping exponents */
  double mp_shift;           /* AES2 radius shift */
  double mp_kexp;            /* AES2 radius steepness */
  double mp_rmax;            /* AES2 maximum radius / bohr */
  double cn_ka, cn_kb;       /* CN counting-function steepness */
  double cn_rshift;          /* CN second counting-function offset / bohr */
  double cn_cutoff;          /* CN real-space cutoff / bohr */
  int nelem;
  tblite_amd_element elem[TBLITE_AMD_MAX_ELEM];
  tblite_amd_d4_param d4;    /* dispersion (element records in the order of elem) */
} tblite_amd_param;

/* Fill *param with the built-in GFN2 data for H, C, N, O and S. */
int tblite_amd_param_default(tblite_amd_param* param);

/* ------------------------------------------------------------- handles */
typedef struct tblite_amd_context_s* tblite_amd_context;
typedef struct tblite_amd_structure_s* tblite_amd_structure;
typedef struct tblite_amd_calculator_s* tblite_amd_calculator;
typedef struct tblite_amd_wavefunction_s* tblite_amd_wavefunction;

/* Context: one device, one stream, last-error message.  device < 0 selects
 * the current device.  No counterpart in the reference (it has no error
 * channel, src/utils/gpu.h:6-14 aborts). */
tblite_amd_context tblite_amd_new_context(int device);
void tblite_amd_delete_context(tblite_amd_context* ctx);
int tblite_amd_get_error(tblite_amd_context ctx, char* buffer, int buffer_size);

/* Structure: replaces structure_type (src/structure.h:5-15). num(nat) atomic
 * numbers, xyz(3,nat) in bohr. */
tblite_amd_structure tblite_amd_new_structure(tblite_amd_context ctx, int nat, const int* num,
                                              const double* xyz, double charge, int uhf);
/* Periodic structure: tblite's tblite_new_structure(error, natoms, numbers,
 * positions, charge, uhf, lattice, periodic) with structure_type's `periodic`
 * flag (src/structure.h:9, never set there).  lattice(3,3) holds the three
 * lattice vectors in bohr, lattice(:, k) = a_k contiguous (tblite's
 * mol%lattice); periodic(3) all nonzero for a three-dimensional cell, all
 * zero (or NULL) for a molecule; other combinations are an error.  Cells are
 * computed at the Gamma point: lattice sums over the images within each
 * cutoff (neighbour list, CN, repulsion, H0 and the multipole integrals, D4
 * two- and three-body) and Ewald sums for ES2 and AES2 (the periodic
 * branches the reference leaves as comments: src/adjlist.cu:26-30,
 * src/h0.cu:359, src/coulomb.cu:94-96, :212-213, :554-557).  Analytic
 * gradients of cells are not available (ERR_ARGUMENT). */
tblite_amd_structure tblite_amd_new_structure_periodic(tblite_amd_context ctx, int nat,
                                                       const int* num, const double* xyz,
                                                       double charge, int uhf,
                                                       const double* lattice,
                                                       const int* periodic);
void tblite_amd_delete_structure(tblite_amd_structure* mol);
/* Atom, shell and orbital counts of a structure under calc's parameter set
 * (calc may be NULL: the built-in GFN2 data). */
int tblite_amd_get_structure_dims(tblite_amd_structure mol, tblite_amd_calculator calc, int* nat,
                                  int* nsh, int* nao);

/* Calculator: replaces xtb_calculator (src/xtb.h:105-126) together with its
 * never-defined constructor new_hamiltonian (src/h0.h:80-86).  param may be
 * NULL for the built-in GFN2 data. Defaults: accuracy 1.0, max_iter 250,
 * mixer damping 0.4, electronic temperature 300 K. */
tblite_amd_calculator tblite_amd_new_gfn2_calculator(tblite_amd_context ctx,
                                                     const tblite_amd_param* param);
void tblite_amd_delete_calculator(tblite_amd_calculator* calc);
int tblite_amd_set_calculator_accuracy(tblite_amd_calculator calc, double accuracy);
int tblite_amd_set_calculator_max_iter(tblite_amd_calculator calc, int max_iter);
int tblite_amd_set_calculator_mixer_damping(tblite_amd_calculator calc, double damping);
int tblite_amd_set_calculator_temperature(tblite_amd_calculator calc, double kelvin);

/* Wavefunction: replaces wavefunction_type (src/wavefunction/type.h:81-99);
 * sized from the structure on first use by a singlepoint. */
tblite_amd_wavefunction tblite_amd_new_wavefunction(tblite_amd_context ctx);
void tblite_amd_delete_wavefunction(tblite_amd_wavefunction* wfn);
/* Copy one tblite wavefunction array to host memory; name is one of
 * "qsh", "qat", "dpat", "qpat", "emo", "focc", "density", "coeff", "n0sh",
 * "n0at".  *count receives the number of doubles; out may be NULL to query. */
int tblite_amd_get_wavefunction_array(tblite_amd_wavefunction wfn, const char* name, double* out,
                                      int* count);

/* Singlepoint: replaces xtb_singlepoint (src/xtb.cu:188-434).  Energy in Eh;
 * tblite_amd_xtb_singlepoint_gradient (below) is the form that also fills
 * gradient and sigma.  accuracy <= 0 uses the calculator's accuracy.
 * iterations may be NULL. */
int tblite_amd_xtb_singlepoint(tblite_amd_context ctx, tblite_amd_structure mol,
                               tblite_amd_calculator calc, tblite_amd_wavefunction wfn,
                               double accuracy, double* energy, int* iterations, int verbosity);

/* Batched singlepoint over nmol independent molecules in one pass of the
 * engine: the data-parallel form of xtb_singlepoint that the reference states
 * as its goal (README.md:21) and runs as <<<1,1>>> (src/xtb.cu:495-497).
 * wfns may be NULL, or an array of nmol handles (entries may be NULL).
 * energies, iterations and status are host arrays of length nmol (iterations
 * and status may be NULL).  Returns OK when every molecule converged,
 * TBLITE_AMD_ERR_NOT_CONVERGED when any status is non-zero. */
int tblite_amd_xtb_singlepoint_batch(tblite_amd_context ctx, int nmol,
                                     const tblite_amd_structure* mols,
                                     tblite_amd_calculator calc, tblite_amd_wavefunction* wfns,
                                     double accuracy, double* energies, int* iterations,
                                     int* status);

/* Flat batched singlepoint: molecules given as one atom list.  mol_nat(nmol)
 * atom counts; num(total) atomic numbers; xyz(3,total) bohr; charge(nmol) and
 * uhf(nmol) may be NULL (neutral, closed shell).  All pointers are host
 * pointers.  The engine keeps its device workspace between calls and grows it
 * only when a batch needs more. */
int tblite_amd_singlepoint_flat(tblite_amd_context ctx, tblite_amd_calculator calc, int nmol,
                                const int* mol_nat, const int* num, const double* xyz,
                                const double* charge, const int* uhf, double accuracy,
                                double* energies, int* iterations, int* status);

/* Flat batch of periodic cells: as tblite_amd_singlepoint_flat, with
 * lattice(9, nmol) the lattice vectors of every cell (layout of
 * tblite_amd_new_structure_periodic).  Cells are batched like molecules. */
int tblite_amd_singlepoint_flat_periodic(tblite_amd_context ctx, tblite_amd_calculator calc,
                                         int nmol, const int* mol_nat, const int* num,
                                         const double* xyz, const double* lattice,
                                         const double* charge, const int* uhf, double accuracy,
                                         double* energies, int* iterations, int* status);

/* Analytic gradients.  The reference's xtb_singlepoint takes gradient(3,nat)
 * and sigma(3,3) arguments (src/xtb.cu:188-197) and sketches the FP64
 * gradient integrals (multipole_grad_cgto, src/multipole.cu:846-981), the
 * repulsion derivative (src/xtb.cu:118-121) and the CN derivative
 * (src/coulomb.cu:289-292) without filling them.  These entry points return
 * dE/dR of the converged SCF energy (Eh/bohr; the free energy is stationary
 * in the orbitals and occupations, so the derivative needs no orbital
 * response) and the strain derivative sigma = sum_a g_a (x) R_a (3 x 3,
 * column-major, Eh).
 *
 * singlepoint_flat_gradient: tblite_amd_singlepoint_flat plus gradients
 * (3, total atoms, in the atom order of xyz; required) and sigmas (9, nmol;
 * may be NULL).  A molecule with a non-zero status still gets the gradient
 * of its last SCF iteration.
 * xtb_singlepoint_gradient: tblite_amd_xtb_singlepoint plus gradient(3,nat)
 * (required) and sigma(3,3) (may be NULL).
 * batch_get_gradient: gradient and sigma of molecule imol of the last
 * gradient call on the context (either may be NULL). */
int tblite_amd_singlepoint_flat_gradient(tblite_amd_context ctx, tblite_amd_calculator calc,
                                         int nmol, const int* mol_nat, const int* num,
                                         const double* xyz, const double* charge, const int* uhf,
                                         double accuracy, double* energies, double* gradients,
                                         double* sigmas, int* iterations, int* status);
int tblite_amd_xtb_singlepoint_gradient(tblite_amd_context ctx, tblite_amd_structure mol,
                                        tblite_amd_calculator calc, tblite_amd_wavefunction wfn,
                                        double accuracy, double* energy, double* gradient,
                                        double* sigma, int* iterations, int verbosity);
int tblite_amd_batch_get_gradient(tblite_amd_context ctx, int imol, double* gradient,
                                  double* sigma);

/* No-allocation variant of tblite_amd_singlepoint_flat, with the same
 * arguments and results.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  It calls no hipMalloc,
 * hipFree or other allocating routine: the parameter block and the whole
 * batch workspace live in one static device array of
 * TBLITE_AMD_NOALLOC_SCRATCH_BYTES that each device the library runs on
 * holds until the process ends.  A batch whose workspace exceeds it returns
 * TBLITE_AMD_ERR_TOO_LARGE (tblite_amd_workspace_estimate gives the size
 * beforehand; at 250 SCF iterations the Broyden history takes about 3 MB per
 * molecule), as does a molecule above the one-workgroup size (padded AO
 * dimension above grid_min_ld).  Every call on a device shares that array
 * and leaves its contents behind: calls that overlap on one device, from any
 * context or thread, race, and the caller must serialise them.  The
 * stage-level accessors below read the array until the next call. */
#define TBLITE_AMD_NOALLOC_SCRATCH_BYTES (64u << 20)
int tblite_amd_singlepoint_flat_noalloc(tblite_amd_context ctx, tblite_amd_calculator calc,
                                        int nmol, const int* mol_nat, const int* num,
                                        const double* xyz, const double* charge, const int* uhf,
                                        double accuracy, double* energies, int* iterations,
                                        int* status);

/* Device-resident batches.  prepare() plans a batch from its composition
 * (atom counts, elements, charges, spins) and allocates the device workspace;
 * set_geometry() loads coordinates (host or device pointer, xyz(3,total)
 * bohr); run() executes the whole singlepoint pipeline on the resident
 * coordinates without host traffic; results() copies energies, iteration
 * counts and status words out.  The same prepared batch can be run again
 * with new coordinates (geometry optimisation, dynamics, benchmarking).
 * The flat and batched entry points replace the prepared batch. */
int tblite_amd_batch_prepare(tblite_amd_context ctx, tblite_amd_calculator calc, int nmol,
                             const int* mol_nat, const int* num, const double* charge,
                             const int* uhf, double accuracy);
int tblite_amd_batch_set_geometry(tblite_amd_context ctx, const double* xyz, int on_device);
/* Device address of the resident coordinate buffer of the prepared batch. */
const double* tblite_amd_batch_geometry_device(tblite_amd_context ctx);
int tblite_amd_batch_run(tblite_amd_context ctx);
int tblite_amd_batch_results(tblite_amd_context ctx, double* energies, int* iterations,
                             int* status);
/* The same with analytic gradients: prepare_gradient also reserves the
 * gradient buffers (and the density and MO coefficients they need),
 * run_gradient runs the singlepoint and then the gradient, gradients()
 * copies gradients(3, total atoms) and sigmas(9, nmol) (either may be NULL)
 * out, and gradient_device() gives the device address of the resident
 * gradient array (3 per atom, atom order of the coordinates; valid until the
 * next prepare), so a geometry step can run on the device between runs. */
int tblite_amd_batch_prepare_gradient(tblite_amd_context ctx, tblite_amd_calculator calc,
                                      int nmol, const int* mol_nat, const int* num,
                                      const double* charge, const int* uhf, double accuracy);
int tblite_amd_batch_run_gradient(tblite_amd_context ctx);
int tblite_amd_batch_gradients(tblite_amd_context ctx, double* gradients, double* sigmas);
const double* tblite_amd_batch_gradient_device(tblite_amd_context ctx);

/* ------------------------------------------------------------- serving
 * Continuous batching of independent singlepoint requests on one device.
 * The server owns `slots` internal contexts (one stream + a device workspace
 * of workspace_bytes / slots reserved at creation each).  With two or more
 * slots the first serves only molecules of the LDS-resident sizes (padded AO
 * dimension <= 96), the others take larger molecules first and fill up with
 * small ones.  submit() queues a request (host arrays are copied) and
 * returns its ticket; poll() never blocks: it reports every molecule whose
 * SCF has ended (per molecule, not per batch), collects finished batches,
 * refills every free slot from its queue while the requests' estimated
 * device footprint fits the slot, launches those batches (batches of
 * different slots run concurrently), and copies out up to `capacity`
 * finished results (ticket, energy in Eh, SCF iterations, status bits,
 * latency from submit to collection in ms).  Calls on one server must come
 * from one host thread.  Requests keep their wavefunction and mixer state on
 * the device until their batch ends.  With two or more slots the server
 * also has a multi-workgroup lane (with four or more slots the last slot
 * becomes the lane, so the server never runs more streams than the
 * device's four hardware queues per process; with two or three it is an
 * extra context with half a slot's share of the workspace): a large request for which the tuning's cost model puts the multi-workgroup
 * path at least twice ahead of one workgroup waits in the lane's queue, and
 * the lane runs the largest waiting one alone on that path, its SCF host
 * loop on a worker thread of the server (poll() still never blocks); beyond
 * two waiting, the regular slots take the oldest of them on one workgroup
 * each.  No counterpart in the reference (one molecule per GPU thread,
 * src/xtb.cu:479-497). */
typedef struct tblite_amd_server_s* tblite_amd_server;
tblite_amd_server tblite_amd_new_server(tblite_amd_context ctx, tblite_amd_calculator calc,
                                        int slots, double workspace_bytes);
void tblite_amd_delete_server(tblite_amd_server* srv);
int tblite_amd_server_submit(tblite_amd_server srv, int nat, const int* num, const double* xyz,
                             double charge, int uhf, long long* ticket);
/* Returns the number of results written (>= 0) or < 0 on error. */
int tblite_amd_server_poll(tblite_amd_server srv, int capacity, long long* tickets,
                           double* energies, int* iterations, int* status, double* latency_ms);
/* grid_requests: requests run alone on the multi-workgroup path so far. */
int tblite_amd_server_stats(tblite_amd_server srv, long long* submitted, long long* completed,
                            int* queued, int* busy_slots, long long* batches,
                            long long* grid_requests);

/* Workspace accounting of the serving admission, on the host (no device):
 * need_bytes receives the device workspace the engine allocates for one batch
 * of these nmol neutral molecules, admit_bytes the footprint the server's
 * admission counts for the same requests.  A slot admits a batch while
 * admit_bytes fits its share of the workspace; admit_bytes >= need_bytes
 * keeps the engine from growing (freeing and reallocating) a slot's reserved
 * workspace.  grid_min_ld as in tblite_amd_tuning; param may be NULL.  Engine
 * instrumentation; no counterpart in the reference. */
int tblite_amd_workspace_estimate(const tblite_amd_param* param, int nmol, const int* mol_nat,
                                  const int* num, int grid_min_ld, double* need_bytes,
                                  double* admit_bytes);

/* ------------------------------------------------- stage-level access
 * Intermediate results of the last flat/batched call, for stage-by-stage
 * parity checks.  All copy to host memory; imol indexes the last batch.
 * Each names the reference routine whose output it returns. */

/* Atom/shell/AO counts of molecule imol. */
int tblite_amd_batch_dims(tblite_amd_context ctx, int imol, int* nat, int* nsh, int* nao);
/* Coordination numbers, ncoord_get_cn (src/coulomb.cu:334-347): cn(nat). */
int tblite_amd_batch_get_cn(tblite_amd_context ctx, int imol, double* cn);
/* Neighbour list, generate (src/adjlist.cu:3-75): inl(nat), nnl(nat) and
 * nlat(nnz) (0-based atom index); returns nnz, or < 0 on error.  nlat may be
 * NULL to query nnz. */
int tblite_amd_batch_get_adjlist(tblite_amd_context ctx, int imol, int* inl, int* nnl, int* nlat,
                                 int capacity);
/* Lattice images of cell imol (generate's trans(:, itr), src/adjlist.cu:26-30
 * text): the translation index of every neighbour-list entry, nltr(nnz), and
 * the translation table trans(3, ntr); either may be NULL.  Returns ntr (0
 * for a molecule) or < 0. */
int tblite_amd_batch_get_images(tblite_amd_context ctx, int imol, int* nltr, int nltr_capacity,
                                double* trans, int trans_capacity);
/* Integrals, get_hamiltonian (src/h0.cu:292-502): overlap(nao,nao),
 * dpint(3,nao,nao), qpint(6,nao,nao), hamiltonian(nao,nao) in tblite's
 * layout; any pointer may be NULL. */
int tblite_amd_batch_get_integrals(tblite_amd_context ctx, int imol, double* overlap,
                                   double* dpint, double* qpint, double* hamiltonian);
/* Self energies, get_selfenergy (src/h0.cu:125-199): selfenergy(nsh). */
int tblite_amd_batch_get_selfenergy(tblite_amd_context ctx, int imol, double* selfenergy);
/* Charges and multipoles after the last SCF step: qsh(nsh), qat(nat),
 * dpat(3,nat), qpat(6,nat), emo(nao), focc(nao); any pointer may be NULL. */
int tblite_amd_batch_get_charges(tblite_amd_context ctx, int imol, double* qsh, double* qat,
                                 double* dpat, double* qpat, double* emo, double* focc);
/* Energy decomposition: erep, eelec (band + entropy + ES2 + AES2 + ES3 + D4
 * two-body); energy = erep + edisp3 + eelec. */
int tblite_amd_batch_get_energy_parts(tblite_amd_context ctx, int imol, double* erep,
                                      double* eelec);
/* Dispersion: D4 two-body energy at the final charges (part of eelec; the
 * dispersion%get_energy hook of next_scf, src/scf/iterators.cu:415-417) and
 * the three-body energy at zero charges (dispersion_get_engrad,
 * src/xtb.cu:274-277); both zero when the parameter set disables D4. */
int tblite_amd_batch_get_dispersion(tblite_amd_context ctx, int imol, double* edisp2,
                                    double* edisp3);

/* Milliseconds spent in each engine phase during the last batched call:
 * times[0] setup (CN, repulsion, Coulomb matrices), [1] neighbour list,
 * [2] integrals/H0, [3] SCF, [4] total.  n = number of entries wanted. */
int tblite_amd_last_timings(tblite_amd_context ctx, double* times, int n);

/* ----------------------------------------------------------------- I/O
 * Host-side helpers around the singlepoint (no device needed).  The
 * reference has no I/O (its driver builds molecules in code, src/main.cu);
 * the formats are the ones tblite reads.
 *
 * read_structure_file: an xyz file (Angstrom; element symbols or numbers) or
 * a Turbomole coord file ($coord in bohr, optional $eht charge= unpaired=).
 * format "xyz", "coord", or NULL to choose by name (*.xyz, else coord).
 * Writes *nat always; fills num[nat] and xyz[3 nat] (bohr) when nat <=
 * capacity (call with capacity 0 to size the buffers).  TBLITE_AMD_ERR_ARGUMENT
 * for an unreadable file or an unparsable line.
 *
 * write_json: one singlepoint result as a JSON object (energy, SCF
 * iterations, numbers, positions, charges, atomic dipoles and quadrupoles,
 * orbital energies and occupations, molecular dipole); null arrays are
 * left out. */
int tblite_amd_read_structure_file(const char* path, const char* format, int capacity, int* nat,
                                   int* num, double* xyz, double* charge, int* uhf);
int tblite_amd_write_json(const char* path, int nat, const int* num, const double* xyz,
                          double energy, int iterations, const double* qat, const double* dpat,
                          const double* qpat, int nao, const double* emo, const double* focc);

/* ------------------------------------------------------------- tuning
 * Which SCF path a molecule takes, by its padded AO dimension ld:
 *   ld <= lds_max_ld                 matrices in LDS, one workgroup per molecule
 *   lds_max_ld < ld <= grid_min_ld   matrices in global memory, one workgroup
 *                                    per molecule (throughput: many molecules
 *                                    share the device)
 *   ld > grid_min_ld                 multi-workgroup path, one molecule across
 *                                    the device (latency)
 * With auto_grid set, molecules of the middle range also take the
 * multi-workgroup path when the batch is too small to fill the device and
 * the cost model (one SCF iteration: wg_a * ld^wg_p seconds on one
 * workgroup, grid_b * ld + grid_c seconds across the device) says the batch
 * finishes sooner.  The coefficients and the work tiles below come from
 * measurement on the device: tblite-cuda_amd/autotune.py times both paths,
 * fits the cost model and picks each tile size by timing batches.  A server's
 * slots never use auto_grid.  TBLITE_AMD_LDS_MAX_LD, TBLITE_AMD_GRID_MIN_LD
 * and TBLITE_AMD_GRID_GRAPH=0 override the defaults at context creation.  No
 * counterpart in the reference (one molecule per GPU thread,
 * src/xtb.cu:479-497). */
typedef struct tblite_amd_tuning {
  int lds_max_ld;   /* 16 .. 96, multiple of 16 */
  int grid_min_ld;  /* 16 .. 1024 */
  int auto_grid;    /* nonzero: cost-model routing of the middle range */
  int grid_graph;   /* nonzero: replay the multi-workgroup eigenvalue phase from a HIP graph */
  double wg_a, wg_p, grid_b, grid_c;
  /* work tiles: neighbour-list entries per integral work item (its
   * workgroup size: 32, 64 or 128), log2 of the atom pairs per pair-plane
   * work item (8..22) and of the triples per three-body work item (10..26);
   * nonzero fine_buckets: SCF size buckets in steps of 16 AOs instead of one
   * per kernel instantiation; big_lds_min_kb: minimum LDS request of the
   * global-memory SCF kernel (KB; above half a CU's LDS it runs one molecule
   * per CU instead of two) */
  int integral_chunk, pair_chunk_log2, atm_chunk_log2, fine_buckets, big_lds_min_kb;
} tblite_amd_tuning;
int tblite_amd_get_tuning(tblite_amd_context ctx, tblite_amd_tuning* out);
int tblite_amd_set_tuning(tblite_amd_context ctx, const tblite_amd_tuning* in);
/* SCF path of molecule imol of the last planned batch: 0 LDS, 1 one
 * workgroup in global memory, 2 multi-workgroup; < 0 on error. */
int tblite_amd_molecule_path(tblite_amd_context ctx, int imol);

/* SCF size buckets of the last batched call, in launch order (largest AO
 * dimension first): ms receives the time of each bucket's SCF kernel (HIP
 * events recorded around the launch on the bucket's stream), ldmax its padded
 * AO dimension, nmol its molecule count, nthreads its workgroup size.  Any
 * array may be NULL.  Returns the number of buckets (entries beyond capacity
 * are not written) or < 0.  Engine instrumentation; no counterpart in the
 * reference. */
int tblite_amd_last_bucket_timings(tblite_amd_context ctx, double* ms, int* ldmax, int* nmol,
                                   int* nthreads, int capacity);

/* Per-molecule SCF phase clocks of the last batch (only when the context was
 * created with TBLITE_AMD_PROFILE=1 in the environment): out receives
 * nmol x 22 values, wall-clock ticks at 100 MHz for init, mixing, potential,
 * Fock, transform, Jacobi (global-memory path: MO basis product), occupation,
 * density, Mulliken, energy, then the Jacobi sweep count, the iteration
 * count, the global-memory eigensolver phases tridiagonalisation (remainder),
 * bisection, inverse iteration and back-transformation, the steps of the
 * blocked tridiagonalisation (column update, reflector, symmetric product,
 * panel corrections, panel update) and the cluster re-orthogonalisation.
 * Returns 22 or < 0. */
int tblite_amd_last_profile(tblite_amd_context ctx, long long* out);

#ifdef __cplusplus
}
#endif

#endif /* TBLITE_AMD_H */
