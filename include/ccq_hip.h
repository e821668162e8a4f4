/*
 * ccq_hip.h -- C ABI of the MI355X (gfx950) coupled-cluster amplitude engine.
 *
 * Drop-in boundary for ccq's CC iteration (piecuch-group/ccq).  ccq calls its
 * amplitude sweep and energy through the procedure pointers update_ptr /
 * calculate_energy_ptr (src/proc_pointers.f90:8-32, bound in
 * src/solve_cc.f90:47-54) from jacobi_iter (src/solve_cc.f90:175-269), which
 * also writes the DIIS ring, prints the residuum and extrapolates
 * (src/diis.F90).  Each entry point below replaces one of those calls; the
 * integrals, T1..T4 and all intermediates stay resident on one GPU between
 * calls and only scalars cross the boundary per iteration.
 *
 * Conventions (ccq is built with -fdefault-integer-8, config/gfortran.mk:10):
 *   - every integer is int64_t, every real is double;
 *   - arrays are Fortran column-major exactly as ccq holds them:
 *       f_a, f_b      (1:orbs, 1:orbs)                    sys%ints%f_a/f_b
 *       v_aa/ab/bb    (froz+1:orbs)^4, <pq|rs> ordering    sys%ints%v_* (src/integrals.f90:280-290)
 *       t_vec         cc%t_vec, blocks T1A,T1B,T2A,T2B,T2C,T3A,T3B,T3C,T3D at
 *                     cc%pos(1..9) (src/utils/cc_utils.f90:479-559)
 *   - every call returns 0 on success and non-zero on error; the message is
 *     available from ccq_hip_last_error (the Fortran side maps it to stop_all,
 *     src/errors.f90:11-27).
 * A context is not thread-safe; it owns one HIP stream on its device.
 */
#ifndef CCQ_HIP_H
#define CCQ_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ccq_hip_ctx ccq_hip_ctx;

/* sys_t sizes (src/system.f90:60-106; src/parser.f90:476-504). */
typedef struct {
  int64_t froz;        /* n0: frozen spatial orbitals              */
  int64_t occ_a;       /* n1: occupied alpha (incl. frozen)        */
  int64_t occ_b;       /* n2: occupied beta  (incl. frozen)        */
  int64_t orbs;        /* n3: spatial orbitals                     */
} ccq_hip_dims;

/* run_t options used by the iteration (src/system.f90:108-170). */
typedef struct {
  int64_t level;       /* 2 CCSD, 3 CCSDT, 4 CCSDTQ (run%lvl_t, run%lvl_q),
                          5 CCSDt: active-space triples (calc_type=CCSDt,
                          run%act_ind_t = 1, src/parser.f90:380-390),
                          6 CC(t;3): CCSDt plus the H-bar / left-CCSD /
                          crcc23 tail (run%hbar, run%lcc, run%mm_23,
                          src/calc_driver.f90:53-72)                       */
  int64_t rhf;         /* run%rhf: closed-shell copies T2C=T2A, T1B=T1A ...
                          (ignored by CCSDt, as update_clusters_t3_opt does) */
  double shift;        /* run%shift: denominator shift                     */
  int64_t diis_space;  /* run%diis_space (<= 12)                           */
  int64_t device;      /* HIP device ordinal                               */
  int64_t act_occ;     /* sys%act_occ, sys%act_unocc (CCSDt only; -1 unset;
                          src/parser.f90:496-499)                          */
  int64_t act_unocc;
} ccq_hip_opts;

/* Create a context: allocate the resident tensors and plan the program.
 * Replaces the set-up part of solve_cc (src/solve_cc.f90:12-99). */
int ccq_hip_create(const ccq_hip_dims* dims, const ccq_hip_opts* opts, ccq_hip_ctx** out);

/* Upload Fock and antisymmetrised integrals in ccq's layouts
 * (sys%ints%f_a, f_b, v_aa, v_ab, v_bb after load_ints, src/integrals.f90:212-298). */
int ccq_hip_set_ints(ccq_hip_ctx* ctx, const double* f_a, const double* f_b,
                     const double* v_aa, const double* v_ab, const double* v_bb);

/* Number of doubles in ccq's t_vec: cc%t_size = pos(10)-1 (get_t_sizes,
 * src/utils/cc_utils.f90:479-559), or pos(16)-1 with the ten P-space T3
 * blocks of CCSDt (get_t_sizes_act, :560-631). */
/* Orbital irreps (ccq's orb_syms from read_sym, src/symmetry.f90:84-114;
 * 1-based, products of the abelian groups C1 .. D2h by XOR of irrep - 1),
 * one per orbital 1..orbs, before ccq_hip_set_ints.  ccq reads them but
 * computes dense blocks; here every product is split into its irrep blocks
 * (run as one grouped launch) and symmetry-forbidden elements are never
 * computed; the results equal the dense ones.  No call: dense (C1). */
int ccq_hip_set_irreps(ccq_hip_ctx* ctx, const int64_t* irreps, int64_t n);
int64_t ccq_hip_t_size(const ccq_hip_ctx* ctx);

/* Copy cc%t_vec to / from the device (T1..T3; T4 stays resident). */
int ccq_hip_set_t(ccq_hip_ctx* ctx, const double* t_vec, int64_t t_size);
int ccq_hip_get_t(ccq_hip_ctx* ctx, double* t_vec, int64_t t_size);

/* T4 (CCSDTQ) stays resident on the device; ccq keeps it in the files
 * <bin_file>_4{a..e}.bin (src/utils/cc_utils.f90:89-145).  Blocks
 * T4A..T4E (aaaa, aaab, aabb, abbb, bbbb) concatenated, each C order
 * occupied-first.  Size 0 unless level == 4. */
int64_t ccq_hip_t4_size(const ccq_hip_ctx* ctx);
int ccq_hip_set_t4(ccq_hip_ctx* ctx, const double* t4, int64_t n);
int ccq_hip_get_t4(ccq_hip_ctx* ctx, double* t4, int64_t n);

/* One in-place amplitude sweep in ccq's order: replaces update_ptr =>
 * update_clusters_t4 (src/update_cc.F90:12-323). */
int ccq_hip_update(ccq_hip_ctx* ctx);

/* Correlation energy: replaces calculate_energy_ptr =>
 * calculate_unsorted_energy (src/energy.f90:138-270). */
int ccq_hip_energy(ccq_hip_ctx* ctx, double* e_cor);

/* DIIS ring record of the current vector: write_vecs (src/diis.F90:289-349). */
int ccq_hip_diis_record(ccq_hip_ctx* ctx, int64_t iter);

/* Printed residuum sqrt(sqrt(|t_iter - t_iter-1|^2)): residuum (src/diis.F90:86-157). */
int ccq_hip_residuum(ccq_hip_ctx* ctx, int64_t iter, double* res);

/* DIIS extrapolation over the ring: calc_diis (src/diis.F90:11-84). */
int ccq_hip_diis_extrapolate(ccq_hip_ctx* ctx);

/* ---- integral input (host code; no GPU needed) ----
 * Fortran column-major arrays with ccq's shapes; files are memory-mapped
 * and parsed in parallel, with ccq's read semantics kept (see each call). */

/* Orbital count from the number of records of a onebody file, n(n+1)/2:
 * count_file_lines + load_ints (src/utils/utils.F90:622-649,
 * src/integrals.f90:248-251). */
int ccq_hip_onebody_orbitals(const char* path, int64_t* orbs);
/* load_e1int (src/integrals.f90:11-41): e1int(orbs, orbs), symmetric. */
int ccq_hip_read_onebody(const char* path, int64_t orbs, double* e1int);
/* load_e2int (src/integrals.f90:43-91): e2int(orbs,orbs,orbs,orbs) =
 * value of line "i a j b" at (i,j,a,b), never symmetrised; the line
 * "0 0 0 0 E" gives en_repul and ends the read.  e2int must be zeroed by the
 * caller (ccq zeroes it in load_ints, src/integrals.f90:240-244). */
int ccq_hip_read_twobody(const char* path, int64_t orbs, double* e2int, double* en_repul);
/* get_fcidump (src/parser.f90:258-311): NORB, NELEC of an FCIDUMP header. */
int ccq_hip_fcidump_header(const char* path, int64_t* norb, int64_t* nelec);
/* load_fcidump (src/integrals.f90:93-210): orbitals reordered by the
 * energies on the "e i 0 0 0" lines (stable rank), 8-fold e2int fill,
 * symmetric e1int, E_nuc.  e1int / e2int zeroed by the caller. */
int ccq_hip_read_fcidump(const char* path, int64_t orbs, double* e1int, double* e2int,
                         double* en_repul);
/* read_sym (src/symmetry.f90:15-119): point group 1..4 (C1, Cs, C2v, D2h)
 * and one irrep per orbital; a missing file gives C1, all irreps 1. */
int ccq_hip_read_sym(const char* path, int64_t orbs, int64_t* point_group, int64_t* irreps);
const char* ccq_hip_read_last_error(void);

/* ---- ccq's amplitude files: Fortran unformatted sequential, one record of
 * doubles (gfortran record markers) ---- */

/* Length of the DIIS vector: conv%vec_size (src/solve_cc.f90:68-77). */
int64_t ccq_hip_vec_size(const ccq_hip_ctx* ctx);
/* write_vecs (src/diis.F90:289-349): the DIIS vector (the first vec_size
 * values of t_vec) to <dir>/t_vec.bin and to the ring slot
 * <dir>/cc_iter_mat-<k>.bin, k = mod(iter, diis_space+1) (0 -> diis_space+1). */
int ccq_hip_write_vecs(ccq_hip_ctx* ctx, int64_t iter, int64_t diis_space, const char* dir);
/* Restart: read a t_vec.bin written by ccq or by ccq_hip_write_vecs into the
 * first vec_size values of t_vec (ccq's restart keyword, src/solve_cc.f90:
 * 56-60, does not read it back; this does). */
int ccq_hip_read_vec(ccq_hip_ctx* ctx, const char* path);
/* T4 (CCSDTQ) to / from <bin_file>_4a.bin .. _4e.bin (open_t4_files,
 * src/utils/cc_utils.f90:89-145), one block per file, ccq's layout. */
int ccq_hip_write_t4(ccq_hip_ctx* ctx, const char* bin_file);
int ccq_hip_read_t4(ccq_hip_ctx* ctx, const char* bin_file);

/* ---- CC(t;3) tail (level 6), after the CCSDt iteration has converged ---- */

/* One- and two-body H-bar from the resident T1, T2: replaces hbar2
 * (src/hbar.f90:10-720) plus add_twobody_hbar (:722-745). */
int ccq_hip_hbar(ccq_hip_ctx* ctx);

/* Start the left-CCSD solve: zero l_vec and fix E_cor for the denominators
 * (the set-up of solve_lcc, src/solve_cc.f90:101-173; l_vec allocated in
 * init_vecs, src/calc_driver.f90:229-233). */
int ccq_hip_left_init(ccq_hip_ctx* ctx, double e_cor);

/* One left-CCSD Jacobi step: replaces update_ptr => update_l2
 * (src/update_lcc.f90:10-104, left_jacobi_update :208-280). */
int ccq_hip_left_update(ccq_hip_ctx* ctx);

/* |LH| + E_cor of the last step: replaces calculate_energy_ptr =>
 * calculate_left_energy (src/energy.f90:103-136). */
int ccq_hip_left_energy(ccq_hip_ctx* ctx, double* e);

/* on != 0: ccq_hip_diis_record / residuum / diis_extrapolate act on l_vec
 * (conv%vec_ptr => cc%l_vec, conv%vec_size = cc%l_size in solve_lcc). */
int ccq_hip_left_mode(ccq_hip_ctx* ctx, int64_t on);

/* cc%l_size = pos(6)-1 (T1/T2 layout of t_vec) and copies of cc%l_vec. */
int64_t ccq_hip_l_size(const ccq_hip_ctx* ctx);
int ccq_hip_set_l(ccq_hip_ctx* ctx, const double* l_vec, int64_t l_size);
int ccq_hip_get_l(ccq_hip_ctx* ctx, double* l_vec, int64_t l_size);

/* Triples correction over the triples outside the P space with the four
 * denominators A-D: replaces crcc23 (src/mm_correct.f90:11-343).
 * e23[0..3] = cc%mm_en_cor_a .. _d (CC(t;3) energy = E_cor + e23[3]). */
int ccq_hip_cct3(ccq_hip_ctx* ctx, double* e23);

/* ---- starting guess (ccq starts from T = 0, src/calc_driver.f90:210-211) ----
 * mode 0: T = 0; 1: T2 = MP2 amplitudes <ij||ab> / D_ijab; 2: the learned
 * guess T2 = MP2 (1 + r), r from a 5-64-64-1 BF16 MLP over per-amplitude
 * features (|MP2| / rms, f_ii + f_jj, f_aa + f_bb, <ij||ij>, <ab||ab>).
 * T1, T3, T4 start at zero.  params: ccq_hip_mlp_nparams() floats
 * (mu[5] sd[5] W1[5][64] b1[64] W2[64][64] b2[64] W3[64] b3). */
int ccq_hip_guess(ccq_hip_ctx* ctx, int64_t mode, const float* params, int64_t nparams);
/* Training data from the current (converged) amplitudes: for every T2 tuple
 * of the T2A, T2B, T2C blocks in order, X[n][5] features, M = MP2 / rms and
 * Y = (t - MP2) / rms. */
int64_t ccq_hip_guess_data_size(const ccq_hip_ctx* ctx);
int ccq_hip_guess_data(ccq_hip_ctx* ctx, float* X, float* M, float* Y, int64_t n);
/* GPU training of the guess network (no context; device of the calling
 * thread): `steps` Adam steps on batches of `batch` samples drawn by a hash
 * of (seed, step, slot); loss = mean (M r - Y)^2.  params in / out. */
int64_t ccq_hip_mlp_nparams(void);
int ccq_hip_mlp_train(const float* X, const float* M, const float* Y, int64_t n, float* params,
                      int64_t nparams, int64_t steps, int64_t batch, double lr, uint64_t seed,
                      double* final_loss);
/* Loss and gradient over all n samples (the training kernels, one step, no update). */
int ccq_hip_mlp_grad(const float* X, const float* M, const float* Y, int64_t n,
                     const float* params, int64_t nparams, float* grad, double* loss);
/* ccq_hip_mlp_grad without device allocation.  No counterpart in ccq: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as ccq_hip_mlp_grad; every device buffer
 * is a static per-GPU array of about 9 MB held until the process ends, so
 * n is limited to CCQ_HIP_MLP_NOALLOC_MAX_SAMPLES (larger n returns 2 and
 * leaves the GPU untouched).  Calls that overlap on one GPU race (they share
 * that array). */
#define CCQ_HIP_MLP_NOALLOC_MAX_SAMPLES 8192
int ccq_hip_mlp_grad_noalloc(const float* X, const float* M, const float* Y, int64_t n,
                             const float* params, int64_t nparams, float* grad, double* loss);
const char* ccq_hip_mlp_last_error(void);

/* Arithmetic of the bulk contractions of later ccq_hip_update calls: 0 FP64
 * MFMA (default, ccq's precision); 1 FP32 MFMA for the LDS-tiled
 * contractions (operands rounded to FP32 on the way into LDS, FP32
 * accumulation, added into FP64 results); 2 as 1, with the intermediates,
 * residuals, integral views and a copy of the amplitudes of the sweep held
 * in FP32 (half the bytes per sweep; float shadows allocated at the first
 * such sweep).  The updated amplitudes, denominators, energy, DIIS and the
 * convergence test stay FP64; the caller refines with FP64 sweeps (mode 0)
 * before it stops.  No counterpart in ccq (FP64 only). */
int ccq_hip_set_precision(ccq_hip_ctx* ctx, int64_t mode);

/* Optional per-operation timing with HIP events on the context's stream.
 * ccq has no counterpart (it prints only wall time per iteration). */
int ccq_hip_profile(ccq_hip_ctx* ctx, int64_t enable);
/* Writes one line per op class: "<name> <calls> <total_ms> <flops> <bytes>". */
int ccq_hip_profile_report(ccq_hip_ctx* ctx, char* buf, int64_t buflen);

/* Copy one dense tensor of the loaded program (an intermediate, residual
 * accumulator or amplitude block, by its IR id) to the host.  Inspection
 * only; ccq has no counterpart.  Reads the FP64 storage: after a precision
 * mode 2 sweep, intermediates and residuals hold the last FP64 / mode-1
 * sweep's values (amplitude blocks are always current). */
int ccq_hip_debug_read(ccq_hip_ctx* ctx, const char* tensor_id, double* out, int64_t n);
/* Run only the first nsections sections of one sweep (inspection only). */
int ccq_hip_debug_sweep(ccq_hip_ctx* ctx, int64_t nsections);

/* Block until the context's stream is idle. */
int ccq_hip_sync(ccq_hip_ctx* ctx);
/* Attainable matrix-core rate (no ccq counterpart; the ceiling the contraction
 * kernels are measured against): `blocks` workgroups of 4 waves each issue
 * `iters` x 8 independent MFMAs (kind 0 v_mfma_f64_16x16x4f64, 1
 * v_mfma_f32_16x16x4f32, 2 v_mfma_f32_16x16x32_bf16) on the current device;
 * returns the second launch's time and its flops.  0 or the HIP error code. */
int ccq_hip_mfma_ceiling(int64_t kind, int64_t blocks, int64_t iters, double* seconds, double* flops);

const char* ccq_hip_last_error(const ccq_hip_ctx* ctx);
int ccq_hip_destroy(ccq_hip_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* CCQ_HIP_H */
