/*
 * C-ABI of the MI355X-native IMSRG(2) scalar-commutator / BCH / Magnus path.
 *
 * Drop-in boundary for n-parzuchowski/EOM-IMSRG: every entry point below
 * replaces one reference routine (cited per declaration) with the same
 * argument meaning.  Operators are device buffers in one flat FP64 layout
 * that is the reference's own storage concatenated:
 *
 *   [E0 | fhh (A x A) | fph (P x A) | fpp (P x P) |
 *    for each channel q = 3*(2J) + 2*(Tz+1) + parity:
 *      gam1 (npp x npp) gam2 (npp x nph) gam3 (npp x nhh)
 *      gam4 (nph x nph) gam5 (nhh x nhh) gam6 (nph x nhh) ]
 *
 * every block column-major exactly as the Fortran sq_op holds it
 * (basic_IMSRG.f90:50-76,619-628).  The herm flag (+1 / -1) of each operand
 * travels beside its pointer, as sq_op%herm does.  Cross-coupled matrices
 * (type cc_mat) are device buffers of imsrg_basis_size(IMSRG_SIZE_CC)
 * doubles; the WCC workspace has imsrg_basis_size(IMSRG_SIZE_WCC) doubles.
 *
 * All compute entry points are asynchronous on `stream` (a hipStream_t; 0 is
 * the default stream) unless they return a host value.  They return 0 on
 * success or an IMSRG_E* code; imsrg_last_error() describes the failure.
 * The reference reports errors with STOP; a Fortran shim maps non-zero
 * codes back to the same STOP text.  RES must not alias L or R.
 */
#ifndef IMSRG_HIP_H
#define IMSRG_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct imsrg_basis_s* imsrg_basis;
typedef struct imsrg_flow_s* imsrg_flow;
typedef struct imsrg_tensor_s* imsrg_tensor;
typedef void* imsrg_stream;

#define IMSRG_OK 0
#define IMSRG_EINVAL 1
#define IMSRG_EHIP 2
#define IMSRG_ECLIP 3 /* BCH_EXPAND did not converge ('clipping BCH') */
#define IMSRG_ESTEPS 4 /* magnus_decouple hit the step limit */

/* which kernels execute a commutator term */
#define IMSRG_PATH_GROUPED 0 /* one grouped launch per term over all channels */
#define IMSRG_PATH_PLAIN 1   /* one launch per channel per term */

/* imsrg_basis_size selectors */
#define IMSRG_SIZE_NORB 0
#define IMSRG_SIZE_NHOLE 1
#define IMSRG_SIZE_NPART 2
#define IMSRG_SIZE_JTOTAL_MAX 3
#define IMSRG_SIZE_NBLOCKS 4
#define IMSRG_SIZE_OP 5   /* doubles per operator */
#define IMSRG_SIZE_NCC 6  /* cross-coupled blocks */
#define IMSRG_SIZE_CC 7   /* doubles per cross-coupled matrix set (cc_mat) */
#define IMSRG_SIZE_WCC 8  /* doubles of the WCC workspace */
#define IMSRG_SIZE_NEQ 9  /* doubles of the vectorize stream (H%neq) */
#define IMSRG_SIZE_SPMAT 10 /* doubles of a single-particle block matrix set (full_sp_block_mat) */

const char* imsrg_last_error(void);

/* ---------------------------------------------------------------- basis */
/* read_sp_basis + find_holes + allocate_blocks + allocate_CCMAT + store_6j
 * (basic_IMSRG.f90:192-440,521-685,2227-2332; cross_coupled.f90:61-177):
 * orbit table (n, l, 2j, 2tz, hole flag) in the .sps order.  Builds host
 * tables only; device tables are uploaded on first device use. */
int imsrg_basis_create(int norb, const int* n, const int* l, const int* jj2, const int* tz2, const int* hole,
                       imsrg_basis* out);
int imsrg_basis_destroy(imsrg_basis b);
int64_t imsrg_basis_size(imsrg_basis b, int what);
/* per channel: dims[3q+c] (npp, nph, nhh), goff[6q+g], jtp[3q+{0,1,2}] = (2J, Tz, parity) */
int imsrg_basis_channels(imsrg_basis b, int* dims, int64_t* goff, int* jtp);
/* pair list of class cls (0 pp, 1 ph, 2 hh) of channel q: out[2i], out[2i+1] (0-based orbits) */
int imsrg_basis_pairs(imsrg_basis b, int q, int cls, int* out);
/* cross-coupled layout: rlen, ncol, ccoff, wcoff, jtp[3q+{0,1,2}] = (2J, |Tz|, parity) */
int imsrg_basis_cc(imsrg_basis b, int* rlen, int* ncol, int64_t* ccoff, int64_t* wcoff, int* jtp);
/* host 6j symbol of the basis table (twice-j arguments), for tests */
double imsrg_basis_sixj(imsrg_basis b, int j1, int j2, int j3, int j4, int j5, int j6);
int imsrg_set_path(imsrg_basis b, int path);

/* ------------------------------------------------- operator storage */
/* allocate_blocks / duplicate_sq_op / allocate_CCMAT (basic_IMSRG.f90:521-685,
 * 2797-2850; cross_coupled.f90:61-177): a zeroed device buffer of `count`
 * doubles (an operator: IMSRG_SIZE_OP, a cc_mat: IMSRG_SIZE_CC, the WCC
 * workspace: IMSRG_SIZE_WCC).  Synchronous. */
int imsrg_alloc(int64_t count, double** out);
int imsrg_free(double* p);
/* copy one component of a host sq_op to / from a device operator, so a
 * Fortran caller can mirror op%E0, op%fhh, op%fph, op%fpp and
 * op%mat(q)%gam(g)%X (basic_IMSRG.f90:50-76) without repacking:
 *   q = -1: g = 0 E0 (one double), 1 fhh, 2 fph, 3 fpp;
 *   q >= 0: gam(g+1) of channel q (0-based: Fortran mat(q+1)%gam(g+1)).
 * Column-major, exactly the Fortran array.  Synchronous. */
int imsrg_op_put(imsrg_basis b, double* op, int q, int g, const double* host);
int imsrg_op_get(imsrg_basis b, const double* op, int q, int g, double* host);

/* vectorize, basic_IMSRG.f90:3929-3981: vec (device, IMSRG_SIZE_NEQ
 * doubles) <- E0, fhh, fpp, fph, then per channel gam(1..6) with the square
 * blocks' lower triangles only. */
int imsrg_vectorize(imsrg_basis b, const double* op, double* vec, imsrg_stream stream);
/* repackage, basic_IMSRG.f90:3984-4053: op <- vec; square blocks get
 * X(i,j) = herm * X(j,i) (a diagonal element ends as herm * v, as there). */
int imsrg_repackage(imsrg_basis b, const double* vec, int herm, double* op, imsrg_stream stream);
/* the .bin operator file of write_omega_checkpoint / read_omega_checkpoint
 * (basic_IMSRG.f90:3215-3407): the vectorize stream as raw native doubles
 * (Fortran unformatted stream access, no record markers).  The file name
 * (prefix, _s<s>.bin) is the caller's.  Reading requires exactly
 * IMSRG_SIZE_NEQ doubles (the reference reads the first neq without a
 * check).  Synchronous. */
int imsrg_write_op_bin(imsrg_basis b, const double* op, const char* path, imsrg_stream stream);
int imsrg_read_op_bin(imsrg_basis b, const char* path, int herm, double* op, imsrg_stream stream);

/* ------------------------------------------------ Hamiltonian set-up */
/* The scalars the reference keeps on the Hamiltonian (hospace, htype,
 * lawson_beta, com_hw, Aprot + Aneut) are arguments here.  htype: 1 = T -
 * Tcm + V, 2 = T + Uho + V, 3 = T + V, 4 = T / A, 5 = the Hcm one-body part
 * without (com_hw/hw)^2. */
#define IMSRG_HTYPE_T_MINUS_TCM 1
#define IMSRG_HTYPE_T_PLUS_UHO 2
#define IMSRG_HTYPE_T 3
/* calculate_h0_harm_osc, basic_IMSRG.f90:2093-2191: fhh, fph and fpp of the
 * device operator H set to the harmonic-oscillator one-body part (zero
 * outside the (l, j, tz) blocks); E0 and the two-body part untouched.
 * Synchronous. */
int imsrg_h0_harm_osc(imsrg_basis b, double hw, int htype, double lawson_beta, double com_hw, int mass, double* H,
                      imsrg_stream stream);
/* read_me2j_interaction, interaction_IO.f90:282-557: the isospin-coupled
 * two-body matrix elements of an me2j file (gzip or plain text: one header
 * line, then ten fixed-width values a line) over the file's orbit table
 * (nfile entries in .sps order: proton and neutron entry of each (n, l, j)
 * adjacent, proton first), converted to proton-neutron form with the
 * centre-of-mass (htype 1) and Lawson (lawson_beta, com_hw) two-body terms
 * and written into the two-body blocks of the device operator H in the
 * reference's loop order.  File orbits the basis lacks are skipped (the
 * reference's eMax / lmax cut); elements the file does not hold keep their
 * values.  Synchronous. */
int imsrg_read_me2j(imsrg_basis b, const char* path, int nfile, const int* n, const int* l, const int* jj,
                    const int* tz, double hw, int htype, double lawson_beta, double com_hw, int mass, double* H,
                    imsrg_stream stream);
/* calc_HF, HF_mod.f90:14-105 (no three-body force): H in = the vacuum
 * Hamiltonian (t, V) in the HO basis; H out = normal ordered in the
 * Hartree-Fock basis (E0 = e_HF, one-body = the Fock matrix, two-body
 * transformed by transform_2b_to_HF).  The self-consistent loop (criterion:
 * summed eigenvalue change < 1e-6, at most maxit iterations, else
 * IMSRG_ESTEPS) runs on the host; the two-body transform runs on the device.
 * coefs (host, optional, IMSRG_SIZE_SPMAT doubles): the HF orbits per
 * (tz, l, j) block, column i of each column-major block = HF orbit i.
 * Synchronous. */
int imsrg_hartree_fock(imsrg_basis b, double* H, int maxit, double* ehf, int* iterations, double* coefs,
                       imsrg_stream stream);
/* sqop_to_HF, HF_mod.f90:181-209, for a hermitian scalar operator (herm =
 * 1; the reference's tensor branch is not built): Op (device, the vacuum
 * operator in the HO basis) <- its one-body part rotated to the HF orbits
 * `coefs` (host, IMSRG_SIZE_SPMAT doubles, as imsrg_hartree_fock returns
 * them), its two-body part transformed, then normal ordered.  Synchronous. */
int imsrg_sqop_to_hf(imsrg_basis b, const double* coefs, double* Op, int herm, imsrg_stream stream);

/* ------------------------------------------------ commutator terms */
/* commutator_110, commutators.f90:10-33 (returns the value; synchronises) */
int imsrg_commutator_110(imsrg_basis b, const double* L, int Lherm, const double* R, int Rherm, double* value,
                         imsrg_stream stream);
/* commutator_220, commutators.f90:399-432 (returns the value; synchronises) */
int imsrg_commutator_220(imsrg_basis b, const double* L, int Lherm, const double* R, int Rherm, double* value,
                         imsrg_stream stream);
/* RES%E0 = commutator_110 + commutator_220 written on the device
 * (IMSRG_MAGNUS.f90:210) */
int imsrg_commutator_e0(imsrg_basis b, const double* L, int Lherm, const double* R, int Rherm, double* RES,
                        imsrg_stream stream);
/* commutator_111, commutators.f90:36-80 (assigns RES one-body) */
int imsrg_commutator_111(imsrg_basis b, const double* L, int Lherm, const double* R, int Rherm, double* RES,
                         int RESherm, imsrg_stream stream);
/* commutator_121, commutators.f90:83-280 (accumulates into RES one-body) */
int imsrg_commutator_121(imsrg_basis b, const double* L, int Lherm, const double* R, int Rherm, double* RES,
                         int RESherm, imsrg_stream stream);
/* commutator_122, commutators.f90:283-396 (assigns RES two-body) */
int imsrg_commutator_122(imsrg_basis b, const double* L, int Lherm, const double* R, int Rherm, double* RES,
                         int RESherm, imsrg_stream stream);
/* commutator_222_pp_hh, commutators.f90:581-768 (accumulates; fills w1, w2) */
int imsrg_commutator_222_pp_hh(imsrg_basis b, const double* L, int Lherm, const double* R, int Rherm, double* RES,
                               int RESherm, double* w1, double* w2, imsrg_stream stream);
/* commutator_221, commutators.f90:435-578 (accumulates; after 222_pp_hh) */
int imsrg_commutator_221(imsrg_basis b, const double* L, int Lherm, const double* R, int Rherm, double* RES,
                         int RESherm, const double* w1, const double* w2, imsrg_stream stream);
/* calculate_cross_coupled, cross_coupled.f90:1190-1300 */
int imsrg_calculate_cross_coupled(imsrg_basis b, const double* op, int herm, double* CC, imsrg_stream stream);
/* commutator_222_ph, commutators.f90:771-925 (accumulates).  WCC is the
 * caller's workspace as in the reference; on return each block holds, in
 * its upper triangle (row <= column), Lh*W - Rh*W^T with W = LCC RCC^T,
 * the only combination the inverse Pandya transform reads. */
int imsrg_commutator_222_ph(imsrg_basis b, const double* LCC, int Lherm, const double* RCC, int Rherm,
                            double* RES, int RESherm, double* WCC, imsrg_stream stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): imsrg_commutator_222_ph with the WCC workspace taken from a
 * fixed 64 MiB device array inside the library (8,388,608 doubles per GPU,
 * held until the process ends) instead of a caller buffer.  Allocates
 * nothing.  Requires IMSRG_SIZE_WCC <= 8,388,608 (the 16O bases up to eMax
 * 6), else IMSRG_EINVAL; requires the basis's device tables to exist (any
 * earlier allocating call on the basis uploads them), else IMSRG_EINVAL.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (other streams, other threads) race.  Same results as
 * imsrg_commutator_222_ph. */
int imsrg_commutator_222_ph_noalloc(imsrg_basis b, const double* LCC, int Lherm, const double* RCC, int Rherm,
                                    double* RES, int RESherm, imsrg_stream stream);

/* ------------------------------------------------ operator algebra */
/* add_sq_op, basic_IMSRG.f90:3551-3583: C = ax*A + bx*B (C may alias A or B) */
int imsrg_add_sq_op(imsrg_basis b, const double* A, double ax, const double* B, double bx, double* C,
                    imsrg_stream stream);
/* mat_frob_norm, basic_IMSRG.f90:4056-4087 (returns the value; synchronises) */
int imsrg_mat_frob_norm(imsrg_basis b, const double* op, double* value, imsrg_stream stream);
/* build_gs_white, generators.f90:4-106 (ETA herm = -1) */
int imsrg_build_gs_white(imsrg_basis b, const double* H, double* ETA, imsrg_stream stream);
/* build_gs_atan, generators.f90:163-269: eta = 0.5 atan(2 H / Eden) on fph
 * and the pphh blocks, Eden as in build_gs_white (ETA herm = -1; E0 kept) */
int imsrg_build_gs_atan(imsrg_basis b, const double* H, double* ETA, imsrg_stream stream);
/* build_gs_atanxx, generators.f90:271-374: the same with the denominators
 * of Hold and the numerators of H */
int imsrg_build_gs_atanxx(imsrg_basis b, const double* Hold, const double* H, double* ETA, imsrg_stream stream);
/* mbpt2, basic_IMSRG.f90:4165-4199 (returns the value; synchronises) */
int imsrg_mbpt2(imsrg_basis b, const double* H, double* value, imsrg_stream stream);
/* normal_order, basic_IMSRG.f90:2419-2550 (in place) */
int imsrg_normal_order(imsrg_basis b, double* H, imsrg_stream stream);

/* ------------------------------------------------ step level */
/* Workspace of the nested-commutator drivers (INT1/INT2/AD/w1/w2 and the
 * cross-coupled buffers that BCH_EXPAND and MAGNUS_EXPAND allocate). */
int imsrg_flow_create(imsrg_basis b, imsrg_flow* out);
int imsrg_flow_destroy(imsrg_flow f);
/* capture each nested commutator term as a hipGraph (grouped path only) */
int imsrg_flow_set_graphs(imsrg_flow f, int on);
/* the generator magnus_decouple / magnus_step build after each step (the
 * reference passes it as build_generator, IMSRG_MAGNUS.f90:15;
 * main_IMSRG.f90:230 passes build_gs_w2).  Default white.  The flow
 * gradient (imsrg_flow_forward / _backward / _grad) requires white. */
#define IMSRG_GEN_WHITE 0
#define IMSRG_GEN_ATAN 1
#define IMSRG_GEN_W2 2
int imsrg_flow_set_generator(imsrg_flow f, int generator);
/* build_gs_w2, generators.f90:109-160: eta1 = build_gs_atan(H), H2 =
 * [eta1, H] (E0 = 110 + 220 included), eta = 1.5 eta1 + 0.5
 * build_gs_atanxx(H, H2); uses the flow's workspace (ETA E0 = 0) */
int imsrg_build_gs_w2(imsrg_flow f, const double* H, double* ETA, imsrg_stream stream);
/* BCH_EXPAND, IMSRG_MAGNUS.f90:149-234: HS = exp(G) H exp(-G), terms until
 * ||term|| < conv (reference: 1e-6 as a single-precision literal), at most
 * maxit-1 nested commutators.  nterms = the loop index iw of the exit.
 * Gherm / Hherm are G%herm and H%herm (HS and every term inherit Hherm).
 * advals (host, optional, length maxit) receives ||term||*coef. */
int imsrg_bch_expand(imsrg_flow f, double* HS, const double* G, int Gherm, const double* H, int Hherm,
                     double conv, int maxit, int* nterms, double* advals, imsrg_stream stream);
/* transform_sq_op_BCH, IMSRG_MAGNUS.f90:131-146, scalar operators: Op <-
 * exp(G) Op exp(-G) in place (G antihermitian, Op of hermiticity Opherm),
 * with BCH_EXPAND's terms and exit (nterms as imsrg_bch_expand). */
int imsrg_transform_sq_op_bch(imsrg_flow f, double* Op, int Opherm, const double* G, double conv, int maxit,
                              int* nterms, imsrg_stream stream);
/* dHds, IMSRG_ODE.f90:416-451: DH = [eta(HS), HS] with E0 = 110 + 220, eta
 * the flow's generator (imsrg_flow_set_generator); DH herm = +1.  The
 * ODE solver's vectorize of DH is imsrg_vectorize. */
int imsrg_dhds(imsrg_flow f, const double* HS, double* DH, imsrg_stream stream);
/* MAGNUS_EXPAND, IMSRG_MAGNUS.f90:620-691: DG <- dOmega/ds from eta (in DG)
 * and Omega G.  ncomm = nested commutators evaluated. */
int imsrg_magnus_expand(imsrg_flow f, double* DG, int DGherm, const double* G, int Gherm, double conv,
                        int* ncomm, imsrg_stream stream);
/* euler_step, IMSRG_MAGNUS.f90:754-764: G = G + ds*DG */
int imsrg_euler_step(imsrg_basis b, double* G, const double* DG, double ds, imsrg_stream stream);
/* One iteration of magnus_decouple's loop (IMSRG_MAGNUS.f90:100-125) with the
 * flow's generator (imsrg_flow_set_generator; white by default):
 * MAGNUS_EXPAND, euler_step, BCH_EXPAND, the generator, mbpt2.  out[0..3] = E0, mbpt2, Magnus commutators, BCH loop index. */
int imsrg_magnus_step(imsrg_flow f, double* HS, double* G, double* DG, const double* H, double ds,
                      double bch_conv, double mag_conv, double* out, imsrg_stream stream);
/* magnus_decouple, IMSRG_MAGNUS.f90:15-128 with the flow's generator: HS in = the
 * bare normal-ordered H, out = decoupled H; G out = Omega.  crit_mode 0:
 * loop while |mbpt2| > crit_tol (reference); 1: until |dE0| < crit_tol.
 * log (host, optional) receives 4 doubles per step: s, E0, E0+mbpt2,
 * nested commutators.  *steps = Magnus steps taken. */
int imsrg_magnus_decouple(imsrg_flow f, double* HS, double* G, double ds, double bch_conv, double mag_conv,
                          int crit_mode, double crit_tol, int maxsteps, double* log, int* steps,
                          imsrg_stream stream);

/* magnus_decouple with the reference's Omega checkpoints (IMSRG_MAGNUS.f90:
 * 64-110 with checkpointing on; write_omega_checkpoint /
 * read_omega_checkpoint, basic_IMSRG.f90:3215-3407).  Files are
 * <prefix>_s<s>.bin (the .bin operator file of imsrg_write_op_bin), s as
 * the reference writes it (Fortran F6.3, a '0' appended below 10; see
 * imsrg_omega_checkpoint_name).  On entry the newest file s = 95.5, 95.0,
 * ... (> 1.0) is looked for; if one exists the flow restarts from it: G <-
 * Omega read from the file, HS (in: the bare H) <- BCH_EXPAND(G, H), the
 * generator rebuilt, s = sx, steps = nint(sx / ds), and the first step runs
 * without the mbpt2 test.  Every chk steps (chk <= 0: the reference's 24,
 * 12 at eMax 14, 4 above) Omega is written before the step, except at the
 * first such step after a restart, and the file of s - 12 (eMax 14: s - 4)
 * is removed.  *restart_s = sx, or -1 for a fresh start; *steps = the step
 * count including the restart's nint(sx / ds); log rows (as
 * imsrg_magnus_decouple) for the steps run by this call; maxsteps limits
 * the steps run by this call. */
int imsrg_magnus_decouple_chk(imsrg_flow f, double* HS, double* G, double ds, double bch_conv, double mag_conv,
                              int crit_mode, double crit_tol, int maxsteps, const char* prefix, int chk,
                              double* log, int* steps, double* restart_s, imsrg_stream stream);
/* the file name write_omega_checkpoint gives Omega at flow parameter s
 * (basic_IMSRG.f90:3239-3249) for the stem `prefix` */
int imsrg_omega_checkpoint_name(const char* prefix, double s, char* out, int outlen);

/* ------------------------------------------------ many-instance batches */
/* Instance-array variants of the step-level calls (IMSRG_MAGNUS.f90:15-234,
 * 620-691) for ninst independent problems on one basis, run together: every
 * kernel launch covers all instances, and each instance takes the BCH,
 * Magnus and convergence exits of its own flow (identical results to ninst
 * separate flows).  Operator arguments hold ninst operators back to back
 * (ninst * IMSRG_SIZE_OP doubles, instance k at offset k * IMSRG_SIZE_OP).
 * A batched flow runs the grouped path and accepts only imsrg_batch_* and
 * imsrg_flow_set_graphs / imsrg_flow_destroy. */
int imsrg_batch_create(imsrg_basis b, int ninst, imsrg_flow* out);
int imsrg_batch_size(imsrg_flow f);
/* RES_k = [L_k, R_k] for every k (RES herm = -Lh*Rh; E0 included) */
int imsrg_batch_commutator(imsrg_flow f, const double* L, int Lh, const double* R, int Rh, double* RES,
                           imsrg_stream stream);
/* BCH_EXPAND per instance (G antihermitian, H hermitian); nterms[k] = the
 * reference's iw at instance k's exit */
int imsrg_batch_bch_expand(imsrg_flow f, double* HS, const double* G, const double* H, double conv, int maxit,
                           int* nterms, imsrg_stream stream);
/* MAGNUS_EXPAND per instance: DG holds eta on entry, dOmega on exit */
int imsrg_batch_magnus_expand(imsrg_flow f, double* DG, const double* G, double conv, int* ncomm,
                              imsrg_stream stream);
/* magnus_decouple per instance with build_gs_white and the |mbpt2| <
 * crit_tol criterion: HS in = normal-ordered H, out = flowed H; G (Omega)
 * must be zero on entry; E0[k], steps[k] per instance; nsteps_max the
 * number of batched steps run */
int imsrg_batch_magnus_decouple(imsrg_flow f, double* HS, double* G, double ds, double bch_conv, double mag_conv,
                                double crit_tol, int maxsteps, double* E0, int* steps, int* nsteps_max,
                                imsrg_stream stream);

/* ------------------------------------------------ reverse mode */
/* The reference has no derivatives.  These are hand-written transposes of
 * the routines above (vector-Jacobian products) for dE0/d(input matrix
 * elements) of the Magnus flow.  Gradients are w.r.t. the flat operator
 * arrays and ACCUMULATE into the outputs (Lb, Rb, Gb, Hb, ETAb) unless
 * stated; they are exact for perturbations that keep the operators'
 * herm symmetry and the zero Pauli-forbidden (a,a) odd-J elements, which
 * is every operator of the flow.  The operators passed in must themselves
 * be (anti)hermitian as flagged (stored in both triangles, as every
 * operator of the flow is): the transposes use X^T = herm X on their
 * square blocks.  Asynchronous on `stream` except the
 * drivers that need the reference's data-dependent exits (bch / magnus /
 * flow), which synchronise like their forward counterparts. */

/* transpose of one [L, R] (E0 = 110 + 220, then the BCH term order of
 * IMSRG_MAGNUS.f90:207-219): Lb += d<RESb,[L,R]>/dL, Rb += d<RESb,[L,R]>/dR */
int imsrg_commutator_vjp(imsrg_basis b, const double* L, int Lherm, const double* R, int Rherm, const double* RESb,
                         int RESherm, double* Lb, double* Rb, imsrg_stream stream);
/* transpose of BCH_EXPAND (IMSRG_MAGNUS.f90:149-234) with G antihermitian:
 * Gb += d<HSb,HS>/dG, Hb += d<HSb,HS>/dH; nterms as imsrg_bch_expand */
int imsrg_bch_vjp(imsrg_flow f, const double* G, const double* H, int Hherm, double conv, int maxit,
                  const double* HSb, double* Gb, double* Hb, int* nterms, imsrg_stream stream);
/* transpose of MAGNUS_EXPAND (IMSRG_MAGNUS.f90:620-691), eta and Omega
 * antihermitian: ETAb += d<DGb,dOmega>/d eta, Gb += d<DGb,dOmega>/dOmega */
int imsrg_magnus_vjp(imsrg_flow f, const double* ETA, const double* G, double conv, const double* DGb,
                     double* ETAb, double* Gb, int* ncomm, imsrg_stream stream);
/* transpose of build_gs_white (generators.f90:4-106): Hb += d<ETAb,eta(H)>/dH */
int imsrg_white_vjp(imsrg_basis b, const double* H, const double* ETAb, double* Hb, imsrg_stream stream);
/* transpose of the flow's generator (build_gs_white, build_gs_atan or
 * build_gs_w2, generators.f90:4-160): Hb += d<ETAb,eta(H)>/dH */
int imsrg_generator_vjp(imsrg_flow f, const double* H, const double* ETAb, double* Hb, imsrg_stream stream);
/* transpose of normal_order (basic_IMSRG.f90:2419-2550): Tb = d<Hb,H(T)>/dT
 * (overwritten) */
int imsrg_normal_order_vjp(imsrg_basis b, const double* Hb, double* Tb, imsrg_stream stream);
/* dE0/dT for magnus_decouple (IMSRG_MAGNUS.f90:15-128, the flow's
 * generator, |mbpt2| > crit_tol loop) started from H = normal_order(T), in two sweeps:
 * imsrg_flow_forward runs the flow, returns E0 and the step count and keeps
 * Omega every `ckpt` steps in the flow object; imsrg_flow_backward then
 * writes dE0/dT into Tb (overwritten), recomputing the steps between
 * checkpoints.  ckpt = 0 keeps every step's states and nested-commutator
 * terms instead (device memory of about (3 steps + commutators) operators,
 * nothing recomputed).  imsrg_flow_grad does both. */
int imsrg_flow_forward(imsrg_flow f, const double* T, double ds, double bch_conv, double mag_conv, double crit_tol,
                       int maxsteps, int ckpt, double* E0, int* steps, imsrg_stream stream);
int imsrg_flow_backward(imsrg_flow f, double* Tb, imsrg_stream stream);
int imsrg_flow_grad(imsrg_flow f, const double* T, double ds, double bch_conv, double mag_conv, double crit_tol,
                    int maxsteps, int ckpt, double* Tb, double* E0, int* steps, imsrg_stream stream);

/* ------------------------------------------------ tuning */
/* No counterpart in the reference: selects the MFMA tile shape of one GEMM
 * family (IMSRG_GEMM_*): tile 0 = 64x64, 1 = 128x64, 2 = 128x128, 3 = 64x32,
 * 4 = 32x64, 5 = 32x32 (operands staged through LDS), 6 = 128x64, 7 =
 * 64x64 (register-direct: each wave loads its own 64x32 / 32x32 block's
 * operands, no LDS), plus IMSRG_TILE_XCD to pack each channel's tiles onto
 * one XCD (one L2). */
#define IMSRG_TILE_XCD 8
#define IMSRG_GEMM_PPPP 0   /* 222_pp_hh L1 R1 */
#define IMSRG_GEMM_PPX 1    /* 222_pp_hh pp x (ph|hh) */
#define IMSRG_GEMM_XX 2     /* 222_pp_hh (ph|hh)^T (ph|hh) */
#define IMSRG_GEMM_PH 3     /* 222_ph W' = Lh LCC RCC^T - Rh RCC LCC^T */
#define IMSRG_GEMM_ADJ 4    /* reverse mode: the transposed 222_pp_hh products */
#define IMSRG_GEMM_ADJ_PH 5 /* reverse mode: the transpose of the W' product */
#define IMSRG_GEMM_ADJ_PP 6 /* reverse mode: the fused transposed products that write or stream pp blocks */
int imsrg_set_gemm_tile(imsrg_basis b, int family, int tile);
int imsrg_get_gemm_tile(imsrg_basis b, int family);
/* No counterpart in the reference: on the grouped path, run the full
 * commutator's GEMM families, the Pandya transform of R and the zero- and
 * one-body terms on three auxiliary streams beside the two-body
 * accumulation (on = 1, the default), or all on the caller's stream (0).
 * Results are identical either way: every output element is accumulated in
 * the same order. */
int imsrg_set_term_streams(imsrg_basis b, int on);

/* ------------------------------------------------ instrumentation */
/* No counterpart in the reference: times `iters` launches of one kernel
 * class of [L, R] between hipEvents recorded on `stream` and returns the
 * mean milliseconds per launch.  Inputs as for the term routines; every
 * kernel class runs once untimed first (fills w1/w2, LCC/RCC, WCC).
 * which: IMSRG_K_* below. */
#define IMSRG_K_PPPP 0      /* 222_pp_hh pppp block: L1 R1 MFMA GEMM + symmetric epilogue (grouped: upper tiles, as in the commutator) */
#define IMSRG_K_PPHH_SMALL 1 /* 222_pp_hh pp x (ph|hh) and (ph|hh)^2 GEMMs + epilogue */
#define IMSRG_K_PH_GEMM 2   /* 222_ph WCC = LCC RCC^T (MFMA) */
#define IMSRG_K_PH_SCATTER 3 /* 222_ph inverse Pandya (grouped: upper triangles, as in the commutator) */
#define IMSRG_K_122 4        /* commutator_122 (grouped: upper triangles, as in the commutator) */
#define IMSRG_K_PANDYA 5
#define IMSRG_K_ONEBODY 6   /* 111 + 121 + 221 */
#define IMSRG_K_COMMUTATOR 7 /* the whole [L, R] */
#define IMSRG_K_PPPP_GEMM 8 /* the pppp L1 R1 MFMA GEMM alone */
#define IMSRG_K_COPY 9      /* plain copy of the first half of WCC into its second half (HBM calibration) */
int imsrg_profile_kernel(imsrg_basis b, int which, const double* L, int Lherm, const double* R, int Rherm,
                         double* RES, int RESherm, double* w1, double* w2, double* LCC, double* RCC, double* WCC,
                         int iters, double* ms, imsrg_stream stream);

/* ------------------------------------------------ tensor operators */
/* A tensor operator of rank lambda = rank/2 and parity change dpar/2 (the
 * reference's sq_op with op%rank > 0, both "multiplied by 2") is one flat
 * FP64 device array:
 *
 *   [E0 | fhh (A x A) | fph (P x A) | fpp (P x P) |
 *    for each block q: tgam1 .. tgam9 ]
 *
 * blocks in allocate_tensor's loop order (2J1 = 0 .. 2 jmax, 2J2 =
 * max(|2J1 - rank|, 2J1) .. 2J1 + rank, Tz = -1..1, P1 = 0..1; P2 = P1, or
 * 1 - P1 when dpar/2 is odd), tgam g = (pairs of class sea1(g) of channel
 * (J1, Tz, P1)) x (pairs of class sea2(g) of channel (J2, Tz, P2)) with
 * sea1 = (pp,pp,pp,ph,hh,ph,hh,ph,hh), sea2 = (pp,ph,hh,ph,hh,hh,pp,pp,ph),
 * column-major (basic_IMSRG.f90:688-951, tblck(q)%tgam(g)%X).  Blocks with
 * 2J2 > 2 jmax are empty.  fhh and fpp hold both triangles; the hole-particle
 * one-body elements are herm (-1)^((ja-jb)/2) times the transposed fph. */
#define IMSRG_TSIZE_OP 0      /* doubles per tensor operator */
#define IMSRG_TSIZE_NBLOCKS 1 /* tensor blocks */
#define IMSRG_TSIZE_NCC 2     /* tensor cross-coupled blocks (pandya_mat) */
#define IMSRG_TSIZE_CC 3      /* doubles of a tensor cross-coupled set: CCX and CCR of every block */
#define IMSRG_TSIZE_W 4       /* doubles of one 222_ph product set (Wx or Wy of every block) */
/* allocate_tensor + allocate_small_tensor_CCMAT + store_6j_3halfint
 * (basic_IMSRG.f90:688-951,2335-2411; cross_coupled.f90:498-700): rank and
 * dpar twice lambda and twice the parity change (STOP 'RANK MUST BE
 * MULTIPLIED BY 2' becomes IMSRG_EINVAL).  Host tables only; device tables
 * on first use. */
int imsrg_tensor_create(imsrg_basis b, int rank, int dpar, imsrg_tensor* out);
int imsrg_tensor_destroy(imsrg_tensor t);
int64_t imsrg_tensor_size(imsrg_tensor t, int what);
/* per block q: info[7q..] = 2J1, 2J2, Tz, P1, P2, channel of J1, channel of
 * J2 (-1: empty); dims[6q..] = npp1, nph1, nhh1, npp2, nph2, nhh2;
 * goff[9q..] offset of tgam g */
int imsrg_tensor_blocks(imsrg_tensor t, int* info, int* dims, int64_t* goff);
/* copy one component of a host tensor sq_op to / from a device tensor
 * operator (basic_IMSRG.f90:50-76, op%tblck(q)%tgam(g)%X): q = -1: g = 0
 * E0, 1 fhh, 2 fph, 3 fpp; q >= 0: tgam(g+1) of block q (Fortran
 * tblck(q+1)%tgam(g+1)).  Column-major, exactly the Fortran array.
 * Synchronous. */
int imsrg_tensor_op_put(imsrg_tensor t, double* op, int q, int g, const double* host);
int imsrg_tensor_op_get(imsrg_tensor t, const double* op, int q, int g, double* host);
/* TS_commutator_111, TS_commutators.f90:10-55: one-body part of [L1, R1],
 * L scalar, R the tensor (assigns RES one-body) */
int imsrg_ts_commutator_111(imsrg_tensor t, const double* L, int Lherm, const double* R, int Rherm, double* RES,
                            int RESherm, imsrg_stream stream);
/* TS_commutator_121, TS_commutators.f90:58-303: one-body part of [L1, R2]
 * (accumulates; fhh / fpp mirrored with RES's herm) */
int imsrg_ts_commutator_121(imsrg_tensor t, const double* L, int Lherm, const double* R, int Rherm, double* RES,
                            int RESherm, imsrg_stream stream);
/* TS_commutator_211, TS_commutators.f90:307-552: one-body part of
 * -[R1, L2]; LCC = the scalar cross-coupled set of L
 * (imsrg_calculate_cross_coupled) */
int imsrg_ts_commutator_211(imsrg_tensor t, const double* LCC, int LCCherm, const double* R, int Rherm, double* RES,
                            int RESherm, imsrg_stream stream);
/* TS_commutator_122, TS_commutators.f90:555-663: two-body part of [L1, R2]
 * (assigns RES two-body) */
int imsrg_ts_commutator_122(imsrg_tensor t, const double* L, int Lherm, const double* R, int Rherm, double* RES,
                            int RESherm, imsrg_stream stream);
/* TS_commutator_212, TS_commutators.f90:666-806: two-body part of -[R1, L2]
 * (accumulates) */
int imsrg_ts_commutator_212(imsrg_tensor t, const double* L, int Lherm, const double* R, int Rherm, double* RES,
                            int RESherm, imsrg_stream stream);
/* TS_commutator_222_pp_hh, TS_commutators.f90:1070-1456 (accumulates; w1,
 * w2: tensor-sized intermediates left for TS_commutator_221) */
int imsrg_ts_commutator_222_pp_hh(imsrg_tensor t, const double* L, int Lherm, const double* R, int Rherm,
                                  double* RES, int RESherm, double* w1, double* w2, imsrg_stream stream);
/* TS_commutator_221, TS_commutators.f90:809-1067 (accumulates; after
 * 222_pp_hh; pm = L%herm * R%herm) */
int imsrg_ts_commutator_221(imsrg_tensor t, const double* w1, const double* w2, int pm, double* RES, int RESherm,
                            imsrg_stream stream);
/* calculate_single_pandya, cross_coupled.f90:1416-1634, of every block and
 * both terms: RCC (IMSRG_TSIZE_CC doubles) <- CCX and CCR of the tensor R */
int imsrg_calculate_single_pandya(imsrg_tensor t, const double* R, int Rherm, double* RCC, imsrg_stream stream);
/* TS_commutator_222_ph, TS_commutators.f90:1459-1849 (accumulates): LCC the
 * scalar cross-coupled set of L, RCC the tensor cross-coupled workspace
 * (filled here, as the reference fills it block by block).  Uses the
 * reference's factor LCC%herm on the Wx products (TS_commutators.f90:1506),
 * which makes it the commutator for an antihermitian L (the BCH generator). */
int imsrg_ts_commutator_222_ph(imsrg_tensor t, const double* LCC, int LCCherm, double* RCC, const double* R,
                               int Rherm, double* RES, int RESherm, imsrg_stream stream);
/* RES = [L, R]: the term sequence of one nested commutator of BCH_TENSOR
 * (IMSRG_MAGNUS.f90:425-446) with its own workspaces */
int imsrg_ts_commutator(imsrg_tensor t, const double* L, int Lherm, const double* R, int Rherm, double* RES,
                        int RESherm, imsrg_stream stream);
/* BCH_TENSOR, IMSRG_MAGNUS.f90:389-464: HS <- e^G HS e^-G for the tensor HS
 * (in place), terms iw = 2..30 until mat_frob_norm(term) * coef < conv (the
 * reference's conv = 1e-6 default-real literal is 9.999999974752427e-07).
 * iterations = the last iw; advals (optional, 30 doubles) as there.
 * Synchronous (one norm read per term). */
int imsrg_bch_tensor(imsrg_tensor t, const double* G, int Gherm, double* HS, int HSherm, double conv,
                     int* iterations, double* advals, imsrg_stream stream);
/* mat_frob_norm, basic_IMSRG.f90:4056-4087, rank > 0 branch (returns the
 * value; synchronises) */
int imsrg_tensor_frob_norm(imsrg_tensor t, const double* op, double* value, imsrg_stream stream);
/* add_sq_op / append_operator (basic_IMSRG.f90:3551-3611) on tensors:
 * C = ax A + bx B */
int imsrg_tensor_add(imsrg_tensor t, const double* A, double ax, const double* B, double bx, double* C,
                     imsrg_stream stream);

#ifdef __cplusplus
}
#endif
#endif
