/*
 * libpropack_amd -- C ABI of the MI355X (gfx950) truncated-SVD engine.
 *
 * Two layers sit behind this header:
 *
 *  1. The drop-in drivers DLANSVD / DLANSVD_IRL / SLANSVD / SLANSVD_IRL with
 *     the reference's Fortran calling sequences (declared in propack.h).
 *     They are host Fortran and call the engine below through iso_c_binding.
 *  2. The engine: a device-resident CSR operator that replaces the user
 *     APROD callback, and the Lanczos bidiagonalization step (the
 *     reference's DLANBPRO with its reorthogonalization and level-1 work)
 *     running on the GPU, plus the basis update GEMM of restarts and Ritz
 *     vectors.
 *
 * Conventions: plain C types only; vector/matrix arguments are column-major;
 * basis column indices are 1-based as in the reference; every function
 * returns PK_OK (0) or a negative PK_ERR_* code and never exits the process
 * (the reference STOPs on workspace errors, double/dgemm_ovwr.F:17-18,68).
 * Handles are not thread-safe; use one engine per host thread.
 */
#ifndef PROPACK_AMD_H
#define PROPACK_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PK_OK = 0,
  PK_ERR_ARG = -101,    /* invalid argument */
  PK_ERR_HIP = -102,    /* HIP runtime error (launch, copy, allocation) */
  PK_ERR_OOM = -103,    /* device memory exhausted */
  PK_ERR_HANDLE = -104, /* not a live handle */
  PK_ERR_NOTSUP = -105  /* combination not supported */
};
/* Element precision of arrays: real double / float, or complex double /
 * float stored as (re, im) pairs (Fortran COMPLEX*16 / COMPLEX). */
enum { PK_F64 = 0, PK_F32 = 1, PK_C128 = 2, PK_C64 = 3 };
enum { PK_HOST = 0, PK_DEVICE = 1 }; /* where caller arrays live */

typedef struct pk_csr_opaque* pk_csr_t;
typedef struct pk_lbd_opaque* pk_lbd_t;

/* Last error message raised in the process, by any thread (the batched
 * drivers raise theirs on host worker threads); the returned copy stays
 * valid until the calling thread's next call. */
const char* pk_last_error(void);

/* ---------------------------------------------------------------------- */
/* Device CSR operator.  Replaces the user-supplied APROD of
 * double/dlansvd.F:20-33 and double/dlanbpro.F:28-41 (Y = A*X for 'n',
 * Y = A^T*X for 't').  rowptr (m+1) and colind (nnz) are 0-based int32;
 * values are double (PK_F64) or float (PK_F32), or complex (PK_C128 /
 * PK_C64, (re, im) pairs) for the complex drivers, whose APROD applies
 * A^H for 'c' (complex16/zlanbpro.F:297).  A^T is built on the device at
 * creation.  Complex operators take column panels (pk_csr_create_panels,
 * at most 8 per direction; by default as many as keep a panel's slice of x
 * within 2 MiB when rows hold 8 or more nonzeros on average), every panel
 * in one launch; pk_csr_set_lanes, pk_csr_set_sort, pk_csr_autotune and
 * pk_csr_spmv_noalloc return PK_ERR_NOTSUP for them.                      */
int pk_csr_create(int m, int n, long long nnz, const int* rowptr, const int* colind,
                  const void* val, int val_prec, int where, pk_csr_t* op);
int pk_csr_destroy(pk_csr_t op);
int pk_csr_dims(pk_csr_t op, int* m, int* n, long long* nnz);
/* y = op(A) x ; x, y in precision prec, on the host or the device.  Direct
 * applications (this call and pk_dcsr_aprod_/pk_scsr_aprod_) use the
 * operator's own scratch for multi-panel and long-row sums, so they must not
 * overlap each other on one operator; engines (pk_lbd_create) carry their
 * own scratch and may share an operator. */
int pk_csr_spmv(pk_csr_t op, char trans, const void* x, void* y, int prec, int where);
/* Layout of the SpMV (no reference counterpart).
 * pk_csr_create_panels: pk_csr_create with the column panels of A and A^T
 * given (powers of two up to 16; a direction keeps one panel when its
 * columns are fewer).  pk_csr_panels reports an operator's layout.
 * pk_csr_set_lanes / pk_csr_lanes: lanes of a wavefront per row in the
 * SpMV's row reduction, as the built-in choice (the power of two nearest a
 * panel's mean row length) times 2^shift, shift in [-2, 2], clamped to
 * [2, 64].
 * pk_csr_set_sort / pk_csr_sort: the column-sorted layout of a direction
 * (cap > 0: row blocks of at most cap nonzeros, 1024 <= cap <= 16384, each
 * block's entries stored in increasing column order so that the gathers of
 * nearby columns share cache lines; built on the device, 12 bytes per
 * nonzero on top of the row-block layout when a panel's width and the cap
 * fit one 32-bit word, 14 otherwise), the row-block layout (cap = 0), the
 * row-wise layout (cap = -1: each row's lanes sum the row-block arrays
 * straight from memory, no LDS staging), or the column-sorted layout with
 * every column panel in one launch (cap in [-16384, -1024], blocks of at
 * most -cap nonzeros: each panel's FP64 row sums go to an array of their
 * own, added in panel order by a second kernel -- the same sums as panel
 * by panel -- and the workgroups are ordered so that each XCD works on one
 * panel at a time; np * rows doubles more per operator and per engine.  A
 * direction with one panel, rows longer than 1024 nonzeros, or panels with
 * different lanes per row runs its panels one launch each; pk_csr_sort
 * reports -cap only when the panels run together; an engine created before
 * the switch keeps the panel-by-panel path).
 * Row sums are bit-identical to the row-block layout's for the same lanes;
 * only the per-block partial sums of the norms are grouped differently.
 * Not while a solve that uses the operator runs.  A batch (pk_batch_create)
 * takes the layout its operators have when it is created; once an
 * operator's lanes or sorted layout change, the batch's calls fail with
 * PK_ERR_ARG until a new batch is created.
 * pk_csr_autotune: builds the operator with 1, 2, 4 and 8 panels per
 * direction (up to 16 when sort_a and sort_t are not NULL) and, when
 * lanes_a and lanes_t are not NULL, lane shifts -1, 0 and 1 on each (-2
 * too when sort_a and sort_t are not NULL), and,
 * when sort_a and sort_t are not NULL, each in the row-block layout, the
 * row-wise layout and the sorted layout with caps 8192 and 16384, panel by
 * panel and all panels in one launch (reported as 0, -1, 8192, 16384,
 * -8192, -16384); times each direction on the
 * device and returns the fastest panel counts (and lane shifts, block caps)
 * and their microseconds per application (inputs as pk_csr_create; nothing
 * is kept).  Settings differ only in how sums are grouped. */
int pk_csr_create_panels(int m, int n, long long nnz, const int* rowptr, const int* colind,
                         const void* val, int val_prec, int where, int panels_a, int panels_t,
                         pk_csr_t* op);
int pk_csr_panels(pk_csr_t op, int* panels_a, int* panels_t);
int pk_csr_set_lanes(pk_csr_t op, int shift_a, int shift_t);
int pk_csr_lanes(pk_csr_t op, int* shift_a, int* shift_t);
int pk_csr_set_sort(pk_csr_t op, int cap_a, int cap_t);
int pk_csr_sort(pk_csr_t op, int* cap_a, int* cap_t);
int pk_csr_autotune(int m, int n, long long nnz, const int* rowptr, const int* colind,
                    const void* val, int val_prec, int where, int* panels_a, int* panels_t,
                    double* us_a, double* us_t, int* lanes_a, int* lanes_t, int* sort_a,
                    int* sort_t);

/* Matrix ingest (no reference counterpart in the library; the reference's
 * example drivers read Harwell-Boeing, coordinate and dense files,
 * README:103-118).
 * pk_matrix_read: a Matrix Market coordinate file (real / integer / pattern;
 * general / symmetric / skew-symmetric), a Harwell-Boeing file of an
 * assembled real or pattern matrix (RUA, RRA, RSA, RZA, P?A), or plain
 * "i j value" triples (1-based, optional "m n nnz" first line), as 0-based
 * coordinate arrays with symmetric storage expanded; free each array with
 * pk_matrix_free.
 * pk_coo_to_csr: coordinate entries (any order, duplicates allowed, indices
 * 0- or 1-based by `base`) -> CSR sorted by (row, column) with duplicates
 * summed in input order, built on the device; inputs and outputs in host
 * or device memory (`where`); colind/val need nnz entries of room,
 * *nnz_out receives the distinct count.
 * Complex values (val_prec PK_C128 / PK_C64, (re, im) pairs) are summed
 * part by part.
 * pk_csr_create_coo: the same straight into a device operator.
 * pk_csr_read: a file (formats above) straight into a device operator;
 * with a complex precision it also reads Matrix Market complex files
 * (general / symmetric / skew-symmetric / Hermitian; a real file reads with
 * zero imaginary parts).  pk_matrix_read and real precisions refuse complex
 * files with PK_ERR_NOTSUP. */
int pk_matrix_read(const char* path, int* m, int* n, long long* nnz, int** rows, int** cols,
                   double** vals);
void pk_matrix_free(void* p);
int pk_coo_to_csr(int m, int n, long long nnz, const int* rows, const int* cols, const void* vals,
                  int val_prec, int where, int base, int* rowptr, int* colind, void* val,
                  long long* nnz_out);
int pk_csr_create_coo(int m, int n, long long nnz, const int* rows, const int* cols,
                      const void* vals, int val_prec, int where, int base, pk_csr_t* op);
int pk_csr_read(const char* path, int val_prec, pk_csr_t* op);
/* Store the handle in IPARM(1:2) so that passing pk_dcsr_aprod_ (or
 * pk_scsr_aprod_) as APROD selects the device operator in the drivers. */
int pk_csr_to_iparm(pk_csr_t op, int* iparm);
/* y = op(A) x with no allocation of any kind (no reference counterpart: it
 * is for callers that cannot allocate -- code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  x and y are
 * device memory in the operator's precision; the work is enqueued on
 * `stream` (a hipStream_t; NULL for the default stream) and the call
 * returns without waiting.  Partial sums, multi-panel running sums and
 * long-row sums go to a fixed 48 MiB device array per GPU, set up on the
 * first call on each device; an operator needing more (about 6M rows in a
 * multi-panel direction) is refused with PK_ERR_NOTSUP.  Results are
 * bit-identical to pk_csr_spmv's.  Calls that overlap on one GPU share that
 * array and race: order them on one stream. */
int pk_csr_spmv_noalloc(pk_csr_t op, char trans, const void* x, void* y, int prec,
                        void* stream);

/* The library's own APROD with the reference callback signature
 * (double/dlanbpro.F:28-41; hidden CHARACTER length last).  Passed to
 * DLANSVD/DLANSVD_IRL together with IPARM from pk_csr_to_iparm, the
 * drivers recognise it and keep the whole Lanczos process on the device;
 * called directly it computes Y = op(A) X for host arrays. */
void pk_dcsr_aprod_(const char* transa, const int* m, const int* n, const double* x, double* y,
                    double* dparm, int* iparm, size_t len_transa);
void pk_scsr_aprod_(const char* transa, const int* m, const int* n, const float* x, float* y,
                    float* dparm, int* iparm, size_t len_transa);
/* The same for complex operators (complex16/zlanbpro.F:28-41: TRANSA 'n'
 * for Y = A*X, 'c' for Y = A^H*X; X and Y COMPLEX*16 / COMPLEX as (re, im)
 * pairs), for ZLANSVD/ZLANSVD_IRL and CLANSVD/CLANSVD_IRL. */
void pk_zcsr_aprod_(const char* transa, const int* m, const int* n, const void* x, void* y,
                    void* zparm, int* iparm, size_t len_transa);
void pk_ccsr_aprod_(const char* transa, const int* m, const int* n, const void* x, void* y,
                    void* cparm, int* iparm, size_t len_transa);

/* ---------------------------------------------------------------------- */
/* Lanczos bidiagonalization engine.  prec = storage type of the bases
 * U (m x (kmax+1)) and V (n x kmax) kept on the device; all reductions and
 * the omega-recurrence run in double.  op may be NULL when a host callback
 * is registered with pk_lbd_set_callback.  A complex prec (PK_C128 /
 * PK_C64) makes the engine of ZLANBPRO / CLANBPRO (complex16/zlanbpro.F):
 * complex bases, the same real B, its operator complex of the same
 * precision and a callback with the complex signature of that precision;
 * its steps run on the host executor.  pk_lbd_residuals, pk_lbd_promote,
 * pk_lbd_seed and the batched calls take real engines only.                */
int pk_lbd_create(int prec, int m, int n, int kmax, pk_csr_t op, pk_lbd_t* eng);
int pk_lbd_destroy(pk_lbd_t eng);
/* Host APROD callback (Fortran ABI, arrays in aprod_prec) used for
 * user-supplied operators: X/Y are staged through pinned host memory. */
int pk_lbd_set_callback(pk_lbd_t eng, void* aprod, int aprod_prec, void* dparm, int* iparm);
/* Copy column col (1-based) of U ('u') or V ('v') from x (host or device
 * memory) / to the host. */
int pk_lbd_put(pk_lbd_t eng, char which, int col, const void* x, int prec);
int pk_lbd_get(pk_lbd_t eng, char which, int ncols, void* X, int ldx, int prec);
/* Device pointer and leading dimension of the current basis. */
int pk_lbd_basis(pk_lbd_t eng, char which, void** dptr, long long* ld);
/* ||column col|| of U or V (pdnrm2, double/dblasext.F:6-31). */
int pk_lbd_nrm2(pk_lbd_t eng, char which, int col, double* nrm);

/* DGETU0 (double/dgetu0.F:11-89): random vector (DLARNV IDIST=2, fixed
 * seed (1,3,5,7)) mapped through op(A) into column j+1 of U ('n') or V
 * ('t'), orthogonalized against columns 1..j.  ierr = -1 if every try
 * gave a zero vector. */
int pk_lbd_getu0(pk_lbd_t eng, char transa, int j, int ntry, double* u0norm, int icgs,
                 double* anormest, int* ierr);

/* DLANBPRO (double/dlanbpro.F:1-549): extend a k0-step Lanczos
 * bidiagonalization to *k steps.  alpha/beta are B(:,1)/B(:,2) (length
 * >= *k, host, in/out); rnorm in/out; doption(1:3) = delta, eta, anorm
 * (anorm written back as the reference does at :547); ioption(1:2) =
 * cgs, elr.  *ierr as the reference's IERR.  *k may shrink on an invariant
 * subspace. */
int pk_lbd_lanbpro(pk_lbd_t eng, int k0, int* k, double* alpha, double* beta, double* rnorm,
                   double* doption, const int* ioption, int* ierr);

/* Basis update of restarts and Ritz vectors, replacing DGEMM_OVWR_LEFT
 * (double/dgemm_ovwr.F:56-89; calls at double/dlansvd_irl.F:387,394 and
 * double/dritzvec.F:160,193):  X(:,1:ncols) <- X(:,1:kdim) * op(B), X the
 * U ('u') or V ('v') basis, B host double column-major, op(B) kdim x ncols:
 * transb 'n' -> B is kdim x ncols, 't' -> B is ncols x kdim. */
int pk_lbd_update(pk_lbd_t eng, char which, char transb, int kdim, int ncols, const double* B,
                  int ldb);

/* Counters with the meaning of COMMON /timing/ (double/stat.h:12-15);
 * out must hold 6 values: out[0] nopx, [1] nreorth, [2] ndot (reference semantics, see
 * double/dreorth.F:92,167), [3] basis columns actually streamed by
 * reorthogonalization, [4] host synchronizations, [5] CGS passes beyond the
 * first of a reorthogonalization. */
int pk_lbd_stats(pk_lbd_t eng, long long* out);
/* How often the rare paths of the reference's DLANBPRO ran (no reference
 * counterpart; the reference only prints warnings): out holds 8 values,
 * out[0..2] = alpha below anorm*epsn (new random direction,
 * double/dlanbpro.F:372-409), |alpha| below the safe minimum (DSAFESCAL's
 * stepwise scaling, double/dsafescal.F:49-53), five failed CGS passes
 * (v_j = 0, double/dreorth.F:96-98); out[3..5] the same for beta
 * (:503-539); out[6] near-invariance warnings (IERR = j > 0, :406-409,
 * :536-539); out[7] = 0.  Reset with pk_lbd_reset_stats. */
int pk_lbd_path_stats(pk_lbd_t eng, long long* out);
int pk_lbd_reset_stats(pk_lbd_t eng);
int pk_lbd_sync(pk_lbd_t eng);
int pk_lbd_shape(pk_lbd_t eng, int* m, int* n, int* kmax, int* prec);
/* Rounding model of the Lanczos thresholds (no reference counterpart; the
 * reference's are double/dlanbpro.F:148-149): epsn = leff * eps and
 * epsn2 = sqrt(leff) * eps with leff = max(m, n) (model 0, the reference's,
 * default) or leff = 4 (model 1: FP32 storage with FP64 accumulation, where
 * only the storage rounding enters; the FP32-basis form uses it).
 * pk_lbd_eps returns eps of the engine's precision and leff. */
int pk_lbd_set_rounding(pk_lbd_t eng, int model);
int pk_lbd_eps(pk_lbd_t eng, double* eps, double* leff);

/* Batched Lanczos steps (no reference counterpart; replaces a per-problem
 * loop of DLANBPRO calls, double/dlanbpro.F:1-2, over independent
 * problems): the engines of a batch (same precision, each with its own
 * device CSR operator) run their Lanczos steps together, one launch per
 * kernel per step for the whole batch, each problem gated by its own
 * control state; a problem gives the same results as alone.  One engine
 * may not appear twice in a batch (PK_ERR_ARG); problems may share an
 * operator (each engine carries its own SpMV scratch).  An operator in the
 * concurrent-panel layout (negative sorted cap) runs its panels one launch
 * each inside a batch: the same row sums, norm partials grouped as panel by
 * panel, so such a problem matches its own concurrent solve to rounding,
 * not bit for bit.  Per-problem host work runs on up to 16 host threads
 * (PK_HOST_THREADS overrides). */
typedef struct pk_batch_opaque* pk_batch_t;
int pk_batch_create(int nprob, const pk_lbd_t* engines, pk_batch_t* batch);
int pk_batch_destroy(pk_batch_t batch);
/* Step chains of batched calls as HIP graphs (on != 0, the default): each
 * call's kernel chain is captured once per (step range, ELR and CGS pass
 * counts) and replayed, up to 8 chains per batch; on = 0 enqueues the
 * kernels directly.  Same results either way (no reference counterpart). */
int pk_batch_set_graph(pk_batch_t batch, int on);
/* Bidiagonal SVDs of the batched restarted driver on the device (replaces
 * the host DBDSQR of its error-bound step, double/dlansvd_irl.F:226-229,
 * for every problem at once; on = 0, the default, keeps the host DBDSQR,
 * which makes a batched problem bit-identical to the same problem alone).
 * pk_batch_bsvd / pk_bsvd_batch: for nb upper bidiagonal matrices (order
 * n[i] <= 128, diagonal d[i*ld + j], superdiagonal e[i*ld + j]), the
 * singular values in decreasing order to sigma[i*ld + r] and the last
 * component of each matching left singular vector to bottom[i*ld + r]
 * (one-sided Jacobi on the device, one workgroup per matrix; host arrays;
 * pk_bsvd_batch allocates its device buffers per call, pk_batch_bsvd keeps
 * them in the batch and runs on its stream). */
int pk_batch_set_bsvd(pk_batch_t batch, int on);
int pk_batch_get_bsvd(pk_batch_t batch, int* on);
int pk_batch_bsvd(pk_batch_t batch, int nb, const int* n, const double* d, const double* e,
                  int ld, double* sigma, double* bottom);
int pk_bsvd_batch(int nb, const int* n, const double* d, const double* e, int ld, double* sigma,
                  double* bottom);
/* pk_lbd_lanbpro for every problem i with active[i] != 0: k0[i] -> k[i]
 * steps; alpha[i], beta[i] its B(:,1), B(:,2); rnorm[i]; doption(4*i+1:
 * 4*i+3) its delta, eta, anorm; ioption shared; ierr[i]. */
int pk_batch_lanbpro(pk_batch_t batch, const int* active, const int* k0, int* k,
                     double* const* alpha, double* const* beta, double* rnorm, double* doption,
                     const int* ioption, int* ierr);
/* pk_lbd_update (DGEMM_OVWR_LEFT, double/dgemm_ovwr.F:56-89) for problems
 * sel[0..nsel) (0-based) of the batch in one grouped launch: problem
 * sel[q]'s basis `which` becomes X(:, 1:ncols[q]) = X(:, 1:kdim[q]) op(B_q),
 * B_q = bm[q] (host, leading dimension ldb[q]); the products are bit for bit
 * the per-engine update's.  A problem may appear once in sel (PK_ERR_ARG
 * otherwise, before anything runs).  kdim past 316 falls back to one
 * update per engine.  Returns when the products are done.  The batched
 * restarted driver issues its restart GEMMs this way. */
int pk_batch_update(pk_batch_t batch, int nsel, const int* sel, char which, char transb,
                    const int* kdim, const int* ncols, const double* const* bm,
                    const int* ldb);
/* Step execution (no reference counterpart): mode 0 runs the Lanczos steps
 * as device-resident kernel chains whose scalar control (the omega
 * recurrence, ELR and reorthogonalization decisions) runs in control
 * kernels on the GPU, holding `tries` CGS passes per side; rare paths and
 * further passes continue on the host.  mode 1 runs every decision on the
 * host after a synchronization (same control code, same results).  The
 * host-callback operator always uses mode 1.  Environment PK_LBD_MODE=host
 * selects mode 1 at creation.  Without a call, mode 0 holds one CGS pass
 * per side and moves to two once a step needed a second pass. */
int pk_lbd_set_mode(pk_lbd_t eng, int mode, int tries);
/* U(:,1) of the double engine dst <- sum of the first ncols columns of
 * src's U (any precision): the start vector of the FP64 polish that follows
 * an FP32-basis solve (no reference counterpart). */
int pk_lbd_seed(pk_lbd_t dst, pk_lbd_t src, int ncols);
/* Residuals of k triplets on the device (no reference counterpart; the
 * reference only estimates them from B_k): the singular vectors are
 * columns 1..k of src's bases (src an engine of eng's shape, any
 * precision), sigma[i] their values, and the operator is eng's (a device
 * CSR):  res[2i] = ||A v_{i+1} - sigma[i] u_{i+1}||, res[2i+1] =
 * ||A^T u_{i+1} - sigma[i] v_{i+1}||, the vectors converted to eng's
 * precision and every sum accumulated in double.  With a double eng and an
 * FP32-basis src this is the FP64 check of FP32-basis triplets.
 * pk_lbd_promote: columns 1..ncols of src's bases copied (converted) into
 * dst's. */
int pk_lbd_residuals(pk_lbd_t eng, pk_lbd_t src, int k, const double* sigma, double* res);
int pk_lbd_promote(pk_lbd_t dst, pk_lbd_t src, int ncols);
/* Live kernel timing (no reference counterpart).  classes: bit mask of the
 * kernel classes to time (0 turns timing off, PK_PROF_ALL times all); each
 * launch of a selected class is bracketed by HIP events on the engine's
 * stream (two marker packets, a few microseconds of idle stream each) and
 * its algorithmic bytes/flops are accumulated.  out holds 6 classes x
 * {launches, total ms, bytes, flops}: 0 SpMV A*v, 1 SpMV A^T*u,
 * 2 Gram-Schmidt V^T w, 3 Gram-Schmidt w -= V h, 4 basis GEMM, 5 level-1. */
#define PK_PROF_SPMV 0x03
#define PK_PROF_ALL 0x3f
int pk_lbd_profile(pk_lbd_t eng, int classes);
int pk_lbd_profile_get(pk_lbd_t eng, double* out);
int pk_lbd_profile_reset(pk_lbd_t eng);

/* ---------------------------------------------------------------------- */
/* Device-resident drivers (no reference counterpart; for callers whose
 * matrix and results live on the GPU).  They run the DLANSVD /
 * DLANSVD_IRL algorithms (double/dlansvd.F:1-291, double/dlansvd_irl.F:1-419)
 * on an engine created by the caller: the start vector is the engine's
 * U(:,1) (zero => the reference's random start), sigma/bnd/doption are
 * host double arrays with the reference's meaning, and the singular vectors
 * are left in columns 1..k of the engine's bases (pk_lbd_basis).  Options
 * are single characters by value.  Return INFO as the reference defines it
 * (0, J > 0 invariant subspace, -1 not converged) or a PK_ERR_* code.     */
int pk_lansvd_run(pk_lbd_t eng, char jobu, char jobv, int* k, int kmax, double* sigma,
                  double* bnd, double tolin, double* doption, int* ioption);
int pk_lansvd_irl_run(pk_lbd_t eng, char which, char jobu, char jobv, int* dim, int p, int* neig,
                      int maxiter, double* sigma, double* bnd, double tolin, double* doption,
                      int* ioption);
/* FP32-basis form with FP64 polish: the restarted solve runs first on
 * eng32 (FP32 operator values and bases, FP64 reductions, omega recurrence,
 * B and its SVD) to tolerance max(tolin, 16*eps32).  When tolin is at least
 * 16*eps32 the FP32 triplets are checked with eng64's operator
 * (pk_lbd_residuals); if every ||A v - s u|| meets tolin * s_1 they are the
 * result (promoted into eng64's bases) and no polish runs.  Otherwise the
 * sum of the FP32 left Ritz vectors seeds the FP64 restarted solve on
 * eng64, which meets tolin.
 * sigma32/bnd32 receive the FP32-phase values, info32 = {INFO, converged}
 * of that phase; the polished vectors are left in eng64's bases. */
int pk_lansvd_irl_mixed(pk_lbd_t eng64, pk_lbd_t eng32, char which, char jobu, char jobv,
                        int* dim, int p, int* neig, int maxiter, double* sigma, double* bnd,
                        double* sigma32, double* bnd32, double tolin, double* doption,
                        int* ioption, int* info32);
/* The same scheme around the non-restarted driver (DLANSVD): the FP32
 * phase runs DLANSVD on eng32 to max(tolin, 16*eps32) with kmax; the check,
 * promotion or seeded FP64 DLANSVD on eng64 follow as above.  k is in/out
 * as in pk_lansvd_run; sigma/bnd/sigma32/bnd32 hold at least kmax values. */
int pk_lansvd_mixed(pk_lbd_t eng64, pk_lbd_t eng32, char jobu, char jobv, int* k, int kmax,
                    double* sigma, double* bnd, double* sigma32, double* bnd32, double tolin,
                    double* doption, int* ioption, int* info32);
/* DLANSVD_IRL on every engine of a batch (engs[i] in pk_batch_create's
 * order; start vectors their U(:,1)); no reference counterpart, it replaces
 * a loop of DLANSVD_IRL calls (double/dlansvd_irl.F:1-3) over independent
 * problems.  Every problem runs the reference's control flow on its own;
 * their Lanczos steps advance together.  sigma/bnd are ldsig x nb
 * (column-major, column i for problem i); nconv[i], info[i] per problem.
 * Returns PK_OK or the first PK_ERR_* code. */
int pk_lansvd_irl_batch(int nb, pk_lbd_t* engs, pk_batch_t batch, char which, char jobu,
                        char jobv, int dim, int p, int neig, int maxiter, double tolin,
                        double* sigma, double* bnd, int ldsig, int* nconv, int* info,
                        double* doption, int* ioption);

/* ---------------------------------------------------------------------- */
/* Host scalar routines, exported so their results can be compared with the
 * reference's bit for bit.  Arrays are 1-based like the reference's
 * (element 0 unused).                                                      */
/* DLARNV(IDIST=2) from the fixed seed (1,3,5,7) of double/dgetu0.F:47-50,
 * n values in double (prec 0) or float (prec 1). */
int pk_larnv_pm1(int prec, long long n, void* x);
/* The same n values generated on the device into x (device memory), as the
 * engine's DGETU0 does; bit-identical to pk_larnv_pm1. */
int pk_larnv_pm1_device(int prec, long long n, void* x);
/* DUPDATE_MU / DUPDATE_NU / DCOMPUTE_INT / DSET_MU (double/dlanbpro.F:555-723). */
void pkh_update_mu(double* mumax, double* mu, const double* nu, int j, const double* alpha,
                   const double* beta, double anorm, double eps1);
void pkh_update_nu(double* numax, const double* mu, double* nu, int j, const double* alpha,
                   const double* beta, double anorm, double eps1);
void pkh_compute_int(const double* mu, int j, double delta, double eta, int* index);
void pkh_set_mu(int k, double* mu, const int* index, double val);
/* DLARTG (double/Lapack_Util/dlartg.f), DBDQR, DBSVDSTEP (jobu=jobv='y'),
 * DREFINEBOUNDS (double/dbsvd.F); 0-based C arrays here. */
void pkh_lartg(double f, double g, double* c, double* s, double* r);
void pkh_bdqr(int ignorelast, char jobq, int n, double* d, double* e, double* c1, double* c2,
              double* qt, int ldq);
void pkh_bsvdstep(int m, int n, int k, double sigma, double* d, double* e, double* u, int ldu,
                  double* v, int ldv);
void pkh_refinebounds(int n, int k, const double* theta, double* bound, double tol,
                      double eps34);

#ifdef __cplusplus
}
#endif
#endif
