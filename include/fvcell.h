/*
 * fvcell.h -- C ABI of the MI355X (gfx950) sparse-solver library that
 * replaces fortranFVCell's linear-solver object (src/Solver.f: solver and
 * getRes, with the Krylov routines of src/Iterativos*.f and the SpMV / dot
 * kernels of src/Matvec*.f underneath).
 *
 * Conventions shared by every entry point:
 *   - arrays are passed as plain pointers to element 1 of the reference's
 *     Fortran arrays; index arrays hold the reference's 1-based int32 values;
 *     values are IEEE double;
 *   - layouts are the reference's (src/Datastruct.f):
 *       matriz 1 CSRD   : iax = ia(neq+1), ja(nad) without the diagonal, sorted;
 *                         unsym: au == al hold both triangles; else al = lower only
 *       matriz 2 CSR    : iax = ia(neq+1), ja(nad) with the diagonal, all values in ad
 *       matriz 3 CSRC   : lower pattern; al = row values, au = transposed upper
 *                         (unsym); symmetric: al only
 *       matriz 4 ELLPACK: iax = {w, neq+1}; ja(w,neq), al(w,neq) slot-fastest,
 *                         padding ja = 1, al = 0; au == al
 *   - "host" entry points take host memory and copy it to and from the GPU;
 *     "_dev" entry points take device memory (hipMalloc'd) and a hipStream_t
 *     passed as void* (NULL = the default stream);
 *   - every entry point returns an fvc status code (FVC_OK = 0).
 */
#ifndef FVCELL_H
#define FVCELL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference stops the program where these are nonzero
 * (src/Iterativos.f:76-77,198-199,405-406; src/Solver.f:225-227,306-308,338-339). */
#define FVC_OK 0            /* converged under the requested stopping rule        */
#define FVC_MAXIT 1         /* maxit reached (x holds the last iterate)           */
#define FVC_BREAKDOWN 2     /* a Krylov denominator became exactly zero           */
#define FVC_NONFINITE 3     /* a NaN or Inf appeared in the iteration scalars     */
#define FVC_UNSUPPORTED 4   /* method x storage combination not available        */
#define FVC_DEVICE_ERROR 5  /* a HIP call failed, or no usable GPU                */
#define FVC_BAD_ARGUMENT 6  /* inconsistent sizes or null pointers                */

/* Krylov methods, the reference's `code` (src/Solver.f:19-21). */
#define FVC_PCG 1
#define FVC_BICGSTAB 2
#define FVC_GMRES 3

/* Storage formats, the reference's `matriz` (src/Datastruct.f:12-15). */
#define FVC_CSRD 1
#define FVC_CSR 2
#define FVC_CSRC 3
#define FVC_ELLPACK 4

/* x0 rules: which initial guess each path uses.  The reference zeroes x
 * (init = .true.) on some paths and warm-starts on others, depending on the
 * openmpSolver flag (src/Solver.f:94-165,196-260; src/Iterativos.f:288-290). */
#define FVC_X0_AS_OPENMP 0      /* the reference with openmpSolver = .true. (the decks' setting) */
#define FVC_X0_AS_SEQUENTIAL 1  /* the reference with openmpSolver = .false.                     */
#define FVC_X0_ZERO 2
#define FVC_X0_GIVEN 3

/* Stopping rules. */
#define FVC_STOP_REFERENCE 0    /* each method's own test (src/Iterativos.f:73,190,376)     */
#define FVC_STOP_RELRES 1       /* ||b - A x||_2 <= tol*||b||_2.  PCG and BiCGSTAB test the
                                   recursive residual, then check the true FP64 residual and
                                   restart from x while it misses the tolerance and each
                                   restart at least halves it (iterations include the
                                   restarts; true_relres reports what was reached).  GMRES
                                   tests the true residual at the start of every cycle, and
                                   runs a cycle until the preconditioned residual has dropped
                                   by the factor the true one still needs; maxit counts
                                   cycles                                                    */

/* Engines. */
#define FVC_ENGINE_AUTO 0        /* KERNELS for PCG and BiCGSTAB, PERSISTENT for GMRES   */
#define FVC_ENGINE_KERNELS 1     /* one kernel per phase; fused recurrences sum the block
                                    partials at the next launch, the others in the last
                                    block of the launch                                 */
#define FVC_ENGINE_PERSISTENT 2  /* one launch per solve, all blocks co-resident, grid
                                    barriers with bounded waits                         */

/* Arithmetic of the iteration. */
#define FVC_PREC_FP64 0         /* everything in FP64 (the reference's real*8)              */
#define FVC_PREC_MIXED 1        /* FP32 matrix + FP32 inner PCG on the Jacobi-scaled system,
                                   FP64 x and FP64 residual recomputation ("reliable
                                   updates"); stops on ||b - A x||_2 <= tol ||b||_2       */

/* Recurrences of PCG and BiCGSTAB (same iterates in exact arithmetic). */
#define FVC_REC_FUSED 0         /* fewest grid reductions: PCG one per iteration
                                   (Chronopoulos-Gear), BiCGSTAB two                       */
#define FVC_REC_REFERENCE 1     /* the reference's scalar expressions and order: PCG two,
                                   BiCGSTAB three reductions per iteration                 */

typedef struct fvc_options {
  int32_t x0_rule;        /* FVC_X0_*; default FVC_X0_AS_OPENMP                      */
  int32_t gmres_k;        /* restart length; 0 = 15 as src/Solver.f:63               */
  int32_t stop_rule;      /* FVC_STOP_*                                              */
  int32_t true_residual;  /* nonzero: also compute ||b - A x||_2 / ||b||_2 in FP64    */
  int32_t engine;         /* FVC_ENGINE_*                                            */
  int32_t block_size;     /* 0 = tuned default                                       */
  int32_t grid_blocks;    /* 0 = tuned default (number of row chunks)                */
  int32_t precision;      /* FVC_PREC_*; mixed applies to PCG on ELLPACK / CSRD      */
  int32_t batch_work_exp10; /* batches: blocks per system ~ neq^(this/10); 0 = 15     */
  int32_t barrier;        /* persistent engine's grid barrier: 0 per-block flags,
                             1 arrival counters                                     */
  int32_t recurrence;     /* FVC_REC_*; PCG and BiCGSTAB                              */
  int32_t refresh_every;  /* mixed: at most this many FP32 iterations between FP64
                             residual recomputations; 0 = 200                         */
  int32_t refresh_drop_e3; /* mixed: recompute when the FP32 (Jacobi-scaled) residual
                             norm fell by this factor (thousandths) since the last
                             one; 0 = 300                                            */
  int32_t row_order;      /* rows to blocks: 0 tuned default (tiles for PCG and
                             BiCGSTAB from one row per thread), 1 one contiguous chunk
                             per block, 2 XCD tiles (each eighth of the rows swept
                             by one XCD's blocks together); single-system solves
                             (batches use chunks)                                    */
  int32_t lane_exchange;  /* fused PCG and BiCGSTAB: the SpMV operand of columns i-1
                             and i+1 comes from the adjacent lanes of the wavefront
                             (DPP wave shifts) instead of a gather of every vector it
                             is formed from; bit-identical either way.  0 default
                             (off: measured slower on MI355X), 1 on, 2 off           */
  int32_t gmres_orth;     /* GMRES orthogonalisation: FVC_ORTH_MGS (default) or
                             FVC_ORTH_CGS2                                           */
  int32_t batch_rows_per_wave; /* batches: at least this many rows per wavefront of a
                             system's block group (a system gets at most
                             ceil(neq / (rows x waves per block)) blocks); 0 = the
                             default of 64 rows per block                            */
  int32_t batch_launch_systems; /* batches: at most this many systems per launch (the
                             launches run one after another, each on all blocks);
                             0 = half the co-resident blocks                         */
} fvc_options;

/* fvc_options.gmres_orth */
#define FVC_ORTH_MGS 0          /* modified Gram-Schmidt in the reference's order
                                   (src/Iterativos.f:343-349): j+1 grid reductions at
                                   Arnoldi step j                                         */
#define FVC_ORTH_CGS2 1         /* classical Gram-Schmidt with one re-orthogonalisation:
                                   three grid reductions per step, the j dot products of
                                   a pass reduced together; gmres_k <= 32                 */
#define FVC_ORTH_CGS2_LAG 2     /* CGS2 with its second pass and normalisation lagged by
                                   one step: one reduction per Arnoldi step, the basis
                                   read twice per step; one kernel per phase only,
                                   gmres_k <= 64, block_size <= 256                       */

typedef struct fvc_stats {
  int32_t status;           /* FVC_* status                                          */
  int32_t iterations;       /* PCG/BiCGSTAB: iterations j; GMRES: cycles l            */
  int32_t inner_iterations; /* GMRES: Arnoldi steps (nit); else = iterations          */
  int32_t engine;           /* engine that ran                                       */
  double energy;            /* the "energy norm" the reference logs: x.Ax (PCG,
                               BiCGSTAB; src/Iterativos.f:82-83,204-205) or x.b
                               (GMRES; src/Iterativos.f:412)                         */
  double conv;              /* stopping threshold                                    */
  double last;              /* last tested quantity                                  */
  double true_relres;       /* ||b - A x||_2 / ||b||_2 (FP64), or -1 if not computed  */
  double seconds;           /* device time of the solve (HIP events)                 */
  int32_t grid_blocks;      /* row chunks used                                       */
  int32_t block_size;
  int32_t flags;            /* FVC_FLAG_* bits: which variant of the kernels ran      */
  int32_t reserved_i;
  double reserved[3];
} fvc_stats;

/* fvc_stats.flags */
#define FVC_FLAG_COL16 1        /* mixed precision read int16 column offsets (else the int32
                                   columns: an offset did not fit)                           */
#define FVC_FLAG_TILES 2        /* rows went to blocks as XCD tiles (else contiguous chunks) */

/* Fills *opt with the defaults listed above. */
void fvc_default_options(fvc_options *opt);

/* Library set-up on the current HIP device (implicit on first use). */
int fvc_init(int32_t device);
/* Human-readable text of the last error of this thread. */
const char *fvc_last_error(void);

/* ----------------------------------------------------------------------------
 * The drop-in solve.  Replaces
 *   subroutine solver(ad,au,al,x,b,iax,ja,neq,nad,unsym,tol,maxit,code,matriz,nlit)
 * of src/Solver.f:28-29.  `unsym` is the LOGICAL unsym as 0/1; `omp_solver` is
 * the caller's openmpSolver flag (include/openmp.fi:16-18), which selects the
 * reference's x0 rule per path.  Differences from the reference, by design:
 * b is preserved (the reference overwrites it, src/Iterativos.f:55,168);
 * non-convergence returns FVC_MAXIT instead of stopping the program; GMRES
 * also runs on ELLPACK and PCG on CSR/CSRC with a proper dot (the reference
 * has no branch / a malformed call there, src/Solver.f:142-149,291-323).
 * `nlit` is accepted for signature parity (it only selects the reference's
 * console output, which the Fortran shim reproduces). */
int fvc_solver(const double *ad, const double *au, const double *al, double *x,
               double *b, const int32_t *iax, const int32_t *ja, int32_t neq,
               int32_t nad, int32_t unsym, double tol, int32_t maxit, int32_t code,
               int32_t matriz, int32_t nlit, int32_t omp_solver, fvc_stats *stats);

/* The same with explicit options (host memory). */
int fvc_solve(int32_t code, int32_t matriz, int32_t unsym, int32_t neq, int32_t nad,
              const double *ad, const double *au, const double *al, double *x,
              const double *b, const int32_t *iax, const int32_t *ja, double tol,
              int32_t maxit, const fvc_options *opt, fvc_stats *stats);

/* Keeps the matrix between the two pressure-correction solves of a SIMPLE
 * step (src/Simple.f:590 and :621, between which pform rebuilds only the
 * right-hand side, src/Simple.f:607-615).  With keep != 0, the next fvc_solver
 * / fvc_solve call on this device does not upload A again if its format, neq,
 * entry count and the addresses of ad, au, al, iax, ja equal those of the
 * previous host-array solve; the caller guarantees that their contents are
 * unchanged.  One-shot: the call after that uploads as usual.  (The
 * host-array calls also keep their device copies of A, x and b between calls,
 * in grow-only buffers, instead of allocating them per call.) */
int fvc_keep_matrix(int32_t keep);

/* The same on device memory; runs on `stream`.  For ELLPACK, `iax` may be a
 * host or device pointer (its first entry, the width, is read once). */
int fvc_solve_dev(int32_t code, int32_t matriz, int32_t unsym, int32_t neq, int32_t nad,
                  const double *ad, const double *au, const double *al, double *x,
                  const double *b, const int32_t *iax, const int32_t *ja, double tol,
                  int32_t maxit, const fvc_options *opt, fvc_stats *stats, void *stream);

/* fvc_solve_dev with no allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, stopping
 * rules, results and status codes as fvc_solve_dev, for ELLPACK storage
 * (matriz 4; other formats return FVC_UNSUPPORTED).  The solver's scratch
 * (vectors, partial sums, counters, state) comes from one fixed workspace of
 * FVC_NOALLOC_WORKSPACE_BYTES per device, which the library holds until the
 * process ends; a solve that needs more returns FVC_UNSUPPORTED (BiCGSTAB:
 * about 9 vectors of neq doubles, PCG 8, GMRES(k) k+2; so up to ~400,000 rows
 * for PCG and BiCGSTAB).  fvc_init(device) must have run on the device first
 * (it sets the workspace up); a call that succeeds makes no allocation of its
 * own (no hipMalloc, hipFree, pinned or heap memory).
 * Every call on a device shares its workspace and leaves its contents there:
 * calls that overlap on one device race (each call waits for its own solve to
 * finish before it returns, so calls from one thread never overlap). */
#define FVC_NOALLOC_WORKSPACE_BYTES (32u << 20)
int fvc_solve_dev_noalloc(int32_t code, int32_t matriz, int32_t unsym, int32_t neq, int32_t nad,
                          const double *ad, const double *au, const double *al, double *x,
                          const double *b, const int32_t *iax, const int32_t *ja, double tol,
                          int32_t maxit, const fvc_options *opt, fvc_stats *stats, void *stream);

/* ----------------------------------------------------------------------------
 * Batched Jacobi-PCG: nsys independent ELLPACK systems of width w, each in
 * the reference's layout (local 1-based ja), concatenated back to back:
 * system s owns rows row_offsets[s] .. row_offsets[s+1]-1 (row_offsets[0] = 0)
 * of ad, b, x and entries w*row_offsets[s] .. of al and ja.  All systems are
 * solved in one persistent launch (or a few, when they outnumber half the
 * co-resident blocks), each with its own scalars and stopping test: a
 * system gets its own group of blocks, or, when it is too small to keep a
 * block busy for the batch, shares a block with other small systems, which
 * that block solves one after another.  stats receives nsys entries.  x0 is zero unless
 * opt->x0_rule == FVC_X0_GIVEN.  opt->precision selects FP64 or mixed.  No
 * counterpart in the reference (its solver takes one system per call). */
int fvc_solve_batch(int32_t code, int32_t nsys, const int32_t *row_offsets, int32_t w,
                    const double *ad, const double *al, double *x, const double *b,
                    const int32_t *ja, double tol, int32_t maxit, const fvc_options *opt,
                    fvc_stats *stats);
int fvc_solve_batch_dev(int32_t code, int32_t nsys, const int32_t *row_offsets, int32_t w,
                        const double *ad, const double *al, double *x, const double *b,
                        const int32_t *ja, double tol, int32_t maxit, const fvc_options *opt,
                        fvc_stats *stats, void *stream);

/* The same for CSRD systems (the reference's matriz 1 with unsym: ad, then
 * the off-diagonals of each row, src/Datastruct.f:40-77): system s owns rows
 * row_offsets[s] .. row_offsets[s+1]-1 of ad, b, x, the stored values and
 * columns val_offsets[s] .. val_offsets[s+1]-1 of a and ja (columns local and
 * 1-based), and the row pointer ia[row_offsets[s] + s .. row_offsets[s+1] + s]
 * (neq_s + 1 entries, local and 1-based: the system's own reference ia).  No
 * counterpart in the reference; the autotuner chooses between this layout
 * and ELLPACK per batch. */
int fvc_solve_batch_csrd(int32_t code, int32_t nsys, const int32_t *row_offsets,
                         const int32_t *val_offsets, const double *ad, const double *a, double *x,
                         const double *b, const int32_t *ia, const int32_t *ja, double tol,
                         int32_t maxit, const fvc_options *opt, fvc_stats *stats);
int fvc_solve_batch_csrd_dev(int32_t code, int32_t nsys, const int32_t *row_offsets,
                             const int32_t *val_offsets, const double *ad, const double *a, double *x,
                             const double *b, const int32_t *ia, const int32_t *ja, double tol,
                             int32_t maxit, const fvc_options *opt, fvc_stats *stats, void *stream);

/* ----------------------------------------------------------------------------
 * Differentiable solve (no counterpart in the reference).
 *
 * fvc_ell_transpose_dev: values of A^T for an ELLPACK matrix with a
 * symmetric pattern, same ja: alt(s,i) = al(back(s,i), ja(s,i)), 0 where
 * back(s,i) < 0 (padding).  back holds 0-based slots.
 *
 * fvc_grad_k_dev: dL/dk = -lam^T (dA/dk T - db/dk) for a finite-volume
 * diffusion operator with harmonic face conductivity
 * (1/k_f = alpha/k_c + (1-alpha)/k_N, D_f = k_f*geo; src/CellibTrans.f:
 * 510-511,527) and Dirichlet faces (a_P += k_c*wall_g, b += k_c*wall_g*T_wall;
 * src/CellibTrans.f:488-491).  Per (slot, cell) arrays geo, alpha, wall_g,
 * wall_t in the ELLPACK layout; geo = 0 marks a non-interior slot. */
int fvc_ell_transpose_dev(int32_t neq, int32_t w, const int32_t *ja, const int32_t *back,
                          const double *al, double *alt, void *stream);
int fvc_grad_k_dev(int32_t neq, int32_t w, const int32_t *ja, const double *k, const double *t,
                   const double *lam, const double *geo, const double *alpha, const double *wall_g,
                   const double *wall_t, double *grad_k, void *stream);

/* fvc_assemble_heat_dev: the operator whose gradient fvc_grad_k_dev takes,
 * assembled on the GPU in the reference's ELLPACK layout (ja as given,
 * al = -a_nb, ad, b) from per-cell k (and an optional source q, NULL for
 * none), with the reference's coefficient forms: harmonic face conductivity
 * (src/CellibTrans.f:510-511), a_nb = k_f*geo + conv (src/CellibTrans.f:527,
 * src/CellibSimple.f:789-816), Dirichlet faces as above, and
 * ad = (sum a_nb + conv_f) + sum k_c*wall_g (src/CellibSimple.f:856-859).
 * conv (per slot) and conv_f (per cell) are the fixed upwind convection
 * coefficients.  Replaces the per-solve upload of a host-assembled A. */
int fvc_assemble_heat_dev(int32_t neq, int32_t w, const int32_t *ja, const double *k, const double *q,
                          const double *geo, const double *alpha, const double *conv,
                          const double *conv_f, const double *wall_g, const double *wall_t,
                          double *ad, double *al, double *b, void *stream);

/* ----------------------------------------------------------------------------
 * The post-solve step of a SIMPLE iteration on device memory, so that p', its
 * gradient and the corrected velocities stay in HBM after the solve.
 *
 * fvc_gform_dev replaces
 *   call gform(pC,gradPc,fluxlPc,x,sedge,e,ls,ie,nelcon,pedge,ix,
 *              numel,ndm,nen,nshared,1,ndf,2,2,2)
 * of src/Simple.f:601,635 (src/Pform.f:466-676) for the cells that take
 * cell_si_p_i_ggl (types 63/65/73/75, src/Celllib.f:127-157): the Green-Gauss
 * gradient over the neighbours' centres, with ghost cells across walls
 * (pedge 3: 0, 6: the cell's value, otherwise extrapolated with the cell's
 * old gradient; src/CellibPressure.f:432-463,611-650).  Arrays are the
 * reference's: x(2,nnode), ix(nen+1,numel), nelcon(nen,numel) (negative
 * across a wall), pedge(nen+1,numel), u(numel); grad(2,numel) holds the old
 * gradient on entry and the new one on return; fluxl(numel) = 1.  nen = 3 or
 * 4 (nshared = nen).  Bit-identical to the reference.  Indices are checked on
 * the device: a cell with one out of range gets NaN. */
int fvc_gform_dev(int32_t numel, int32_t nnode, int32_t nen, const double *x, const int32_t *ix,
                  const int32_t *nelcon, const int32_t *pedge, const double *u, double *grad,
                  double *fluxl, void *stream);

/* fvc_assemble_pc_dev replaces the pressure-correction assembly
 *   call pform(adPc,auPc,alPc,iaPc,jaPc,num,un,pC,p,p,gradPc,gradP,gradPc,
 *              div,fluxlPc,iM,rO,wP,bPc,rCPc,x,sedge,e,ddU,eddyVisc,ie,
 *              nelcon,pedge,ix,numel,ndm,nen,nshared,1,ndf,dt,0.0d0,
 *              matrizPc,unsymPc,1,2,2,.true.,.true.,.false.,.false.,.true.,bs)
 * of src/Simple.f:583-588 (src/Pform.f:65-436) for ELLPACK (matrizPc = 4,
 * the reference's default), cells of types 63/65/73/75 (cell_si_p_i_ggl,
 * src/CellibPressure.f:390-571, iws = 1).  int_vel is the /simple_/ intVel
 * (5 in src/Simple.f:186; 1 and 2 also taken).  Inputs as fvc_gform_dev plus
 * num(numel) (1-based equation numbers, <= 0: no equation), e(10,nmat)
 * (density in row 2, the material in ix(nen+1,:)), sedge(3,nen+1,numel)
 * (prescribed edge velocities), w(2,numel) (velocity), d(2,numel) (the
 * SIMPLE d field, ddU), gradp(2,numel) and p(numel) (the pressure and its
 * gradient).  Writes, for each cell c with an equation eq = num(c):
 * ad(eq), al(1:nen, eq) = -a_nb (0 across walls and to cells without an
 * equation; ellPack, src/EllPack.f:68-93) and b(eq); rescell(c) = the cell's
 * mass residual.  Rows of equations that no cell owns are not written.
 * Bit-identical to the reference.  Indices are checked on the device: a
 * cell with one out of range gets NaN. */
int fvc_assemble_pc_dev(int32_t numel, int32_t nnode, int32_t nen, int32_t neq, int32_t nmat,
                        int32_t int_vel, const double *x, const int32_t *ix, const int32_t *nelcon,
                        const int32_t *pedge, const int32_t *num, const double *e, const double *sedge,
                        const double *w, const double *d, const double *gradp, const double *p,
                        double *ad, double *al, double *b, double *rescell, void *stream);

/* fvc_assemble_mom_dev replaces the momentum assembly
 *   call pform(adU1,auU1,alU1,iaU1,jaU1,num,un,u1,p,p,gradU1,gradU2,gradP,
 *              div,fluxlU1,iM,rO,w,bU1,rCu1,x,sedge,e,ddU,eddyVisc,ie,
 *              nelcon,pedge,ix,numel,ndm,nen,nshared,1,ndf,dt,0.0d0,
 *              matrizU1,.true.,1,1,1,.true.,.true.,.true.,.true.,.true.,bs)
 * of src/Simple.f:425-431 (comp = 1; :456-462 for u2, comp = 2) for
 * ELLPACK and cells of types 63/65 (cell_si_eb_i_ggl,
 * src/CellibSimple.f:591-901, iws = 1).  under_u, s_pressure and simple_c
 * are the /simple_/ underU, sPressure and simpleC; bs the second-order
 * backward time term.  Inputs as fvc_assemble_pc_dev with e(10,nmat) =
 * (viscosity, density, ..., e(6): transient coefficient, e(7): TVD limiter
 * code 1..7), u(numel) the velocity component, pres(numel) the pressure,
 * w(2,numel) the velocity, grad1/grad2/gradp(2,numel) the gradients of u,
 * of the other component and of the pressure.  Writes ad, al, b by
 * equation as fvc_assemble_pc_dev, rescell(numel), im(comp,:) and
 * im(comp+2,:) of im(4,numel) (the momentum interpolation terms) and
 * d(comp,:) of d(2,numel) (the SIMPLE d field).  Bit-identical to the
 * reference.  A cell with an index out of range or a limiter code outside
 * 1..7 gets NaN. */
int fvc_assemble_mom_dev(int32_t numel, int32_t nnode, int32_t nen, int32_t neq, int32_t nmat, int32_t comp,
                         double dt, double under_u, int32_t s_pressure, int32_t simple_c, int32_t bs,
                         const double *x, const int32_t *ix, const int32_t *nelcon, const int32_t *pedge,
                         const int32_t *num, const double *e, const double *sedge, const double *u,
                         const double *pres, const double *w, const double *grad1, const double *grad2,
                         const double *gradp, double *d, double *ad, double *al, double *b, double *rescell,
                         double *im, void *stream);

/* fvc_simple_update_dev replaces
 *   subroutine simpleUpdate(w,pre,u1,u2,preC,gradPc,d,numel,ndm,underU,underPc,simpleR)
 * of src/Simple.f:871-898 (called at :642): w(:,i) = (u1(i),u2(i)) - d(:,i)*gradPc(:,i); and
 * unless simpleR (SIMPLEC), pre(i) = pre(i) + underPc*preC(i).  w, d and
 * gradPc are (2,numel); ndm must be 2.  Bit-identical to the reference. */
int fvc_simple_update_dev(int32_t numel, int32_t ndm, double *w, double *pre, const double *u1,
                          const double *u2, const double *prec, const double *gradpc, const double *d,
                          double under_u, double under_pc, int32_t simple_r, void *stream);

/* ----------------------------------------------------------------------------
 * SpMV y = A x.  Replaces the matvec_* routines of src/Matvec.f:21-300 and
 * src/Matvec_omp.f (matvec_csrd[_ilu2]_omp, matvec_csrd_sym_omp family,
 * matvec_ellpack_omp).  Each row is summed in the reference's sequential
 * order without FMA contraction, so results are bit-identical to the
 * reference's sequential routines. */
int fvc_matvec(int32_t matriz, int32_t unsym, int32_t neq, int32_t nad,
               const double *ad, const double *au, const double *al,
               const int32_t *iax, const int32_t *ja, const double *x, double *y);
int fvc_matvec_dev(int32_t matriz, int32_t unsym, int32_t neq, int32_t nad,
                   const double *ad, const double *au, const double *al,
                   const int32_t *iax, const int32_t *ja, const double *x, double *y,
                   void *stream);

/* Inner product a.b.  Replaces dot / dot_ompL4 (src/Matvec.f:383-402,
 * src/Matvec_omp.f:1247-1296).  Deterministic fixed-tree reduction: repeated
 * calls give bit-identical results.  *result is written on the host. */
int fvc_dot(const double *a, const double *b, int32_t n, double *result);
int fvc_dot_dev(const double *a, const double *b, int32_t n, double *result, void *stream);

/* getRes (src/Solver.f:368-391): flag != 0: x(num(i)) = u(i) (cells ->
 * equations); flag == 0: u(i) = x(num(i)).  num holds 1-based equation ids. */
int fvc_getres(double *u, double *x, const int32_t *num, int32_t neq, int32_t numel,
               int32_t flag);
int fvc_getres_dev(double *u, double *x, const int32_t *num, int32_t neq, int32_t numel,
                   int32_t flag, void *stream);

/* ----------------------------------------------------------------------------
 * Device memory helpers, so that callers without a GPU framework can keep
 * systems resident in HBM between solves. */
int fvc_malloc(void **ptr, size_t bytes);
int fvc_free(void *ptr);
int fvc_memcpy_htod(void *dst, const void *src, size_t bytes);
int fvc_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int fvc_memset_dev(void *dst, int value, size_t bytes);
/* Device-to-device copy by a plain streaming kernel (16-byte aligned
 * pointers, size a multiple of 16); also the bandwidth reference the
 * benchmark calibrates hardware byte counters against. */
int fvc_memcpy_dtod(void *dst, const void *src, size_t bytes, void *stream);
int fvc_synchronize(void);
/* Number of compute units and the device name (name may be NULL). */
int fvc_device_info(int32_t *compute_units, char *name, int32_t name_len);

/* Timing helpers: record an event on `stream`, and elapsed ms between two. */
int fvc_event_create(void **ev);
int fvc_event_destroy(void *ev);
int fvc_event_record(void *ev, void *stream);
int fvc_event_elapsed_ms(void *start, void *stop, float *ms);

#ifdef __cplusplus
}
#endif
#endif /* FVCELL_H */
