/*
 * momwire_hip.h -- C ABI of the MI355X (gfx950) B-spline port-solution path.
 *
 * Replaces, for BSplineSolver.compute_port_solution(), the reference's
 * pybind11 accelerator module (reference src/momwire/_accelerators.cpp) and
 * the LAPACK solve it calls through scipy (bspline.py:2971-2989).  Every entry
 * point takes plain pointers and sizes: device pointers for arrays (any HIP
 * allocation, e.g. torch tensors' data_ptr()), a hipStream_t passed as void*,
 * and returns an int status (0 = ok; see mw_last_error() for the message).
 * Work is enqueued asynchronously on the stream unless stated otherwise.
 *
 * Complex values are interleaved (re, im) pairs: double2 for the FP64 entry
 * points (z*), float2 for the FP32 ones (c*).  Matrices are row-major with an
 * explicit leading dimension counted in complex elements.  Z is complex
 * symmetric only when every wire has the same radius (the fill uses the
 * observer row's radius); with per-wire radii it is not, so a binding that
 * hands over Fortran-order (column-major) storage solves with Z^T.  Pass
 * row-major storage.
 *
 * Cancellation.  Entry points that take `cancel` say which side reads it:
 *  - HOST-polled (mw_?getrf, mw_?gesv, mw_?sysv, mw_sweep_run): the host
 *    reads the int32 between launches; any host int32 address works.
 *    mw_sweep_run's persistent forms (gemm_variant 2-5) run as one launch:
 *    they also read the flag on the device before each problem claim when
 *    it is device-accessible mapped host memory (mw_cancel_create); an
 *    ordinary host address is then read only before the launch.
 *  - DEVICE-polled (mw_somm_six_integrals, mw_somm_surfaces,
 *    mw_somm_surfaces_varied): the kernels read it while they run, so it
 *    must be device-accessible memory -- a flag from mw_cancel_create()
 *    (mapped pinned host memory) or device memory.  An ordinary host
 *    address is refused with MW_EINVAL before anything is launched.
 * A flag from mw_cancel_create() serves both kinds.  Reference:
 * _accelerators.cpp:79-113 (the polled flag), _cancel.py:20-79 (the token).
 */
#ifndef MOMWIRE_HIP_H
#define MOMWIRE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MW_OK = 0,
  MW_EINVAL = 1,
  MW_ESINGULAR = 2,
  MW_EHIP = 3,
  MW_EABORT = 4,
  MW_ERANGE = 5
};

/* Maximum segments one basis tile may touch (tiles hold <= 16 bases). */
#define MW_TSMAX 18
/* Basis tile edge (bases per tile side). */
#define MW_TB 16

/*
 * Device-resident, k-independent description of a meshed structure.  Built
 * once per geometry by the host (momwire_amd/_mesh.py) and reused for every
 * frequency.  All pointers are device pointers.
 */
typedef struct mw_mesh {
  int32_t n_seg;     /* segments over all wires                              */
  int32_t n_edge;    /* straight edges (uniform segmentation on each)         */
  int32_t n_basis;   /* B-spline unknowns                                     */
  int32_t n_tile;    /* basis tiles partitioning [0, n_basis)                 */
  int32_t n_se;      /* same-edge table entries: sum over edges of 2 n_e - 1  */
  int32_t degree;    /* B-spline degree, 1 or 2 (tables padded to 3 wings,
                        band slots and powers)                              */
  int32_t n_qp;      /* Gauss-Legendre points per segment per axis, 1 .. 8
                        (the batched sweep: 4)                              */
  int32_t ek;        /* 1: extended thin-wire kernel on coaxial equal-radius
                        pairs (labels below), 0: the reduced kernel
                        (reference _bspline_kernels.py:28-80, 198-322)     */
  const double* seg;       /* [n_seg][12] l.xyz r.xyz t.xyz length radius 0   */
  const int32_t* seg_edge; /* [n_seg][2]  edge id, index along the edge       */
  const double* edge;      /* [n_edge][2] segment length h, wire radius a     */
  const int32_t* edge_i;   /* [n_edge][2] n_e, offset of the edge's table     */
  const int32_t* se_edge;  /* [n_se] edge owning each same-edge table entry   */
  void* se_tab;            /* [n_se][9] double2: static + smooth same-edge
                              moments J_pq(j - i), written by
                              mw_same_edge_table()                            */
  const double* band_poly; /* [n_seg][3 slot][3 power] polynomial of the basis
                              occupying each band slot of a segment           */
  const int32_t* bw_loc;   /* [n_basis][3] wing segment's index in the tile's
                              segment list, -1 for an absent wing             */
  const int32_t* bw_slot;  /* [n_basis][3] band slot of the basis on that
                              segment                                         */
  const int32_t* tile_b0;  /* [n_tile+1] first basis of each tile            */
  const int32_t* tile_seg; /* [n_tile][MW_TSMAX] sorted segment list          */
  const int32_t* tile_nseg;/* [n_tile]                                        */
  const double* gl;        /* [2 n_qp] t, w on [0,1]; then [32] 16-point
                              Gauss-Legendre x, w on [-1,1] (far static rule) */
  const int32_t* ek_label; /* [n_seg] coaxial labels (ek): a pair with equal
                              labels takes the extended kernel; NULL if !ek  */
  const int32_t* ek_img_obs; /* [n_seg] labels of the observers and of the  */
  const int32_t* ek_img_src; /* mirrored sources in the ground image block,
                              from one scan over both (NULL if !ek or no
                              ground)                                        */
  const double* se_static; /* [n_se][9] k-independent static part of the
                              same-edge table (mw_same_edge_static), or NULL:
                              mw_same_edge_table then computes it per call  */
} mw_mesh;

/*
 * One problem of a batched sweep: a structure (mesh index) at one wavenumber.
 * A is the problem's [Z | B] buffer (n x (n + ncols) complex double, row
 * stride lda); the port columns and KCL rows come from the mesh's port
 * tables.  All pointers are device pointers.
 */
typedef struct mw_sweep_prob {
  int32_t mesh;           /* index into the mesh / port tables               */
  int32_t symmetric;      /* every wire of the mesh shares one radius        */
  double k, omega;        /* wavenumber, angular frequency                   */
  void* A;                /* [Z | B], written by the fill and the solve      */
  int64_t lda;
  int64_t n;              /* unknowns                                        */
  int64_t ncols;          /* port columns + KCL constraint columns           */
  void* se_tab;           /* same-edge table of this (mesh, k): [n_se][9]    */
  const int32_t* tile_pairs;  /* the mesh's (ti, tj) list                     */
  int32_t n_pairs;
  int32_t n_se;
  int32_t* ipiv;          /* [n]                                             */
  void* work;             /* mw_getrf_workspace_bytes(n) bytes               */
  int32_t* info;          /* [2]: factorisation, KCL Schur                   */
  void* Y;                /* [n_ports][n_ports] output                       */
  void* S;                /* [ncon][ncon] scratch                            */
  void* T;                /* [ncon][n_ports] scratch                         */
} mw_sweep_prob;

/* Sparse port columns and KCL rows of one mesh (device pointers). */
typedef struct mw_ports {
  int32_t n_ports, n_con;
  int32_t col_max, con_max, port_max, pad;
  const int32_t* col_idx;  /* [n_ports + n_con][col_max]: right-hand sides  */
  const double* col_val;
  const int32_t* con_idx;  /* [n_con][con_max]: KCL constraint rows         */
  const double* con_val;
  const int32_t* port_idx; /* [n_ports][port_max]: readout rows (= columns) */
  const double* port_val;
} mw_ports;

typedef struct mw_sweep_plan mw_sweep_plan;

/*
 * A leaf block of the block-low-rank (H-matrix) operator.  Its rows and
 * columns are cluster members (basis ids, ascending) listed from row_off /
 * col_off in two index pools.  Storage sits in one complex double store at
 * data_off: a dense block as D (m x n, row-major); a compressed block as U^T
 * (cap x m) then V (cap x n), U V ~ the block.
 */
typedef struct mw_hblock {
  int32_t m, n;        /* rows, columns                                       */
  int32_t rank;        /* compressed: rank of U V; -1: stored dense           */
  int32_t cap;         /* compressed: factor capacity in ranks; dense: 0      */
  int64_t row_off;     /* first row member in the row pool                    */
  int64_t col_off;     /* first column member in the column pool              */
  int64_t data_off;    /* complex offset of the block's storage in the store  */
  int64_t aux_off;     /* mw_hm_aca: byte offset of its m + n used flags in the
                          workspace; mw_hm_invert: complex offset of its
                          2 m^2 scratch in the workspace                      */
  int64_t t_off;       /* compressed: first row of V X[cols] in mw_hm_apply's
                          scratch T (rows of nrhs complex)                    */
  int64_t row_start;   /* leaf-order position of the first row (the cluster
                          tree's leaves one after another: every cluster is
                          one contiguous range)                               */
  int64_t col_start;   /* leaf-order position of the first column            */
} mw_hblock;

/* mw_hm_apply work items: 64 V rows of the compressed blocks that share a
   column cluster (their (block, rank) pairs from t_rows[row0]); k columns
   whose leaf-order positions are colpos[col_off ..]. */
typedef struct mw_hm_titem {
  int32_t row0, nrows, k, pad;
  int64_t col_off;
} mw_hm_titem;
/* mw_hm_apply work items: up to 32 leaf-order rows from row0, summed over
   the (block, row offset) pairs r_ent[ent0 .. ent0 + nent). */
typedef struct mw_hm_ritem {
  int64_t row0;
  int32_t nrows, ent0, nent, pad;
} mw_hm_ritem;

/* Last error message of the calling thread ("" if none). */
const char* mw_last_error(void);
/* ABI version of this library. */
int mw_abi_version(void);

/*
 * A cancel flag both the host and the kernels can read: one int32 (zeroed)
 * in mapped pinned host memory.  Write 1 through *flag to cancel, 0 to
 * reset.  Replaces the token storage of CancelToken (reference
 * src/momwire/_cancel.py:20-79, handed to the accelerators as an address,
 * _accelerators.cpp:79-113).  mw_cancel_destroy frees it (NULL is a no-op).
 */
int mw_cancel_create(int32_t** flag);
int mw_cancel_destroy(int32_t* flag);

/*
 * Same-edge moment table: J_pq(delta) = static (closed form, 1/sqrt kernel)
 * + smooth remainder (exp(-jkR)-1)/(4 pi R) for every edge and offset delta.
 * Replaces seg_seg_static_moments_bspline_uniform (_accelerators.cpp:2822-2951)
 * and seg_seg_reg_moments_bspline_swept (_accelerators.cpp:165-327), which the
 * reference calls per edge (_bspline_kernels.py:333-517).
 */
int mw_same_edge_table(const mw_mesh* mesh, double k, void* stream);

/*
 * The k-independent half of the same-edge table, once per mesh: the static
 * closed-form moments h^{p+q+1} I(p, q, delta; a/h) / (4 pi) of every entry,
 * out: [n_se][9] doubles (device).  Point mesh->se_static at it afterwards and
 * every mw_same_edge_table call adds only the k-dependent smooth remainder.
 * Replaces the reference's per-edge seg_seg_static_moments_bspline_uniform
 * (_accelerators.cpp:2822-2951), which its swept path also computes once per
 * geometry (_bspline_kernels.py:333-422).
 */
int mw_same_edge_static(const mw_mesh* mesh, double* out, void* stream);

/*
 * Fused Galerkin fill Z[m][n] for all basis pairs, free space.  Replaces
 * seg_seg_full_moments_bspline (_accelerators.cpp:349-542) followed by
 * assemble_Z_bspline (_accelerators.cpp:1196-1309) or its windowed twin
 * (_accelerators.cpp:1312-1476), as driven by bspline.py:2222-2550.
 * tile_pairs: [n_pairs][2] (ti, tj) device array; symmetric != 0 means only
 * ti <= tj is listed and the transposed tile is written too (valid when every
 * wire shares one radius).  Z: n_basis x n_basis double2, row stride ldz.
 * Needs mw_same_edge_table() for the same k first.
 */
int mw_fill_z(const mw_mesh* mesh, double k, double omega, double eps, double mu,
              void* Z, int64_t ldz, const int32_t* tile_pairs, int32_t n_pairs,
              int32_t symmetric, void* stream);

/*
 * mw_fill_z with the far-interaction kernel in FP32: segment pairs whose
 * midpoints are more than near_ratio x the longer segment apart take
 * exp(-jkR)/(4 pi R) with its phase reduced in FP64 and sin, cos and the
 * reciprocal in single precision; the near pairs, the same-edge table and the
 * extended-kernel pairs stay FP64.  The mixed form's optional fill; no
 * counterpart in the reference (which fills every pair in FP64,
 * _accelerators.cpp:370-542).
 */
int mw_fill_z_far32(const mw_mesh* mesh, double k, double omega, double eps, double mu,
                    void* Z, int64_t ldz, const int32_t* tile_pairs, int32_t n_pairs,
                    int32_t symmetric, double near_ratio, void* stream);

/*
 * LU factorisation with row pivoting (threshold pivoting per 32-column panel,
 * tournament partial pivoting where the threshold fails), in place:
 * A = P L U, L unit lower.  Replaces LAPACK zgetrf/cgetrf behind
 * scipy.linalg.solve (bspline.py:2971-2989).  ipiv: [n] device int32,
 * LAPACK-style sequential row interchanges (0-based).  work: device scratch of
 * mw_getrf_workspace_bytes(n) bytes.  info: device int32, set to 1 + the
 * first zero-pivot column (0 if none).  cancel: HOST-polled int32 read
 * between panels (may be NULL); a set flag stops the factorisation with
 * MW_EABORT.
 */
int64_t mw_getrf_workspace_bytes(int64_t n);

/*
 * Threshold pivoting: a panel keeps its diagonal pivots when every multiplier
 * of the panel satisfies |re| + |im| <= bound (checked on the GPU); otherwise
 * the panel is refactored with tournament partial pivoting.  bound <= 0
 * pivots every panel.  Default 8.  The setting belongs to the calling
 * thread: it applies to the factorisations that thread starts afterwards,
 * and threads driving other solvers keep their own (read once per call).
 */
int mw_set_pivot_threshold(double bound);
double mw_get_pivot_threshold(void);
int mw_zgetrf(int64_t n, void* A, int64_t lda, int32_t* ipiv, void* work,
              int32_t* info, const volatile int32_t* cancel, void* stream);
int mw_cgetrf(int64_t n, void* A, int64_t lda, int32_t* ipiv, void* work,
              int32_t* info, const volatile int32_t* cancel, void* stream);

/*
 * Multi-RHS solve with the factors of mw_?getrf: X <- A^{-1} X for nrhs
 * columns (X: n x nrhs, row stride ldx).  work: the same workspace the
 * factorisation wrote (it keeps the inverted 32x32 diagonal blocks).
 * ipiv NULL: factors without row exchanges (mw_?sysv's).
 * Replaces LAPACK zgetrs.
 */
int mw_zgetrs(int64_t n, int64_t nrhs, const void* A, int64_t lda,
              const int32_t* ipiv, const void* work, void* X, int64_t ldx,
              void* stream);
int mw_cgetrs(int64_t n, int64_t nrhs, const void* A, int64_t lda,
              const int32_t* ipiv, const void* work, void* X, int64_t ldx,
              void* stream);

/*
 * Factor and solve in one pass (LAPACK ?gesv, which scipy.linalg.solve calls
 * in bspline.py:2971-2989): A is n x (n + nrhs), row stride lda >= n + nrhs;
 * columns [n, n + nrhs) hold the right-hand sides on entry and the solution
 * on return.  The right-hand sides ride through the factorisation as extra
 * columns, so only the backward substitution runs afterwards.  ipiv, work,
 * info and cancel as for mw_?getrf; the factors stay in A[:, :n] and work
 * for later mw_?getrs calls.
 */
int mw_zgesv(int64_t n, int64_t nrhs, void* A, int64_t lda, int32_t* ipiv,
             void* work, int32_t* info, const volatile int32_t* cancel,
             void* stream);
int mw_cgesv(int64_t n, int64_t nrhs, void* A, int64_t lda, int32_t* ipiv,
             void* work, int32_t* info, const volatile int32_t* cancel,
             void* stream);

/*
 * mw_?gesv for a complex SYMMETRIC A (A[:, :n] = A[:, :n]^T exactly: the
 * Galerkin Z when every wire has the same radius), factored WITHOUT row
 * exchanges, so U = D L^T: past n = 4096 each 512-column block's U rows come
 * from its L columns by a transpose-scale and the trailing update covers only
 * the blocks on or below the block diagonal (half the factorisation's work).
 * The inverses of the 512 x 512 diagonal blocks of L and U are then formed
 * (recursive doubling from the 32 x 32 panel inverses) and the right-hand
 * sides solved by one block product per 512 rows in each direction.  Same
 * arguments and outputs as mw_?gesv except work, which holds
 * mw_sysv_workspace_bytes(n) bytes (the block inverses stay there for
 * mw_?sytrs; ipiv is returned as the identity).
 * first_row: NULL, or a HOST array of nrhs int32 giving the first row at
 * which each right-hand side may be nonzero (its rows above are zero on
 * entry: a port's drive column starts at its feed row); the forward pass
 * then skips, per 512-row block, the columns that have not started by the
 * block's end, which takes a third of its work on an array whose ports are
 * numbered in feed-row order.  Read during the call.
 * When a panel's threshold test fails (a multiplier beyond the bound of
 * mw_set_pivot_threshold, or a zero pivot) the panel is NOT pivoted and
 * info[0] = -1 on return: the result is not valid and the caller refactors
 * the original matrix with mw_?gesv.
 * n <= 4096: one level, the look-ahead chain of mw_?gesv without row
 * exchanges, its trailing update on the lower part and the panel-diagonal
 * blocks only (the U12 rows formed as D L21^T); first_row is not used and
 * work is mw_getrf_workspace_bytes(n).  The reference has no counterpart:
 * it solves Z with scipy.linalg.solve (bspline.py:2971-2989) whatever its
 * symmetry.
 */
int mw_zsysv(int64_t n, int64_t nrhs, void* A, int64_t lda, int32_t* ipiv,
             void* work, int32_t* info, const volatile int32_t* cancel,
             const int32_t* first_row, void* stream);
int mw_csysv(int64_t n, int64_t nrhs, void* A, int64_t lda, int32_t* ipiv,
             void* work, int32_t* info, const volatile int32_t* cancel,
             const int32_t* first_row, void* stream);

/*
 * Workspace of mw_?sysv in bytes: mw_getrf_workspace_bytes(n) plus, past
 * n = 4096, the 512 x 512 diagonal-block inverses of L and U and a
 * 512 x 1024 staging block (2 x 8192 n + 8 MiB more).
 */
int64_t mw_sysv_workspace_bytes(int64_t n);

/*
 * X <- A^{-1} X with the factors and workspace of mw_?sysv (X: n x nrhs,
 * row stride ldx): past n = 4096 one block product per 512 rows each way;
 * n <= 4096 (one level), mw_?getrs with ipiv (the identity).  No
 * counterpart in the reference (see mw_?sysv).
 */
int mw_zsytrs(int64_t n, int64_t nrhs, const void* A, int64_t lda,
              const int32_t* ipiv, const void* work, void* X, int64_t ldx,
              void* stream);
int mw_csytrs(int64_t n, int64_t nrhs, const void* A, int64_t lda,
              const int32_t* ipiv, const void* work, void* X, int64_t ldx,
              void* stream);

/*
 * mw_zgesv without the workspace argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference: it
 * exists for those callers only.  The LU workspace is a static device array
 * of 17 MiB per GPU (held from the first call on that GPU until the process
 * ends), so n <= 16384; larger n returns MW_ERANGE.  Everything else as
 * mw_zgesv (ipiv and info are outputs the caller passes).  The workspace
 * keeps the last call's contents and is shared by every call on the GPU:
 * calls that overlap on one GPU (two streams or two threads) race, and the
 * factors cannot be reused by mw_zgetrs afterwards.  Allocates nothing.
 */
int mw_zgesv_noalloc(int64_t n, int64_t nrhs, void* A, int64_t lda, int32_t* ipiv,
                     int32_t* info, const volatile int32_t* cancel, void* stream);

/* C <- C - A B (complex, MFMA).  A: m x k, B: k x n, C: m x n. */
int mw_zgemm_sub(int64_t m, int64_t n, int64_t k, const void* A, int64_t lda,
                 const void* B, int64_t ldb, void* C, int64_t ldc, void* stream);
int mw_cgemm_sub(int64_t m, int64_t n, int64_t k, const void* A, int64_t lda,
                 const void* B, int64_t ldb, void* C, int64_t ldc, void* stream);

/*
 * C <- C + A^T B (complex, MFMA; A stored k x m, plain transpose).  The
 * adjoint-corrected port readout of the mixed form, Y + X^T R with
 * R = B - Z X: Z is complex symmetric, so its error is quadratic in X's.
 */
int mw_zgemm_tn_add(int64_t m, int64_t n, int64_t k, const void* A, int64_t lda,
                    const void* B, int64_t ldb, void* C, int64_t ldc, void* stream);
/* The same in FP32 (complex64): the mixed form's C^T R, a term ~1e-6 of Y,
 * for which single precision leaves ~1e-11 of Y. */
int mw_cgemm_tn_add(int64_t m, int64_t n, int64_t k, const void* A, int64_t lda,
                    const void* B, int64_t ldb, void* C, int64_t ldc, void* stream);

/*
 * out[i][j] = sum_t val[i][t] * X[idx[i][t]][j]  (sparse rows times X).
 * The port readout Y = B^T X of bspline.py:4178-4179 (B's columns are sparse:
 * a delta-gap drive touches d+1 bases) and the KCL products A X of
 * bspline.py:2981-2989.  idx < 0 marks padding.
 */
int mw_zsparse_rows(int64_t nrows, int32_t maxnz, const int32_t* idx,
                    const double* val, const void* X, int64_t ldx, int64_t ncols,
                    void* out, int64_t ldo, void* stream);

/* X[:, :] = 0, then X[idx[c][t]][c] = val[c][t]: dense right-hand sides. */
int mw_zscatter_cols(int64_t n, int64_t ncols, int32_t maxnz, const int32_t* idx,
                     const double* val, void* X, int64_t ldx, void* stream);

/*
 * Block-low-rank operator of Z (H-matrix / ACA), free space.  Replaces the
 * reference's HMatrixSolver operator build and apply (hmatrix.py:982-1083,
 * _aca.py:137-248) and its off-edge block assembler for ACA sampling
 * (_accelerators.cpp:1949-2208).  wing_seg / wing_slot: [n_basis][3]
 * device tables, the global segment and band slot of every wing of every
 * basis (-1: absent wing).
 *
 * mw_zentries: out[t] = Z[rows[t]][cols[t]], t < n.  same_edge_table != 0
 * takes same-edge segment pairs from the mesh's same-edge table (the dense
 * fill's entries; call mw_same_edge_table for this k first), 0 integrates
 * every pair by quadrature (the admissible blocks' off-edge kernel,
 * hmatrix.py:1097-1155).
 */
int mw_zentries(const mw_mesh* mesh, const int32_t* wing_seg, const int32_t* wing_slot, double k,
                double omega, double eps, double mu, const int32_t* rows, const int32_t* cols,
                int64_t n, int32_t same_edge_table, void* out, void* stream);
/*
 * mw_hm_fill_dense: D of the dense blocks, work items[t] = (block index << 32
 * | first entry), 256 row-major entries each (near blocks, hmatrix.py:1010-
 * 1022, with same_edge_table = 1; admissible blocks that did not compress,
 * hmatrix.py:1054-1058, with 0).
 */
int mw_hm_fill_dense(const mw_mesh* mesh, const int32_t* wing_seg, const int32_t* wing_slot,
                     double k, double omega, double eps, double mu, const int32_t* rows_pool,
                     const int32_t* cols_pool, const mw_hblock* blocks, const int64_t* items,
                     int64_t n_items, int32_t same_edge_table, void* store, void* stream);
/*
 * mw_hm_aca: partial-pivoted ACA of the blocks list[0 .. n_list), one
 * workgroup each, all at once: the reference's aca_partial (_aca.py:137-211)
 * with its pivoting, hop and stopping rules at tolerance tol.  Writes U^T, V
 * and rank into each block (rank -1: the factors would not compress, or the
 * rank outgrew cap; the caller then stores the block dense).  work: the used
 * flags (aux_off bytes per block).  list[0 .. n_large) run in 512-thread
 * workgroups (large blocks), the rest in 256-thread ones.
 */
int mw_hm_aca(const mw_mesh* mesh, const int32_t* wing_seg, const int32_t* wing_slot, double k,
              double omega, double eps, double mu, const int32_t* rows_pool,
              const int32_t* cols_pool, mw_hblock* blocks, const int32_t* list, int32_t n_list,
              int32_t n_large, double tol, void* store, void* work, void* stream);
/*
 * mw_hm_apply: Y = H X for nrhs columns in leaf order (X, Y: n x nrhs, row
 * strides ldx, ldy), the batched product of the reference's HMatrix.matmat
 * (_aca.py:238-248).  Dense blocks are stored in leaf order; compressed
 * blocks keep the ACA's member order: sigma[row_off + t] is the position of
 * leaf-order row t among the block's rows, colpos[col_off + j] the
 * leaf-order position of column member j.  Phase 1 forms T = V X[cols]
 * (t_items; T: sum of ranks x nrhs, rows from t_off), phase 2 the rows of Y
 * (r_items), each row's terms summed in block order.
 */
int mw_hm_apply(const mw_hblock* blocks, const void* store, const mw_hm_titem* t_items,
                int32_t n_t, const int32_t* t_rows, const int32_t* colpos,
                const mw_hm_ritem* r_items, int32_t n_r, const int32_t* r_ent,
                const int32_t* sigma, const void* X, int64_t ldx, int32_t nrhs, void* T,
                void* Y, int64_t ldy, void* stream);
/*
 * mw_hm_invert: each square dense block of list replaced by its inverse
 * (Gauss-Jordan, partial pivoting; info[t] = 1 for an exactly singular
 * block, left unchanged).  The leaf diagonal blocks' inverses precondition
 * the Krylov solve (block Jacobi).  No counterpart in the reference, which
 * factors the whole near field with a sparse LU (hmatrix.py:95-112,
 * 1468-1493).
 */
int mw_hm_invert(const mw_hblock* blocks, const int32_t* list, int32_t n_list, void* store,
                 void* work, int32_t* info, void* stream);

/*
 * Small dense solve T <- S^{-1} T in one workgroup (S: ns x ns, T: ns x nt,
 * ns, nt <= 64): the Schur-complement multiplier solve of bspline.py:2988.
 */
int mw_zsmall_solve(int32_t ns, int32_t nt, void* S, int64_t lds, void* T,
                    int64_t ldt, int32_t* info, void* stream);

/*
 * Batched sweep over many (structure, wavenumber) problems of uneven size:
 * plan once (copies the descriptors to the device, sorts problems by size
 * and fixes the launch schedule), then run: the same-edge tables and fills
 * of every problem, the right-hand sides, the factorisations with the
 * forward solve riding along, the back substitutions, the KCL Schur steps
 * and the Y readouts -- a fixed sequence of batched launches (about 4 per
 * 32-column panel of the largest problem), so the whole sweep can be
 * captured as one HIP graph.  No counterpart in the reference, which sweeps
 * one structure at a time (bspline.py:4181-4439); per problem it computes
 * what compute_port_solution computes.  gemm_variant (chosen by the
 * autotuner): 0 = per-step launches with 64x32 update tiles, 1 = the same
 * with 128x64 tiles; 2..5 = after the fill, ONE persistent launch in which a
 * workgroup takes one problem at a time (largest first) through its whole
 * solve -- 2: one-level blocking (a K = 32 update per 32-column panel),
 * 3: two levels (128-column outer blocks, one K = 128 update of the rest
 * per block), both with two workgroups per compute unit; 4, 5: the same
 * with one per compute unit.  No workgroup of that launch
 * waits for another.  Its cancel flag is read by the kernel when it is
 * device-readable (mw_cancel_create): problems claimed after it is set are
 * left unsolved with info[0] = -1.  An ordinary host flag is read once,
 * before the launches.  pivot_bound: the plan's own threshold-pivoting multiplier bound, meant as
 * in mw_set_pivot_threshold (<= 0 pivots every panel).  32 (tau = 1/32)
 * is the Python default: the unpivoted multipliers of Galerkin wire
 * matrices stay in the tens, and a panel that falls back to the tournament
 * holds up the next step of every problem in the batch.
 */
int mw_sweep_plan_create(int32_t n_mesh, const mw_mesh* meshes, const mw_ports* ports,
                         int32_t nprob, const mw_sweep_prob* probs, int32_t gemm_variant,
                         double pivot_bound, mw_sweep_plan** plan);
int mw_sweep_run(mw_sweep_plan* plan, double eps, double mu,
                 const volatile int32_t* cancel, void* stream);
/*
 * The fill half of mw_sweep_run alone: every problem's Z into the first n
 * columns of its A (what mw_fill_z writes for that mesh and wavenumber,
 * same-edge tables included), for callers that factor the matrices
 * elsewhere.  Problems of one mesh whose wavenumbers step evenly share their
 * geometry: one workgroup per tile pair walks up to 16 of them.  No
 * counterpart in the reference, which fills one structure at one wavenumber
 * at a time (bspline.py:2222-2550).
 */
int mw_sweep_fill(mw_sweep_plan* plan, double eps, double mu, void* stream);
int mw_sweep_plan_destroy(mw_sweep_plan* plan);

/*
 * Precision changes for the mixed form (FP32 factorisation, FP64 residual):
 * dst <- (float) src, and dst (+)= (double) src, over a rows x cols block.
 */
int mw_z2c(int64_t rows, int64_t cols, const void* src, int64_t lds, void* dst,
           int64_t ldd, void* stream);
int mw_c2z(int64_t rows, int64_t cols, const void* src, int64_t lds, void* dst,
           int64_t ldd, int32_t accumulate, void* stream);

/*
 * The mixed form's one-residual port readout and its check.  Inputs, all
 * FP64 complex and row-major on the device: X0 (n x p, the FP32 solve of
 * Z X = B), C (n x p, the FP32 solve of Z C = R with R = B - Z X0 in FP64),
 * T = C^T R (p x p) and Y = B^T X0 + X0^T R (p x p).  On return
 * Y <- Y + T, the readout B^T X0 + X0^T R + C^T R of B^T Z^{-1} B for
 * Z = Z^T, whose error is (Z^{-1} R - C)^T R; stats (4 doubles, device)
 * = {delta, max|T|, max|Y|, delta max|T|} with delta = max_j |C_j| / |X0_j|
 * (the FP32 solve's relative error, which C shares), and gate (1 int32,
 * device) = 0 when delta max|T| <= tol max|Y|, else 1: the caller then
 * reads out again from a fresh FP64 residual.  scratch: 16 (p + 1) bytes
 * of device memory (cleared here).  No counterpart in the reference, which
 * solves in FP64 only (bspline.py:2971-2989).
 */
int mw_zreadout_check(int64_t n, int64_t p, const void* X0, int64_t ldx, const void* C,
                      int64_t ldc, const void* T, int64_t ldt, void* Y, int64_t ldy,
                      double tol, void* scratch, double* stats, int32_t* gate,
                      void* stream);

/* ------------------------------------------------------------------------
 * Sommerfeld ground (ground_model="sommerfeld")
 * ------------------------------------------------------------------------ */

/*
 * The tabulated remainder surfaces I_rho^V, I_z^V, I_rho^H, I_phi^H of one
 * ground: 4 or 6 (R1, theta) regions, each a uniform n_r x n_th lattice
 * starting at (r0, th0) with spacings (dr, dth).  vals holds, region after
 * region, [n_r][n_th][4] interleaved complex values; off[g] is region g's
 * first node.  Lengths are physical (the lattice of a normalised master
 * scaled by lambda); every interpolated value is multiplied by value_scale
 * (the omega*mu ratio that rescales a normalised master).  Regions: 0/1 are
 * R1 <= r_break, 2/3 up to r_near, 4/5 beyond; even ones theta <= th_split.
 * Mirrors SommerfeldGrid (reference src/momwire/_sommerfeld.py:642-914) and
 * its flattened form grid_cpp_args (:917-937).
 */
typedef struct mw_somm_grid {
  int32_t n_reg, pad;
  double r1_max, r_break, th_split, r_near, value_scale;
  double r0[6], dr[6], th0[6], dth[6];
  int32_t n_r[6], n_th[6];
  int64_t off[6];
  const double* vals;
} mw_somm_grid;

/*
 * The six lambda-integrals (V_rr, V_zz, V_rz, V_r1, V, U) of NEC eqs 148-153 at
 * n (rho, h) nodes (device arrays), out: [n][6] complex.  form 0 = automatic
 * contour (Bessel for rho < 2h), 1 = Bessel, 2 = Hankel (needs rho > 0).
 * Replaces somm_six_integrals_batch (reference _accelerators.cpp:5771-5823,
 * python twin _sommerfeld.py:468-502).  Nodes must satisfy rho, h >= 0 and
 * rho + h > 0 (not checked on the device: the arrays are device memory).
 * cancel: DEVICE-polled (see the top of this file), may be NULL; a node
 * whose kernel sees it set stops early and leaves its output undefined, so
 * the caller checks the flag after the stream has finished.
 */
int mw_somm_six_integrals(double eps_re, double eps_im, double k2, const double* rho,
                          const double* h, int64_t n, double rtol, int32_t form, void* out,
                          const volatile int32_t* cancel, void* stream);

/*
 * The four surfaces at n (R1, theta) nodes, out: [n][4] complex (R1 = 0 takes
 * the limits of eqs 169-172).  Replaces iv_surfaces_direct
 * (reference _sommerfeld.py:535-584), the grid's fill function.  cancel as
 * for mw_somm_six_integrals (DEVICE-polled).
 */
int mw_somm_surfaces(double eps_re, double eps_im, double k2, double omega, double mu,
                     const double* R1, const double* theta, int64_t n, double rtol,
                     void* out, const volatile int32_t* cancel, void* stream);

/*
 * 4x4 Lagrange interpolation of a grid at n points, out: [n][4] complex.
 * Replaces SommerfeldGrid.eval (reference _sommerfeld.py:831-876).
 */
int mw_somm_eval(const mw_somm_grid* grid, const double* R1, const double* theta, int64_t n,
                 void* out, void* stream);

/*
 * Ground remainder of E at points in space for p sets of source moments:
 * the reflected-wave combination of eqs 143-147 kept as a vector and summed
 * over the sources, E_rem(point) = sum_s F(point, src_s) . p_s.  Replaces
 * reflected_field_at's ("above", "above") regime (reference
 * src/momwire/_field_point.py:292-331, with _reflected_wave_surfaces and
 * _reflected_wave_combination at :145-233), batched over port columns.
 * pts [n_pts][3], mid [n_src][3] (device); moment [n_src][ldm] complex with
 * column c of component d at d * p + c; out [n_pts][ldo] likewise, written
 * (accumulate = 0) or added to.  Only the remainder: the direct and
 * C2-scaled image terms are the caller's, as in the reference.
 */
int mw_ground_fields(const mw_somm_grid* grid, double ground_z, double k, int32_t n_pts,
                     const double* pts, int32_t n_src, const double* mid, const void* moment,
                     int64_t ldm, int32_t p, int32_t accumulate, void* out, int64_t ldo,
                     void* stream);

/*
 * Grounded fill: Z -= C2 * Z_image + Q, accumulated into a Z already holding
 * the free-space matrix (row-major, ld ldz).  Z_image is the exact image
 * (mirrored sources, tangent dot t_m . (t_x, t_y, -t_z), charge weight 1);
 * Q the Galerkin remainder block of the tabulated field (n_qp_sommerfeld-
 * point Gauss per segment, 1 .. 8; 3 is the reference's default).  grid =
 * NULL drops Q (image only).  Same tiling and symmetric
 * tile-pair list as mw_fill_z.  Replaces the weighted image assembly and the
 * fused remainder kernel: assemble_Z_bspline_weighted_windowed and
 * sommerfeld_remainder_bspline_Q (reference _accelerators.cpp:1498-1608,
 * 6081-6253; callers bspline.py:2552-2661, 1791-1874, 3730-3770).
 */
int mw_fill_ground(const mw_mesh* mesh, double k, double omega, double eps, double mu,
                   double ground_z, double c2_re, double c2_im, const mw_somm_grid* grid,
                   int32_t n_qp_sommerfeld, void* Z, int64_t ldz, const int32_t* tile_pairs,
                   int32_t n_pairs, int32_t symmetric, void* stream);

/*
 * Reflection-coefficient finite ground (NEC gn 0 style): Z -= Z_image with
 * per-segment-pair Fresnel weights -- the A term by
 * w_A = rho_v (t_m . M t_n - P) - rho_h P and the charge term by
 * w_Phi = c0 + c1 rho_v, rho_v and rho_h taken at the specular angle of
 * the ray from the source segment's image midpoint to the observer's
 * midpoint for eps~ = eps_t (P: the tangents' products with the horizontal
 * normal of the plane of incidence).  (c0, c1) select the charge weighting
 * (reference _ground_refl.phi_mode_coeffs).  The weights are formed in the
 * kernel; no Q term.  Replaces the weighted image assembly with Fresnel
 * tables: assemble_Z_bspline_weighted(_windowed) (reference
 * _accelerators.cpp:1790-1917; weights _ground_refl.py:65-289; caller
 * bspline.py:1654-1740).
 */
int mw_fill_ground_refl(const mw_mesh* mesh, double k, double omega, double eps, double mu,
                        double ground_z, double eps_t_re, double eps_t_im, double phi_c0_re,
                        double phi_c0_im, double phi_c1_re, double phi_c1_im, void* Z,
                        int64_t ldz, const int32_t* tile_pairs, int32_t n_pairs,
                        int32_t symmetric, void* stream);

/*
 * The same surfaces with a complex permittivity per node (eps: [n][2]):
 * the sample generator of the learned surrogate.  No counterpart in the
 * reference (its engine takes one ground per call, _sommerfeld.py:468-502).
 * cancel: DEVICE-polled, as for mw_somm_six_integrals.
 */
int mw_somm_surfaces_varied(const double* eps, double k2, double omega, double mu,
                            const double* R1, const double* theta, int64_t n, double rtol,
                            void* out, const volatile int32_t* cancel, void* stream);

/*
 * Z[rows[e]][cols[e]] += vals[e] (complex), e < nnz, for distinct pairs with
 * every index < n: the distributed wire loading term sum_w Z'_w(omega) S_w
 * (series impedance per unit length times the basis Gram on wire w), merged
 * per frequency by the caller.  Replaces BSplineSolver._apply_loading
 * (reference bspline.py:2732-2753; Gram :2667-2730; Z'_w
 * _wire_loading.py:88-178).
 */
int mw_zadd_sparse(int64_t n, int64_t nnz, const int32_t* rows, const int32_t* cols,
                   const void* vals, void* Z, int64_t ldz, void* stream);

/* ------------------------------------------------------------------------
 * Readout of solved port columns: every column at once.
 *
 * Complex buffers are interleaved double2, row-major.  A structure's
 * readout mesh has n_el straight elements (knot to knot along each wire)
 * and n_node knots (n_el + n_wires); its tables are built once on the host.
 * ------------------------------------------------------------------------ */

/*
 * Element current moments m_e = 0.5 (I_a + I_b) dl_e (mom [n_el][3][p],
 * component stride p, row stride ldm >= 3p) and continuity steps
 * delta_n = el_in - el_out (delta [n_node][p], ldd >= p; el = 0 past a wire
 * end) of the solution columns X [n_b][p] (ldx >= p).  The current at sample
 * s is sum_t tap_v[s][t] X[tap_i[s][t]] over three taps (tap_i = -1: none,
 * every other index < n_b); el_tab [n_el][2] holds each element's end
 * samples, dl [n_el][3] its vector length, node_tab [n_node][2] the element
 * arriving at and leaving each knot (-1: none).  Replaces the per-column
 * currents_at_knots + element_currents walk (reference bspline.py:4523-4685,
 * _element_currents.py:33-120).
 */
int mw_element_moments(int32_t n_el, int32_t n_node, int32_t p, const int32_t* tap_i,
                       const double* tap_v, const int32_t* el_tab, const double* dl,
                       const int32_t* node_tab, const void* X, int64_t ldx, void* mom,
                       int64_t ldm, void* delta, int64_t ldd, void* stream);

/*
 * Far-field current moments on a theta x phi grid for p columns:
 * M_theta, M_phi [n_theta][n_phi][p] (direction stride ldo >= p) from
 * element midpoints mid [n_el][3] and moments mom as above, with
 * trig_theta [n_theta][2] = (sin, cos) of each theta and trig_phi likewise.
 * ground: 0 free space, 1 PEC image, 2 finite ground through the Fresnel
 * image multipliers at eps_c = eps_re + j eps_im, the image mirrored about
 * z = ground_z.  Replaces _far_moments (reference portal/_portal.py:1520-1598,
 * image terms :1384-1423), one column per call there.
 */
int mw_far_moments(int32_t n_el, const double* mid, const void* mom, int64_t ldm, int32_t p,
                   int32_t n_theta, int32_t n_phi, const double* trig_theta,
                   const double* trig_phi, double k, int32_t ground, double ground_z,
                   double eps_re, double eps_im, void* m_theta, void* m_phi, int64_t ldo,
                   void* stream);

/*
 * Near fields in mixed-potential form at pts [n_pts][3] for p columns:
 * out [n_pts][3][p] (point stride ldo >= 3p) is E (magnetic = 0; element
 * moments plus node charges delta * node_sign, nodes [n_node][3]) or H
 * (magnetic = 1; moments only), thin-wire distance sqrt(|R|^2 + radius^2),
 * eta the wave impedance.  accumulate = 1 adds to out (a PEC image is a
 * second call on mirrored elements with node_sign = -1).  Replaces
 * _element_fields (reference portal/_portal.py:1601-1651).
 */
int mw_near_fields(int32_t n_pts, const double* pts, int32_t n_el, const double* mid,
                   const void* mom, int64_t ldm, int32_t n_node, const double* nodes,
                   const void* delta, int64_t ldd, double node_sign, int32_t p, double k,
                   double radius, double eta, int32_t magnetic, int32_t accumulate, void* out,
                   int64_t ldo, void* stream);

/* ------------------------------------------------------------------------
 * Learned Sommerfeld surrogate (no counterpart in the reference)
 *
 * A SiLU MLP, 4 hidden layers of 256, BF16 weights, FP32 accumulation on
 * MFMA: inputs (rho/lambda, (z+z')/lambda, eps_r, sigma/(omega eps0)) of one
 * node, outputs the four remainder surfaces (8 reals) divided by
 * -j omega mu/(4 pi).  All buffers are device memory owned by the caller:
 * params/grads/adam_m/adam_v hold mw_mlp_param_count() floats, w16/wt16 hold
 * mw_mlp_w16_count() bf16 values (the BF16 weights and their transposes).
 * ------------------------------------------------------------------------ */
typedef struct mw_mlp {
  float* params;
  float* grads;
  float* adam_m;
  float* adam_v;
  uint16_t* w16;
  uint16_t* wt16;
} mw_mlp;

int64_t mw_mlp_param_count(void);
int64_t mw_mlp_w16_count(void);
/* workspace bytes for a batch of n nodes (inference or training) */
int64_t mw_mlp_workspace_bytes(int64_t n);
/* refresh the BF16 copies from the FP32 master parameters */
int mw_mlp_sync_weights(const mw_mlp* net, void* stream);
/* out [n][8] float = network(x [n][4] float) */
int mw_mlp_forward(const mw_mlp* net, const float* x, int64_t n, float* out, void* ws,
                   void* stream);
/* forward, mean-square loss against target [n][8] (written to *loss, device
 * float), and grads = dLoss/dparams (overwritten) */
int mw_mlp_backward(const mw_mlp* net, const float* x, const float* target, int64_t n,
                    float* loss, void* ws, void* stream);
/* one Adam step (step counts from 1) on params from grads, then the BF16 copies */
int mw_mlp_adam_step(const mw_mlp* net, float lr, float beta1, float beta2, float eps,
                     int32_t step, void* stream);
/* surfaces of nodes idx[0..n) of a normalised master (R1, theta device arrays,
 * lambda = 1) for one ground, written into vals[idx[i]][4] (complex);
 * value_re/im = -j omega mu/(4 pi) of the master */
int mw_mlp_grid_values(const mw_mlp* net, double eps_re, double eps_im, const double* R1,
                       const double* theta, const int32_t* idx, int64_t n, void* vals,
                       double value_re, double value_im, void* ws, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MOMWIRE_HIP_H */
