/*
 * cloudfem_hip.h -- C ABI of libcloudfem_hip.so, the MI355X (gfx950) kernels
 * behind the tetrahedral P1/Tet10 assembly and sparse SPD solve of CloudFEM.
 *
 * The reference (CloudFEM) is pure Python; its boundary for this path is the
 * set of Python functions in cloudfem/fem/{tet_poisson,tet_heat,tet10}.py and
 * cloudfem/physics/{elastic_tet10,elastic3d,modal3d}.py.  Each entry point
 * below names the reference code whose arithmetic it replaces.  The Python
 * package cloudfem_amd binds these with ctypes and presents the reference's
 * function signatures.
 *
 * Conventions
 *  - Every function returns an int status (CF_OK = 0); cf_last_error() gives
 *    the message of the last failure on the calling thread.
 *  - All array arguments are DEVICE pointers unless the name ends in _host.
 *    Device buffers are owned by the caller; the library allocates no device
 *    memory.  Scratch space is passed in as (ws, ws_bytes); the matching
 *    *_workspace_bytes() call gives the size.
 *  - `stream` is a hipStream_t (NULL = default stream).  Calls are
 *    asynchronous on that stream unless stated otherwise.
 *  - Indices are int32 (scipy's CSR index type for these sizes); node
 *    coordinates are FP64 structure-of-arrays x[N], y[N], z[N];
 *    connectivity is int32 row-major (n_elem, k), k = 4 (P1) or 10 (Tet10).
 *  - Node-pair pattern: the CSR pattern of the node graph induced by the
 *    elements (row = node, sorted unique columns, duplicates merged, explicit
 *    zeros kept) -- exactly the pattern scipy's coo_matrix(...).tocsr()
 *    produces for the reference's element triplets.
 *  - Vector (3 DOF per node, interleaved DOF 3i+c) matrices are stored in
 *    "NB3" order: the value array of the scipy CSR of the DOF matrix, whose
 *    node row i occupies values[9*ptr[i] .. 9*ptr[i+1]), DOF row 3i+c holding
 *    the 3*len_i entries block(i,j_0)[c][0..2], block(i,j_1)[c][0..2], ...
 *    The node pattern (ptr, idx) plus this array is enough for SpMV; the
 *    DOF-level indptr/indices are produced only for export.
 */
#ifndef CLOUDFEM_HIP_H
#define CLOUDFEM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define CF_OK 0
#define CF_ERR_ARG 1            /* bad argument (size, null pointer, op)    */
#define CF_ERR_HIP 2            /* HIP runtime error                        */
#define CF_ERR_WORKSPACE 3      /* workspace too small                      */
#define CF_ERR_DEGENERATE 4     /* zero-volume element (LinAlgError)        */
#define CF_ERR_NOT_CONVERGED 5  /* PCG hit maxit                            */
#define CF_ERR_BREAKDOWN 6      /* PCG breakdown: NaN/Inf or p.Ap <= 0      */

/* ---- matrix operators (cf_assemble_matrix) ------------------------------ */
#define CF_OP_P1_STIFF 1   /* k_e V G G^T              tet_poisson.py:51-66   */
#define CF_OP_P1_MASS 2    /* rho_e V (1+d_ab)/20      tet_heat.py:72-94    */
#define CF_OP_T10_STIFF 3  /* V sum_q w_q k_q G G^T    tet10.py:138-165       */
#define CF_OP_T10_MASS 4   /* rho_e V MASS_REF         tet10.py:185-249       */
#define CF_OP_T10_ELAST 5  /* V sum_q w_q B^T D B      elastic_tet10.py:164-173 */
#define CF_OP_P1_ELAST 6   /* V B^T D B                elastic3d.py:145-155   */
#define CF_OP_P1_VMASS 7   /* rho V (1+d)/20 per comp  modal3d.py:62-75       */

/* ---- right-hand-side operators (cf_assemble_rhs) ------------------------- */
#define CF_RHS_P1_SOURCE 1    /* f_e V/4                  tet_poisson.py:62      */
#define CF_RHS_T10_SOURCE 2   /* V sum_q w_q f_q N_q      tet10.py:155           */
#define CF_RHS_P1_BODY 3      /* b_e,c V/4                elastic3d.py:157-162   */
#define CF_RHS_T10_BODY 4     /* V sum_q w_q N_q b_q,c    elastic_tet10.py:175-179 */
#define CF_RHS_P1_THERMAL 5   /* V (3l+2m) s_e g_a        elastic3d.py:164-169   */
#define CF_RHS_T10_THERMAL 6  /* V sum_q w_q (3l+2m) s_q g_aq  elastic_tet10.py:181-190 */

/* ---- face operators (cf_face_rhs, cf_face_matrix) ------------------------ */
#define CF_FACE_P1_FLUX 1       /* q_f A/3                tet_heat.py:149-154    */
#define CF_FACE_P1_ROBIN_K 2    /* h_f A (1+d)/12         tet_heat.py:156-168    */
#define CF_FACE_T10_FLUX 3      /* A sum_q w6 q_q N6      tet10.py:393-401       */
#define CF_FACE_T10_ROBIN_K 4   /* A sum_q w6 h_q N6 N6^T tet10.py:413-427       */
#define CF_FACE_P1_TRACTION 5   /* t_f,c A/3              elastic3d.py:171-186   */
#define CF_FACE_T10_TRACTION 6  /* A sum_q w3 N3_q t_q,c  elastic_tet10.py:192-206 */

/* ---- solver options ------------------------------------------------------ */
#define CF_FMT_CSR 0    /* scalar CSR: rows = DOFs                           */
#define CF_FMT_NB3 1    /* 3x3 node blocks in NB3 order: rows = nodes        */
#define CF_PREC_FP64 0  /* plain FP64 PCG                                    */
#define CF_PREC_MIXED 1 /* NB3: FP64 flexible CG preconditioned by an FP32
                           Chebyshev polynomial of the Jacobi-scaled matrix
                           (FP32 stored values, 23 products on [b/3000, b];
                           FP16 stored: 15 products).  CSR: FP32 inner CG on the
                           Jacobi-scaled matrix with FP64 residual
                           replacement + solution accumulation              */
#define CF_PREC_MIXED16 2 /* as CF_PREC_MIXED, the polynomial's matrix copy
                           stored in FP16 (FP32 arithmetic)                  */
#define CF_PC_JACOBI 0
#define CF_PC_BLOCK3 1  /* 3x3 node-block Jacobi (NB3 only, FP64 only)       */

typedef struct cf_matrix {
  int32_t format;        /* CF_FMT_CSR or CF_FMT_NB3                          */
  int32_t n_rows;        /* CSR: DOF rows; NB3: node rows                     */
  int64_t nnz;           /* CSR: stored entries; NB3: stored node blocks      */
  const int32_t* indptr; /* n_rows + 1                                        */
  const int32_t* indices;/* nnz (node columns for NB3)                        */
  const double* values;  /* CSR: nnz; NB3: 9*nnz                              */
} cf_matrix;

typedef struct cf_pcg_stats {
  int32_t iterations;        /* CG iterations (flexible CG: outer FP64 ones)  */
  int32_t residual_updates;  /* scalar CSR mixed: FP64 residual replacements;
                                Chebyshev-preconditioned solvers: products in
                                low precision                               */
  double rel_residual;       /* true ||b_f - A_ff x_f|| / ||b_f||, FP64, at exit */
  double b_norm;             /* ||b_f||                                      */
  int32_t restarts;          /* FP64 and flexible CG: restarts from the true
                                residual after the recursive one met tol     */
  int32_t floor_limited;     /* 1: a restart no longer halved the true
                                residual, i.e. tol lies below what FP64 resolves
                                for this system (~eps ||A_ff| |x_f|| / ||b_f||);
                                the call returns CF_OK with rel_residual > tol */
} cf_pcg_stats;

/* ---- library ------------------------------------------------------------ */
int cf_version(void);
const char* cf_last_error(void);

/* ---- node-pair pattern (replaces the pattern half of
 *      scipy coo_matrix(...).tocsr() at tet_poisson.py:68, tet_heat.py:94,
 *      tet10.py:162-164 / 247-249, elastic_tet10.py:208-210,
 *      elastic3d.py:188, modal3d.py:75) ------------------------------------
 * Two phases around one host read of nnz:
 *  1. cf_pattern_sort: radix-sorts all n_elem*k*k element node pairs by
 *     (row, col), stable in pair id p = e*k*k + a*k + b, writes the sorted
 *     pair ids to contrib[] (the contributor list of every stored entry, in
 *     ascending element order) and nnz to *nnz_dev.
 *  2. cf_pattern_finalize (same ws): indptr[N+1], indices[nnz] and the
 *     contributor segment offsets seg[nnz+1] (entry s sums contrib[seg[s] ..
 *     seg[s+1])).                                                           */
int cf_pattern_workspace_bytes(int64_t n_elem, int32_t k, int64_t n_nodes,
                               size_t* bytes);
int cf_pattern_sort(const int32_t* conn, int64_t n_elem, int32_t k,
                    int64_t n_nodes, void* ws, size_t ws_bytes,
                    uint32_t* contrib, int64_t* nnz_dev, void* stream);
int cf_pattern_finalize(int64_t n_elem, int32_t k, int64_t n_nodes,
                        int64_t nnz, const void* ws, int32_t* indptr,
                        int32_t* indices, int32_t* seg, void* stream);

/* ---- node incidence: for every node the (element, local slot) items that
 *      touch it, ascending; used for deterministic RHS gathers
 *      (replaces rhs[el] += ... at tet_poisson.py:62, tet10.py:161,
 *      elastic_tet10.py:179/190/206, tet_heat.py:154/291) ---------------- */
int cf_incidence_workspace_bytes(int64_t n_items, int64_t n_nodes,
                                 size_t* bytes);
int cf_incidence_build(const int32_t* targets, int64_t n_items,
                       int64_t n_nodes, void* ws, size_t ws_bytes,
                       uint32_t* perm, int32_t* offs, void* stream);

/* ---- element geometry: inverse corner Jacobian (row-major, Jinv[m][i] =
 *      d xi_m / d x_i) and volume |det J|/6 per element, geo[e*10 + 0..9].
 *      Replaces _tet_gradients (tet_poisson.py:17-32) and the Jacobian part
 *      of element_operators (tet10.py:113-124).  *n_bad_dev receives the
 *      number of elements with det J == 0 (the reference raises LinAlgError).
 */
int cf_tet_geometry(const double* x, const double* y, const double* z,
                    const int32_t* conn, int64_t n_elem, int32_t k,
                    double* geo, int32_t* n_bad_dev, void* stream);

/* ---- matrix values: one thread per stored node pair sums its
 *      contributors' local entries, recomputed from geo + constant
 *      reference tables, in ascending element order (deterministic, no
 *      atomics).  coef: per-element (P1_STIFF, P1_MASS, T10_MASS, P1_VMASS)
 *      or per-element-per-point [e*5+q] (T10_STIFF); unused for *_ELAST,
 *      which take Lame lam, mu.  Output: scalar ops write values[nnz];
 *      *_ELAST write NB3 values[9*nnz]; P1_VMASS writes the
 *      component-diagonal layout values[3*nnz] (DOF row 3i+c holds
 *      block(i,j)[c][c] for the row's columns j).  conn: the connectivity
 *      the pattern was built from (the row node of an entry is conn[e][a]
 *      of its first contributor).                                          */
int cf_assemble_matrix(int32_t op, const int32_t* indptr,
                       const int32_t* indices, const int32_t* seg,
                       const uint32_t* contrib, const int32_t* conn, int64_t n_nodes,
                       int64_t nnz, int32_t k, const double* geo, const double* coef,
                       double lam, double mu, double* values, void* stream);

/* ---- right-hand sides: rhs[i] (or rhs[3i+c]) += sum over the node's
 *      incidence list; coef layouts: P1_SOURCE f[e]; T10_SOURCE f[e*5+q];
 *      P1_BODY b[e*3+c]; T10_BODY b[(e*5+q)*3+c]; P1_THERMAL s[e];
 *      T10_THERMAL s[e*5+q] with s = alpha (T - T_ref).                      */
int cf_assemble_rhs(int32_t op, const int32_t* offs, const uint32_t* perm,
                    int64_t n_nodes, int32_t k, const double* geo,
                    const double* coef, double lam, double mu, double* rhs,
                    void* stream);

/* ---- boundary-face loads: faces (n_faces, nf) int32 with nf = 3 (P1) or
 *      6 (Tet10, [c0,c1,c2,m01,m12,m20]); offs/perm = cf_incidence_build of
 *      the face array.  coef: P1_FLUX q[f]; T10_FLUX q[f*6+q];
 *      P1_TRACTION t[f*3+c]; T10_TRACTION t[(f*3+q)*3+c].  Adds into rhs.  */
int cf_face_rhs(int32_t op, const double* x, const double* y,
                const double* z, const int32_t* faces, int64_t n_faces,
                int32_t nf, const int32_t* offs, const uint32_t* perm,
                int64_t n_nodes, const double* coef, double* rhs,
                void* stream);

/* ---- boundary-face matrices (Robin): values of the face pattern built by
 *      cf_pattern_* over the face array.  coef: P1_ROBIN_K h[f];
 *      T10_ROBIN_K h[f*6+q].                                                */
int cf_face_matrix(int32_t op, const double* x, const double* y,
                   const double* z, const int32_t* faces, int32_t nf,
                   const int32_t* indptr, const int32_t* indices,
                   const int32_t* seg, const uint32_t* contrib,
                   int64_t n_nodes, int64_t nnz, const double* coef,
                   double* values, void* stream);

/* ---- sparse add of a sub-pattern into a super-pattern (scalar CSR):
 *      for every stored (i, j) of B, A[i, j] += alpha * B[i, j].  Every
 *      stored entry of B must be stored in A (the face pattern is a subset
 *      of the element pattern).  Replaces K + K_add (tet_heat.py:230).      */
int cf_csr_add_subpattern(int64_t n_rows, const int32_t* a_ptr,
                          const int32_t* a_idx, double* a_val,
                          const int32_t* b_ptr, const int32_t* b_idx,
                          const double* b_val, double alpha,
                          int32_t* n_missing_dev, void* stream);

/* ---- element strain/stress at the centroid, Voigt [xx,yy,zz,yz,xz,xy],
 *      engineering shear: eps = B u_e, sig = D (eps - s_e (1,1,1,0,0,0)) with
 *      isotropic D(lam, mu); s_thermal[e] = alpha (T_c - T_ref) or NULL.
 *      Replaces element_stress_strain_3d (elastic3d.py:252-289) for k = 4
 *      and element_stress_tet10 (elastic_tet10.py:232-279) for k = 10.
 *      u is the interleaved (3N) displacement; sig, eps are (n_elem, 6).    */
int cf_element_stress(const int32_t* conn, int64_t n_elem, int32_t k,
                      const double* geo, const double* u,
                      const double* s_thermal, double lam, double mu,
                      double* sig, double* eps, void* stream);

/* ---- adjoint sensitivity of a steady diffusion solve to a per-element
 *      conductivity (no reference counterpart; the reference has no
 *      adjoint): dk[e] = -lambda_e^T (dK_e/dk_e) T_e, i.e. -V grad(lambda).grad(T)
 *      for P1 (k = 4) and -V sum_q w_q grad(lambda)_q.grad(T)_q for Tet10
 *      (k = 10), where K lambda = dJ/dT on the free DOFs, lambda = 0 on fixed
 *      ones.  T, lambda are nodal (N); dk is (n_elem).                       */
int cf_adjoint_dk(const int32_t* conn, int64_t n_elem, int32_t k,
                  const double* geo, const double* T, const double* lam,
                  double* dk, void* stream);

/* ---- export: DOF-level CSR indptr/indices of a vector matrix from its node
 *      pattern.  dense3 = 1: full 3x3 blocks (elasticity, 9*nnz entries);
 *      dense3 = 0: component-diagonal blocks (vector mass, 3*nnz entries). */
int cf_expand_pattern3(int64_t n_nodes, const int32_t* ptr,
                       const int32_t* idx, int32_t dense3,
                       int32_t* dof_ptr, int32_t* dof_idx, void* stream);

/* ---- scalar Dirichlet elimination with the reference's structure
 *      (tet_poisson.apply_dirichlet, tet_poisson.py:72-110): fixed rows become
 *      identity rows, rhs_out = rhs - A[:, F] v on free rows and v on fixed
 *      rows, and every row stores every fixed column (explicit zeros), as
 *      scipy's A[:, F] = 0 does.  A must be canonical CSR.  fixed_val[i] is the
 *      value of fixed node i; fixed_list = sorted F.  Two phases around one
 *      host read of the output nnz, like cf_pattern_*.                      */
int cf_dirichlet_workspace_bytes(int64_t n, size_t* bytes);
int cf_dirichlet_count(int64_t n, const int32_t* ptr, const int32_t* idx,
                       const double* val, const double* rhs,
                       const uint8_t* fixed, const double* fixed_val,
                       int64_t n_fixed, void* ws, size_t ws_bytes,
                       int32_t* ptr_out, double* rhs_out, int64_t* nnz_dev,
                       void* stream);
int cf_dirichlet_fill(int64_t n, const int32_t* ptr, const int32_t* idx,
                      const double* val, const uint8_t* fixed,
                      const int32_t* fixed_list, int64_t n_fixed,
                      const int32_t* ptr_out, int32_t* idx_out,
                      double* val_out, void* stream);

/* ---- y = a x + b y (theta-method operators C + th h K, per-step loads;
 *      replaces the scipy sums at tet_heat.py:239-240, 386).              */
int cf_axpby(int64_t n, double a, const double* x, double b, double* y,
             void* stream);
/* ---- dst[i] = src[i] where mask[i] (re-imposing Dirichlet values,
 *      tet_heat.py:266-270).                                              */
int cf_copy_masked(int64_t n, const uint8_t* mask, const double* src,
                   double* dst, void* stream);

/* ---- y = A x (FP64), optional row mask (fixed[i] != 0 -> y_i = 0). ------ */
int cf_spmv(const cf_matrix* A, const double* x, double* y,
            const uint8_t* fixed, void* stream);

/* ---- y = |A| |x| (fixed rows 0): no reference counterpart; it gives the
 *      FP64 rounding floor eps ||A_ff| |x_f|| / ||b_f|| below which the
 *      relative residual of a solve cannot be resolved.                  */
int cf_spmv_abs(const cf_matrix* A, const double* x, double* y, const uint8_t* fixed,
                void* stream);

/* ---- preconditioned CG on the free-DOF system (replaces spsolve /
 *      apply_dirichlet_vec3 at elastic3d.py:220-232, tet_poisson.py:121,
 *      tet10.py:177, elastic_tet10.py:225-228).
 *  fixed (nullable): fixed[i] != 0 marks a Dirichlet DOF.  On entry x holds
 *  the fixed values at fixed DOFs and the initial guess at free DOFs; on exit
 *  x[free] solves A_ff x_f = b_f - A_fF x_F to ||res||/||b_f - A_fF x_F||
 *  <= tol (FP64 residual) and x[fixed] is unchanged.  Synchronous: returns
 *  after the solve, stats_host filled.  CF_ERR_NOT_CONVERGED if maxit is
 *  reached, CF_ERR_BREAKDOWN on NaN/Inf or a non-positive curvature.       */
int cf_pcg_workspace_bytes(const cf_matrix* A, int32_t precision,
                           int32_t precond, size_t* bytes);
int cf_pcg_solve(const cf_matrix* A, const double* b, double* x,
                 const uint8_t* fixed, int32_t precision, int32_t precond,
                 double tol, int32_t maxit, void* ws, size_t ws_bytes,
                 cf_pcg_stats* stats_host, void* stream);

/* ---- matrix-free Tet10 elasticity operator (no assembled matrix): element
 *      forces K_e x_e from the geometry record (exact 4-point rule for the
 *      degree-2 integrand of affine elements; the reference's 5-point rule
 *      gives the same matrix up to rounding), summed into the nodes in fixed
 *      order.  offs/perm = cf_incidence_build of the Tet10 connectivity
 *      (conn, in the element order of the caller's choice; consecutive
 *      elements sharing nodes make the node sums cheaper).  Each call first
 *      builds a run plan in the workspace: per tile of 256 elements, the
 *      forces are summed per node in LDS and one partial per (tile, node) is
 *      written, so no per-element force goes through HBM.  Same operator as
 *      CF_OP_T10_ELAST (isotropic lam, mu; affine elements), so it replaces
 *      the stiffness in the solve of elastic_tet10.py:214-229 without the
 *      12 B/value matrix traffic.  Workspace: about 52 doubles per element
 *      + 9 per node + the PCG vectors.                                       */
int cf_mf_t10_workspace_bytes(int64_t n_nodes, int64_t n_elem, int32_t precond,
                              size_t* bytes);
int cf_mf_t10_apply(int64_t n_nodes, int64_t n_elem, const int32_t* conn,
                    const double* geo, double lam, double mu,
                    const int32_t* offs, const uint32_t* perm,
                    const double* x, double* y, const uint8_t* fixed,
                    void* ws, size_t ws_bytes, void* stream);
int cf_pcg_solve_mf_t10(int64_t n_nodes, int64_t n_elem, const int32_t* conn,
                        const double* geo, double lam, double mu,
                        const int32_t* offs, const uint32_t* perm,
                        const double* b, double* x, const uint8_t* fixed,
                        int32_t precond, double tol, int32_t maxit, void* ws,
                        size_t ws_bytes, cf_pcg_stats* stats_host, void* stream);

/* ---- chunked plan of the matrix-free Tet10 operator.  Elements are put in
 *      Morton order of their centroids (bbox: host xmin, ymin, zmin, xmax,
 *      ymax, zmax of the nodes) and cut into chunks of 256; each chunk
 *      computes its element forces into LDS and sums them per node, so no
 *      per-element forces go through HBM.  Nodes shared between chunks are
 *      finished from per-chunk partials in chunk order (deterministic).
 *      offs = node incidence offsets (cf_incidence_build of the
 *      connectivity).  plan: caller buffer of cf_mf_t10_plan_bytes;
 *      counts (host, 3): pairs, shared pairs, shared nodes, written by the
 *      build and passed back to apply/solve.  Same operator and solve as
 *      cf_mf_t10_apply / cf_pcg_solve_mf_t10 (which they replace in the
 *      elastic_tet10.py:214-229 solve); conn/geo/offs/perm of the solve are
 *      the original (unpermuted) arrays, used for the preconditioner.      */
int cf_mf_t10_plan_bytes(int64_t n_elem, size_t* bytes);
int cf_mf_t10_plan_build(int64_t n_nodes, int64_t n_elem, const int32_t* conn,
                         const double* geo, const double* x, const double* y,
                         const double* z, const int32_t* offs, const double* bbox,
                         void* plan, size_t plan_bytes, int64_t* counts, void* stream);
int cf_mf_t10_plan_workspace_bytes(int64_t n_nodes, int64_t n_elem,
                                   const int64_t* counts, int32_t precond,
                                   size_t* bytes);
int cf_mf_t10_plan_apply(int64_t n_nodes, int64_t n_elem, const void* plan,
                         const int64_t* counts, double lam, double mu,
                         const double* x, double* y, const uint8_t* fixed,
                         void* ws, size_t ws_bytes, void* stream);
int cf_pcg_solve_mf_t10_plan(int64_t n_nodes, int64_t n_elem, const int32_t* conn,
                             const double* geo, const int32_t* offs,
                             const uint32_t* perm, const void* plan,
                             const int64_t* counts, double lam, double mu,
                             const double* b, double* x, const uint8_t* fixed,
                             int32_t precond, double tol, int32_t maxit, void* ws,
                             size_t ws_bytes, cf_pcg_stats* stats_host, void* stream);

/* ---- mesh processing on the GPU.
 *  cf_tet10_edges: Tet4 -> Tet10 numbering of tet10_mesh.py:77-112 -- the
 *  midpoint of corner pair (a < b) gets n_corners + rank of (a, b) among the
 *  unique pairs in lexicographic order (np.unique order).  conn4 has `stride`
 *  ints per element (4, or 10 to reuse Tet10 corners); writes conn10 (M x 10),
 *  the unique edges' endpoints edge_a/edge_b (capacity 6 M) and *n_edges
 *  (device).  cf_edge_midpoints: x/y/z[n_corners + i] = endpoint mean.
 *  cf_boundary_faces: exterior faces (sorted corner triple in exactly one
 *  element) whose corners are all in `member` (nullable: all), in first-
 *  occurrence order with that occurrence's node order (tet_heat.py:97-120,
 *  tet10.py:329-352, elastic_tet10.py:68-91); width 3 (P1 corner faces) or 6
 *  (Tet10 faces, corners + midpoints); faces capacity 4 M x width; *count
 *  (device).  n_nodes < 2^21.  Workspace: cf_mesh_workspace_bytes.       */
int cf_mesh_workspace_bytes(int64_t n_elem, size_t* bytes);
int cf_tet10_edges(int64_t n_elem, const int32_t* conn4, int32_t stride,
                   int64_t n_corners, int32_t* conn10, int32_t* edge_a,
                   int32_t* edge_b, int64_t* n_edges, void* ws, size_t ws_bytes,
                   void* stream);
int cf_edge_midpoints(int64_t n_edges, int64_t n_corners, const int32_t* edge_a,
                      const int32_t* edge_b, double* x, double* y, double* z,
                      void* stream);
int cf_boundary_faces(int64_t n_elem, const int32_t* conn, int32_t k, int64_t n_nodes,
                      const uint8_t* member, int32_t width, int32_t* faces,
                      int64_t* count, void* ws, size_t ws_bytes, void* stream);

/* ---- mixed precision with the matrix-free operator: FP64 flexible CG on
 *      the FP64 matrix-free product (as cf_pcg_solve_mf_t10), preconditioned
 *      by 23 Chebyshev steps on S K S (S = diag(K)^-1/2) whose products are
 *      matrix-free in FP32 (geometry rounded once to FP32, forces and node
 *      sums in FP32, same run plan).  The outer residual, recurrence and
 *      solution are FP64, so the result meets the same residual bar.        */
int cf_mf_t10_mixed_workspace_bytes(int64_t n_nodes, int64_t n_elem, size_t* bytes);
int cf_pcg_solve_mf_t10_mixed(int64_t n_nodes, int64_t n_elem, const int32_t* conn,
                              const double* geo, double lam, double mu,
                              const int32_t* offs, const uint32_t* perm,
                              const double* b, double* x, const uint8_t* fixed,
                              double tol, int32_t maxit, void* ws, size_t ws_bytes,
                              cf_pcg_stats* stats_host, void* stream);

/* ---- two-level mixed precision with the matrix-free operator (no
 *      reference counterpart; replaces the spsolve of elastic_tet10.py:214-229
 *      like cf_pcg_solve_mf_t10).  FP64 flexible CG on the FP64 matrix-free
 *      product, preconditioned by z1 = C r, z2 = z1 + Cheb(r - K z1),
 *      z = z2 + C (r - K z2): Cheb the FP32 matrix-free Chebyshev polynomial
 *      of cf_pcg_solve_mf_t10_mixed, C = P A_c^-1 P^T a coarse correction on
 *      the six rigid-body modes of node aggregates (the group of 512 to
 *      1536 consecutive elements, about 250 groups per mesh, that first
 *      touches a node), A_c = P^T K_ff P
 *      assembled by probing with distance-2 coloured aggregates (no atomics)
 *      and inverted by a banded Cholesky factorisation and a column-parallel
 *      inverse when its half-bandwidth is at most 62 (the aggregates follow
 *      the element order), else by rocSOLVER (dpotrf + dpotri; rocBLAS
 *      allocates its own small workspace and handle, once per device).  The
 *      Lanczos estimate for the Chebyshev bound runs on a side stream (created
 *      once per device) beside the coarse set-up, launched from a second host
 *      thread; on return the caller's stream is ordered after both.  Solves
 *      that overlap on one device share that stream: run them one at a time.
 *      x_nodes/y_nodes/z_nodes: node coordinates (device, FP64).  Same FP64
 *      residual bar as the other solvers.                                    */
int cf_mf_t10_twolevel_workspace_bytes(int64_t n_nodes, int64_t n_elem, size_t* bytes);
int cf_pcg_solve_mf_t10_twolevel(int64_t n_nodes, int64_t n_elem, const int32_t* conn,
                                 const double* geo, double lam, double mu,
                                 const int32_t* offs, const uint32_t* perm,
                                 const double* x_nodes, const double* y_nodes,
                                 const double* z_nodes, const double* b, double* x,
                                 const uint8_t* fixed, double tol, int32_t maxit, void* ws,
                                 size_t ws_bytes, cf_pcg_stats* stats_host, void* stream);

/* ---- two-level solver, general form (no reference counterpart; replaces
 *      the spsolve of elastic_tet10.py:214-229 like cf_pcg_solve).  The
 *      preconditioner is the two-level cycle of cf_pcg_solve_mf_t10_twolevel,
 *      its Chebyshev polynomial and coarse prolongation run matrix-free in
 *      `precision`: CF_PREC_MIXED = FP32, CF_PREC_FP64 = FP64 (the same
 *      algorithm in FP64 throughout, for measuring what precision alone
 *      gains).  Without A the outer FP64 flexible CG multiplies by the FP64
 *      matrix-free operator (the same matrix as the assembled one up to
 *      rounding).  With A (the assembled NB3 stiffness of the same mesh,
 *      node rows = n_nodes; the solver builds a sliced copy of it in the
 *      workspace) the outer products start matrix-free and, once that
 *      converges, switch to A: the solver restarts from A's true residual
 *      for at least one segment, so the system solved is A's own -- the
 *      matrix cf_assemble_matrix(CF_OP_T10_ELAST) returns and the reference
 *      builds.  Same stop rule and residual replacement as cf_pcg_solve.    */
int cf_t10_twolevel_workspace_bytes(const cf_matrix* A, int64_t n_nodes, int64_t n_elem,
                                    int32_t precision, size_t* bytes);
int cf_pcg_solve_t10_twolevel(const cf_matrix* A, int64_t n_nodes, int64_t n_elem,
                              const int32_t* conn, const double* geo, double lam, double mu,
                              const int32_t* offs, const uint32_t* perm,
                              const double* x_nodes, const double* y_nodes,
                              const double* z_nodes, const double* b, double* x,
                              const uint8_t* fixed, int32_t precision, double tol,
                              int32_t maxit, void* ws, size_t ws_bytes,
                              cf_pcg_stats* stats_host, void* stream);

/* ---- no-allocation PCG (no reference counterpart): for callers that cannot
 *      allocate device memory or pass a workspace -- code under HIP graph
 *      capture, tight serving loops, harnesses that forbid hipMalloc.  Same
 *      solve as cf_pcg_solve(A, ..., CF_PREC_FP64, CF_PC_JACOBI, ...) for a
 *      scalar CSR matrix (CF_FMT_CSR only), with the workspace in a static
 *      device array of the library (8.6 MB per GPU, held until the process
 *      ends): at most 262,144 unknowns, larger systems return CF_ERR_ARG.
 *      The call allocates nothing.  Every call on a GPU shares that GPU's
 *      array, so calls that overlap on one GPU (two streams or two host
 *      threads) race: run them one at a time.                             */
int cf_pcg_solve_noalloc(const cf_matrix* A, const double* b, double* x,
                         const uint8_t* fixed, double tol, int32_t maxit,
                         cf_pcg_stats* stats_host, void* stream);

/* ---- the inner product of the Chebyshev / two-level solvers on its own
 *      (no reference counterpart; for tests and kernel timing):
 *      y = S K S x with S = diag(scale) (scale NULL: S = I) in the solvers'
 *      inner precision -- x, y are float vectors for CF_PREC_MIXED, double
 *      for CF_PREC_FP64 -- by the same run plan and kernels.  reps > 0
 *      repeats the product reps more times between two HIP events and
 *      stores the average milliseconds in *avg_ms_host (synchronous then). */
int cf_mf_t10_lp_workspace_bytes(int64_t n_nodes, int64_t n_elem, int32_t precision,
                                 size_t* bytes);
int cf_mf_t10_lp_apply(int64_t n_nodes, int64_t n_elem, const int32_t* conn, const double* geo,
                       double lam, double mu, const int32_t* offs, const uint32_t* perm,
                       int32_t precision, const double* scale, const void* x, void* y,
                       int32_t reps, float* avg_ms_host, void* ws, size_t ws_bytes,
                       void* stream);

/* ---- inv[perm[j]] = j for a permutation of 0..n-1.                     */
int cf_invert_permutation(int64_t n, const uint32_t* perm, uint32_t* inv,
                          void* stream);

/* ---- timing of the CG product kernel (no reference counterpart; used by
 *      the benchmark to report achieved bandwidth).  Returns the average
 *      duration (ms) and count of the sampled FP64 and FP32 SpMV launches of
 *      cf_pcg_solve since the last reset; enable >= 0 then resets the
 *      counters and switches sampling on (1) or off (0).  Not thread-safe. */
int cf_spmv_profile(int32_t enable, double* avg_ms_fp64, int64_t* n_fp64,
                    double* avg_ms_fp32, int64_t* n_fp32);

/* ---- batched theta-method heat (many meshes, one block-diagonal CSR).
 *      Replaces the per-mesh loop of transient_heat_tet (tet_heat.py:197-274:
 *      (C + th h K) T+ = (C - (1-th) h K) T + h F, factorized once) for a
 *      batch of meshes without Dirichlet nodes: every launch covers all
 *      systems, CG scalars and stop tests are per system (tol relative to
 *      each system's step rhs), one time step is a replayed hipGraph.
 *  sys_rows (host, n_sys + 1): first row of each system, sys_rows[n_sys] =
 *  rows.  A_vals / R_vals: C + th h K and C - (1-th) h K on one pattern.
 *  T (device, rows): initial state in, final state out.  hist (device,
 *  nullable): (n_steps + 1) x rows history, row 0 = the initial state.
 *  precision CF_PREC_FP64 (Jacobi CG) or CF_PREC_MIXED (FP32 CG on S A S
 *  inside FP64 iterative refinement).  lanes (4, 8, 16) per row and
 *  rows_per_tile come from cf_batch_autotune.                              */
typedef struct cf_batch_stats {
  int32_t n_steps;
  int32_t lanes, rows_per_tile;  /* launch configuration used              */
  int32_t step_iterations;  /* CG iterations (mixed: sweeps) in the step graph */
  int32_t graph_replays;    /* hipGraph launches (step + iteration graphs)  */
  int32_t extra_replays;    /* iteration-graph launches after a step graph  */
  int64_t iterations;       /* CG iterations summed over systems and steps
                               (mixed: FP32 inner iterations)              */
  int64_t sweeps;           /* mixed: FP64 refinement sweeps, summed        */
  int32_t worst_step;       /* most iterations (sweeps) one system needed   */
  int32_t failed;           /* system-steps that did not converge           */
  double max_rel_residual;  /* largest final ||r|| / ||b|| over systems     */
} cf_batch_stats;

int cf_batch_workspace_bytes(int32_t n_sys, const int32_t* sys_rows, int64_t nnz,
                             int32_t precision, size_t* bytes);
/* Times y = A x on the whole batch for the 9 (lanes, rows per tile)
 * configurations (reps launches each) and returns the fastest; us_out
 * (nullable, 9 floats) receives every timing, lanes-major.  No reference
 * counterpart: it is the launch-shape tuner of the batched path.         */
int cf_batch_autotune(int32_t n_sys, const int32_t* sys_rows, const int32_t* indptr,
                      const int32_t* indices, const double* values, int64_t nnz,
                      int32_t reps, int32_t* lanes_out, int32_t* rows_out,
                      float* us_out, void* ws, size_t ws_bytes, void* stream);
int cf_batch_theta_run(int32_t n_sys, const int32_t* sys_rows, const int32_t* indptr,
                       const int32_t* indices, int64_t nnz, const double* A_vals,
                       const double* R_vals, const double* F, double h, double* T,
                       double* hist, int32_t n_steps, int32_t precision, double tol,
                       int32_t maxit, int32_t lanes, int32_t rows_per_tile,
                       void* ws, size_t ws_bytes, cf_batch_stats* stats,
                       void* stream);

#ifdef __cplusplus
}
#endif
#endif /* CLOUDFEM_HIP_H */
