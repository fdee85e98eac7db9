/*
 * yasps_amd - C ABI of the MI355X (gfx950) Newton linear-solve path.
 *
 * Every entry point takes plain device pointers and host scalars, never torch types.
 * Entry points are asynchronous on the given HIP stream (NULL = default stream)
 * unless they return a host value (a count or an iteration status), in which case
 * they synchronise the stream before returning. Return value: 0 (or a non-negative
 * count) on success, a negative YSP_ERR_* code on failure; ysp_last_error() gives the
 * message. The library never frees caller memory.
 *
 * Each declaration names the reference interface it replaces (paths relative to the
 * reference repository root).
 */
#ifndef YASPS_AMD_H
#define YASPS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define YSP_OK 0
#define YSP_ERR_HIP (-1)          /* device/runtime failure (reference launchers return -1) */
#define YSP_ERR_ARG (-2)          /* invalid argument or size */
#define YSP_ERR_UNSUPPORTED (-6)  /* shape outside what the kernels implement */

#define YSP_MAX_UNION_CHILDREN 8
#define YSP_MAX_PROGRAM_CONNS 8
#define YSP_MAX_PROGRAM_UNIONS 8

/* ---------------------------------------------------------------- runtime */
const char* ysp_last_error(void);
int ysp_version(void);
/* Entry points whose kernels flag an unsupported input (the element evaluations' slot
 * layouts) do not synchronise: the flag is copied back in stream order and examined by
 * the next ysp_check_deferred (or ysp_sync) on the same stream, which synchronises the
 * stream and returns YSP_ERR_UNSUPPORTED (message in ysp_last_error) if any queued flag
 * was raised. The reference's kernel calls are synchronous and raise at once
 * (hessianAndGradientKernel.pyx:348-349); the Python facade restores that by checking
 * at the end of every assembly, or after the solve in a Newton step. */
int ysp_check_deferred(void* stream);
int ysp_sync(void* stream);   /* ysp_check_deferred */
/* Frees the library's grow-on-demand scratch buffers. */
void ysp_release_scratch(void);

/* A union of per-child FP64 attribute arrays, e.g. [fixed_vertices.position,
 * moving_vertices.position] stacked in one union primitive
 * (yasps/yasps/primitive/primitiveUnion.pyx:96-105,115-210). counts is a device
 * array of the children's instance counts (children_primitive_counts_gpu).
 * Union instance v belongs to the first child c with sum(counts[0..c]) > v and reads
 * data[c][(v - sum(counts[0..c-1])) * width + k]. */
typedef struct ysp_union_f64 {
  uint32_t num_children;
  uint32_t width;                          /* doubles per instance (3 for a position) */
  const uint32_t* counts;                  /* device [num_children] */
  const double* data[YSP_MAX_UNION_CHILDREN];
} ysp_union_f64;

/* ------------------------------------------------------- index pipeline */
/* Index program bytecode (int32, device memory). Every node starts with
 * [op, k (index slots), g (gradient size), len (ints in this node incl. children)]:
 *   LEAF   [0, 1, size, 6, wrt_pos (-1 = not a target), per_instance (0/1)]
 *   JOIN   [1, k, g, len, conn_slot, dim, n_children, child...]
 *   UNION  [2, k, g, len, counts_slot, n_union_children,
 *           for each union child: n_members, member...]
 *   CONCAT [3, k, g, len, n_children, child...]
 * conns[conn_slot] is a row-major uint32 connectivity [instances x dim];
 * union_counts[counts_slot] the device child-count array of that union.
 * Replaces the generated get_indices(<join conns>, <union counts>, wrtStartIndices,
 * outputIndices, outputSizes, NUM_INSTANCES)
 * (yasps/yasps/kernel/Coordinate/gradientIndicesKernel.pyx:904-926): same slot
 * encoding (scalar start + 2; 1 = not a target; 0 = union padding, ``:699-713,809-811``).
 * out_idx / out_sizes must be zeroed by the caller (``:969-972``). */
int ysp_get_indices(const int32_t* program, int32_t program_len,
                    const uint32_t* const* conns, int32_t num_conns,
                    const uint32_t* const* union_counts, int32_t num_unions,
                    const uint32_t* wrt_start_indices,
                    uint32_t* out_indices, uint16_t* out_sizes,
                    uint32_t num_instances, uint32_t num_indices_for_each_instance,
                    void* stream);

/* Replaces compress_indices (gradientIndicesKernel.pyx:132-214), same arguments:
 * local permutations, compressed gradient sizes, coordinate-count prefix
 * (coordinate_counts_outer[0] = 0 is written here), stable grouping of instances by
 * compressed size (unique_sizes, grouped_indices, offsets[U+1]) and the device
 * count of unique sizes. */
int ysp_compress_indices(const uint32_t* indices, const uint16_t* index_sizes,
                         int16_t* permutations, uint16_t* total_gradient_sizes,
                         uint32_t* coordinate_counts_outer, uint16_t* unique_sizes,
                         uint32_t* grouped_indices, uint32_t* offsets,
                         uint16_t* num_unique,
                         uint32_t num_instances, uint32_t num_indices_for_each_instance,
                         uint32_t compute_coordinate_metadata, void* stream);

/* Replaces computeCoordinates (gradientIndicesKernel.pyx:234-318). */
int ysp_compute_coordinates(const uint32_t* indices, const int16_t* permutations,
                            const uint16_t* index_sizes, uint32_t* coordinates,
                            uint16_t* dimensions, const uint32_t* coordinate_counts_outer,
                            uint32_t num_indices_for_each_instance, uint32_t num_instances,
                            void* stream);

/* 16-byte record layout of coord_dim_tmp (coordinateCompressionKernel.pyx:14-20). */
typedef struct ysp_coord_dim {
  uint32_t row, col;
  uint16_t h, w;
  uint32_t placeholder;
} ysp_coord_dim;

/* Replaces get_unique_coords (coordinateCompressionKernel.pyx:93-144): sorts the
 * coordinates by (h, w, row, col), writes each occurrence's unique ordinal to lookup
 * and the unique records to coord_dim_tmp[0..num_unique). Synchronous; the count is
 * returned through *num_unique_coords (host). */
int ysp_get_unique_coords(const uint32_t* coords, const uint16_t* dims,
                          uint32_t num_coordinates, ysp_coord_dim* coord_dim_tmp,
                          uint32_t* lookup, uint32_t* num_unique_coords, void* stream);

/* Replaces compress_unique_coords (coordinateCompressionKernel.pyx:264-419):
 * unique (row, col) pairs, per-(h,w) category dims / value starts (doubles) / counts,
 * and lookup -> value offset in doubles. Synchronous (returns the category count
 * through *num_unique_dims). */
int ysp_compress_unique_coords(const ysp_coord_dim* coord_dim_tmp, uint32_t num_unique_coords,
                               uint32_t num_coordinates,
                               uint32_t* unique_coords_out, uint16_t* unique_dims_out,
                               uint32_t* unique_dims_outer_indices,
                               uint32_t* unique_dims_block_counts,
                               uint32_t* lookup, uint32_t* num_unique_dims, void* stream);

/* ----------------------------------------------- element gradient/Hessian */
/* Shared index/output block of every element entry point; the same arguments the
 * generated compute_hessian_and_gradient_with_compression receives
 * (yasps/yasps/kernel/Hessian/hessianKernelHost.pyx:27-50, consumed as in
 * hessianKernelNoProject.pyx:16-264). gradient / hessian_blocks / diagonal_blocks are
 * accumulated into (FP64 atomics); the caller zeroes them. */
typedef struct ysp_scatter_args {
  const uint32_t* segment_indices;       /* [N x K] */
  const uint16_t* segment_sizes;         /* [N x K] */
  const int16_t* local_permutations;     /* [N x K] */
  const uint32_t* lookups;               /* value offset per coordinate */
  const uint32_t* coordinates_outer;     /* [N + 1] */
  uint32_t max_num_indices;              /* K */
  uint32_t num_attributes;               /* number of minimisation targets */
  const uint32_t* diagonal_blocks_start; /* device [num_attributes + 1] */
  const uint32_t* gradient_segments_start; /* device [num_attributes + 1] */
  double* gradient;
  double* hessian_blocks;
  double* diagonal_blocks;
  /* Staged (atomic-free) mode, used when block_staging != NULL: the element writes its
   * k-th oriented 3x3 block to block_staging[9 * (coordinates_outer[e] + k)] and its local
   * gradient to gradient_staging[e * gradient_local_size ...] with plain stores;
   * ysp_gather_blocks3 / ysp_gather_gradient3 then sum them in a fixed order. gradient,
   * hessian_blocks and diagonal_blocks are not touched in this mode. */
  double* block_staging;
  double* gradient_staging;
  uint32_t gradient_local_size;
} ysp_scatter_args;

/* Rest-shape data of the Stable Neo-Hookean tet (examples/one_bunny/helpers.py:42-54):
 * dm_inv = (TB^T)^-1 row-major [T x 9] and the signed vol = det(TB^T)/6 [T]. */
int ysp_snh_rest_data(const ysp_union_f64* rest_positions, const uint32_t* tet2v,
                      uint32_t num_tets, double* dm_inv, double* vol, void* stream);

/* Stable Neo-Hookean tet term (examples/one_bunny/helpers.py:42-67), PSD projection
 * of the 12x12 local Hessian (hessianKernelHeader.pyx:79-252, method 0/1/2) and the
 * NoProject scatter. mu/lam are device arrays read at index tet*stride (stride 0 =
 * one value for every tet). Replaces the generated kernel of the energy
 * ``stable_neo_hookean`` with projection_method = 2 (one_bunny.py:149). */
int ysp_eval_snh_tet(const ysp_union_f64* positions, const uint32_t* tet2v,
                     uint32_t num_tets, const double* dm_inv, const double* vol,
                     const double* mu, uint32_t mu_stride,
                     const double* lam, uint32_t lam_stride, double dt,
                     uint32_t projection_method, const ysp_scatter_args* out,
                     void* stream);

/* Backward-Euler inertia on the moving vertices (examples/one_bunny/helpers.py:69-71):
 * x~ = x_prev + v dt - (0, 9.8 dt^2, 0), E = 1/2 m |x - x~|^2. */
int ysp_eval_inertia(const double* position, const double* last_position,
                     const double* velocity, const double* mass, double dt,
                     uint32_t num_vertices, uint32_t projection_method,
                     const ysp_scatter_args* out, void* stream);

/* Per-element energies (value kernels of kernel/Compute/globalKernel.pyx:35-434, used
 * by minimizer.computeTotalEnergy, minimizer.pyx:396-407). */
int ysp_energy_snh_tet(const ysp_union_f64* positions, const uint32_t* tet2v,
                       uint32_t num_tets, const double* dm_inv, const double* vol,
                       const double* mu, uint32_t mu_stride, const double* lam,
                       uint32_t lam_stride, double dt, double* energy, void* stream);
int ysp_energy_inertia(const double* position, const double* last_position,
                       const double* velocity, const double* mass, double dt,
                       uint32_t num_vertices, double* energy, void* stream);

/* Material adjoint (no reference counterpart; the form of a material optimisation like
 * examples/bunny_material_optimization): dtheta[e] = -adjoint_e . dg_e/dtheta_e for
 * theta_e = log E_e of every tet, where g_e is the SNH element gradient (mu, lam
 * proportional to E make dg_e/dtheta_e = g_e) and adjoint_e gathers the adjoint vector
 * (the solution of H lambda = dL/dx) at the tet's target segments; `slots` carries the
 * term's segment indices / sizes / permutations (the other fields are unused).
 * Synchronous. */
int ysp_snh_theta_adjoint(const ysp_union_f64* positions, const uint32_t* tet2v,
                          uint32_t num_tets, const double* dm_inv, const double* vol,
                          const double* mu, uint32_t mu_stride, const double* lam,
                          uint32_t lam_stride, double dt, const ysp_scatter_args* slots,
                          const double* adjoint, double* dtheta, void* stream);

/* ------------------------------------------------ generated element energies */
/* Energies of the reference examples other than SNH/inertia, each evaluated per element
 * from its vertex data (arity vertices of `width` doubles gathered through the union
 * view and elem2v), differentiated exactly offline (yasps_amd/codegen/gen_energies.py):
 *   POINT_POINT/POINT_EDGE/POINT_TRIANGLE/EDGE_EDGE: IPC barriers
 *     kappa (d - dhat)^2 log^2(d / dhat) on squared distances
 *     (examples/one_bunny/helpers.py:74-120); scene params [dhat, kappa].
 *   BENDING: k (theta - theta0)^2 |e0| (examples/many_bunnies_one_cloth/helpers.py:193-212);
 *     element params [theta0, |e0|], scene params [k].
 *   BARAFF_WITKIN: stretch/shear (helpers.py:239-259); element params [Minv (2x2), area],
 *     scene params [stretch, shear, dt, -, -] (the last two are set per element).
 *   STVK_MEMBRANE: area dt^2 (mu |E|^2 + lam/2 tr(E)^2), no reference counterpart;
 *     element params [Minv, area], scene params [mu, lam, dt].
 *   AFFINE_ORTHOGONALITY: k/2 |A^T A - I|_F^2 on a 9-vector A
 *     (examples/one_bunny_partial_abd/helpers.py:132-139); scene params [k].
 * They replace the generated compute_hessian_and_gradient kernels of these energies
 * (hessianKernelHost.pyx:27-50) with the same scatter arguments. */
#define YSP_ENERGY_POINT_POINT 0
#define YSP_ENERGY_POINT_EDGE 1
#define YSP_ENERGY_POINT_TRIANGLE 2
#define YSP_ENERGY_EDGE_EDGE 3
#define YSP_ENERGY_BENDING 4
#define YSP_ENERGY_BARAFF_WITKIN 5
#define YSP_ENERGY_STVK_MEMBRANE 6
#define YSP_ENERGY_AFFINE_ORTHOGONALITY 7
#define YSP_ENERGY_SNH 8       /* p [Dm^-1 (9), vol], s [mu, lam, dt] (helpers.py:42-67) */
#define YSP_ENERGY_INERTIA 9   /* p [x_prev (3), v (3), m], s [dt] (helpers.py:69-71) */
/* QUADRATIC_FORM: 4 vertices (12 DoFs), y = J x + c, E = y^T Q y / 2 + l^T y; the map is
 * linear, so the projection acts on the inner Hessian Q and H = J^T spd(Q) J
 * (differentiator.pyx:795-801), as the energy of
 * tests/differentiator/test_hessian_sparsity_permutation_gpu.py:17-38 is projected there.
 * Scene params [J (6 x 12, row-major), Q (6 x 6), l (6), c (6)]: a map with m < 6 rows
 * leaves rows m.. of J, l and c zero and Q identity there. */
#define YSP_ENERGY_QUADRATIC_FORM 10

/* Union children whose vertices follow an affine body, x = t + A r
 * (examples/one_bunny_partial_abd/one_bunny_partial_abd.py:127-144): bit c of mask marks
 * child c; vertex_body[c] maps its vertices to bodies, rest_position[c] holds r (3 per
 * vertex). The union view's data for such a child holds the current x. */
typedef struct ysp_affine_children {
  uint32_t mask;
  const uint32_t* vertex_body[YSP_MAX_UNION_CHILDREN];
  const double* rest_position[YSP_MAX_UNION_CHILDREN];
} ysp_affine_children;

/* position[v] = translation[b] + affine_matrix[b] rest_position[v], b = vertex_body[v]
 * (A row-major); the current x of affine vertices for the union view. */
int ysp_affine_positions(const double* affine_matrix, const double* translation,
                         const uint32_t* vertex_body, const double* rest_position, uint32_t n,
                         double* position, void* stream);

/* Shape of an energy kind: vertices per element, doubles per vertex, element and scene
 * parameter counts. */
int ysp_energy_arity(int32_t kind, uint32_t* arity, uint32_t* width, uint32_t* element_params,
                     uint32_t* scene_params);

/* Element parameters from rest positions: BENDING [theta0, |e0|] per hinge,
 * BARAFF_WITKIN / STVK_MEMBRANE [Minv (4), thickness * rest area] per triangle (rest
 * frame of helpers.py:214-237). */
int ysp_element_rest_data(int32_t kind, const ysp_union_f64* rest_positions,
                          const uint32_t* elem2v, uint32_t n, double thickness,
                          double* element_params, void* stream);

/* Gradient and (projected) Hessian of every element scattered as ysp_eval_snh_tet does;
 * segments of any size (3 for vertices, 9 for an affine matrix). With affine children
 * (affine may be NULL) the element Hessian in vertex positions is projected and mapped to
 * the bodies' A and t segments by the chain rule. Synchronous. */
int ysp_eval_element_energy(int32_t kind, const ysp_union_f64* positions,
                            const ysp_affine_children* affine,
                            const uint32_t* elem2v, uint32_t n, const double* element_params,
                            const double* scene_params, uint32_t projection_method,
                            const ysp_scatter_args* out, void* stream);

/* Per-element energy values. */
int ysp_element_energy_values(int32_t kind, const ysp_union_f64* positions,
                              const uint32_t* elem2v, uint32_t n, const double* element_params,
                              const double* scene_params, double* energy, void* stream);

/* ------------------------------------------- mixed second derivatives d g / d theta */
/* Element blocks of a rectangular second-order Jacobian (secondOrderJacobian.compute,
 * matrixAndVector/matrix/secondOrderJacobian.pyx:531-559, with the generated kernels of
 * kernel/Jacobian/secondOrderJacobianKernelNoProject.pyx): rows = the element's DoF
 * (arity x width, connectivity order), w columns of one column target theta. For each
 * element, B = (d g / d p) dparams + (d g / d s) dscene, with dparams [P x w] (the
 * derivatives of the element parameters w.r.t. theta; inertia only, null = zero) and
 * dscene [S x w] (of the scene parameters; null = zero), each *_stride doubles apart per
 * element (0: shared); scene_params likewise scene_stride apart (per-element Lamé
 * coefficients, for instance). blocks: [n, arity x width, w]. No projection. Vertex
 * data as in ysp_eval_element_energy (no affine-body children). Asynchronous. */
int ysp_element_param_jacobian(int32_t kind, const ysp_union_f64* positions,
                               const uint32_t* element_vertices, uint32_t n,
                               const double* element_params, const double* scene_params,
                               uint32_t scene_stride, const double* dparams,
                               uint32_t dparams_stride, const double* dscene,
                               uint32_t dscene_stride, uint32_t w, double* blocks, void* stream);

/* Global row of each DoF of each element from the get_indices output of the row targets
 * (secondOrderJacobianIndicesKernel.pyx:23-299, row stream): rows[n, dofs], -1 for DoF
 * outside the targets. YSP_ERR_ARG when an element's segments do not cover its dofs.
 * Synchronous. */
int ysp_jacobian_rows(const uint32_t* segment_indices, const uint16_t* segment_sizes,
                      uint32_t max_num_indices, uint32_t n, uint32_t dofs, int32_t* rows,
                      void* stream);

/* y += B x (transpose = 0: x over the w columns of each element's block, y over rows) or
 * y += B^T x (transpose = 1) for blocks of ysp_element_param_jacobian (the matrix spmv of
 * matrixAndVector/matrix/matrix.pyx:21-86 with transpose). Columns start at e w for
 * element e, or at 0 for every element when shared_columns != 0 (theta with one
 * instance). Accumulates into y with FP64 atomics where elements share an output.
 * Asynchronous. */
int ysp_jacobian_spmv(const double* blocks, uint32_t n, uint32_t dofs, uint32_t w,
                      const int32_t* rows, int32_t shared_columns, const double* x, double* y,
                      int32_t transpose, void* stream);

/* --------------------------------------------------------- block storage */
/* Symmetric upper block storage of one Hessian part (hessian.pyx:783-808,
 * matrix.pyx:138-282). Host arrays describe the (h, w) categories. */
typedef struct ysp_block_matrix {
  const double* values;              /* blocks_flattened */
  const uint32_t* positions;         /* device [2 x nb] (row, col) scalar starts */
  uint32_t num_categories;
  const uint32_t* starts;            /* host [num_categories] value offsets */
  const uint32_t* counts;            /* host [num_categories] */
  const uint32_t* dims;              /* host [2 x num_categories] (h, w) */
} ysp_block_matrix;

/* y (+)= H x for symmetric upper storage: y_r += B x_c and, r != c, y_c += B^T x_r
 * (matrix.pyx:21-86 / solverKernel.pyx:157-274). transpose applies H^T (identical for
 * square symmetric storage). accumulate=0 overwrites y. */
int ysp_block_spmv(const ysp_block_matrix* part, const ysp_block_matrix* dynamic_part,
                   const double* x, double* y, uint32_t n, int transpose, int accumulate,
                   void* stream);

/* ------------------------------------------------------------ solver */
/* Replaces invert_diagonal_blocks (kernel/Solver/diagonalBlockInverseKernel.pyx:
 * 107-136): per-attribute blocks [count x size x size]; eigenvalues |l| < 1e-6 are
 * kept as |l|, others become |1/l| (``:69,91``). starts/counts/sizes are host arrays
 * (starts in doubles). */
int ysp_invert_diagonal_blocks(const double* diagonal_blocks, double* inverse,
                               const uint32_t* starts, const uint32_t* counts,
                               const uint32_t* sizes, int32_t num_attributes, void* stream);

/* Scalar diagonal of the assembled Hessian from its diagonal blocks:
 * diagonal[gradient_segments_start[a] + i * sizes[a] + k] = block (a, i) entry (k, k).
 * Replaces the `diagonal` output of the FullProject element kernels
 * (kernel/Hessian/hessianKernelFullProject.pyx:187-192), read through hessian.diagonal
 * (matrixAndVector/matrix/hessian.pyx:188) and scene.diagonal (scene/scene.pyx:186-188).
 * starts (doubles), counts, sizes and gradient_segments_start are host arrays. */
int ysp_diagonal_from_blocks(const double* diagonal_blocks, double* diagonal,
                             const uint32_t* starts, const uint32_t* counts,
                             const uint32_t* sizes, const uint32_t* gradient_segments_start,
                             int32_t num_attributes, void* stream);

typedef struct ysp_pcg_stats {
  int32_t iterations;
  int32_t status;              /* return code, see ysp_pcg_solve */
  double delta0;               /* r0^T M^-1 r0 */
  double delta;                /* final r^T M^-1 r */
  double rr0;                  /* ||b||^2 */
  double rr;                   /* final recurrence ||r||^2 */
} ysp_pcg_stats;

#define YSP_PCG_REFERENCE_RATIO 0   /* stop when r^T M^-1 r <= tol * r0^T M^-1 r0 */
#define YSP_PCG_RELATIVE_RESIDUAL 1 /* stop when ||r|| <= tol * ||b|| (recurrence r) */

/* Replaces computeSolution (kernel/Solver/solverKernel.pyx:477-687): block-Jacobi PCG
 * with device-resident recurrence scalars. threshold_mode selects the stopping rule.
 * work must hold 5 * n doubles, cg_scalars YSP_PCG3_SCALARS doubles. Synchronous. Returns the
 * iteration count (>= 0), -4 when max_iterations is reached, -5 for an invalid
 * initial residual, -(it + 4) for a non-positive curvature at iteration it. */
int ysp_pcg_solve(uint32_t max_iterations, double threshold, int32_t threshold_mode,
                  const ysp_block_matrix* part, const ysp_block_matrix* dynamic_part,
                  const double* diagonal_blocks_inverse,
                  const uint32_t* diagonal_blocks_start, const uint32_t* diagonal_blocks_count,
                  const uint32_t* diagonal_blocks_size, const uint32_t* gradient_segments_start,
                  int32_t num_attributes,
                  const double* rhs, uint32_t n,
                  double* solution, const double* initial_guess, int32_t zero_initial_guess,
                  double* work, double* cg_scalars, ysp_pcg_stats* stats, void* stream);

/* ysp_pcg_solve with the operator given as the split mirrored full block-CSR of
 * ysp_bsr3_build_split / ysp_bsr3_gather_split (gathers instead of FP64 atomics; same
 * block-Jacobi preconditioner over the diagonal blocks of every target, recurrence,
 * stopping rules and return codes). n must be a multiple of 3. Block rows with more than
 * long_row_threshold entries (0: 128) get a workgroup each. values FP64 or, with
 * values_f32, FP32 (the inner solve of mixed precision; vectors, dots and the diagonal
 * block inverses stay FP64). */
int ysp_pcg_solve_full(uint32_t max_iterations, double threshold, int32_t threshold_mode,
                       const uint32_t* row_ptr, const uint32_t* col, const void* values,
                       int32_t values_f32, uint32_t nnz, uint32_t long_row_threshold,
                       const double* diagonal_blocks_inverse,
                       const uint32_t* diagonal_blocks_start,
                       const uint32_t* diagonal_blocks_count,
                       const uint32_t* diagonal_blocks_size,
                       const uint32_t* gradient_segments_start, int32_t num_attributes,
                       const double* rhs, uint32_t n, double* solution,
                       const double* initial_guess, int32_t zero_initial_guess, double* work,
                       double* cg_scalars, ysp_pcg_stats* stats, void* stream);

/* --------------------------------------------------------------- MAS preconditioner */
/* Multilevel additive Schwarz (solver/mas): the hierarchy (domains, identity
 * aggregation maps) is built on the host once per static structure
 * (solver/mas/hierarchy.py:150-271); these entry points do the per-solve numeric work
 * and the apply on the device. Nodes are 3-scalar slots (a larger node of the
 * hierarchy is several slots sharing its domain); domains are padded to 32 scalars (the
 * configurations' leaf size) or 48 (the reference's default, solver/mas/solver.py:91). */

/* values (level l, AoS 9 doubles per block) = sum of the level-(l-1) blocks listed in
 * items[offsets[e] .. offsets[e+1]) (identity-aggregation Galerkin product); the previous
 * level is either the solve loop's full block-CSR (component-major) or an AoS level. */
int ysp_mas_coarsen(const double* prev_values, int32_t prev_is_full_bsr, uint32_t prev_nnz,
                    const uint32_t* offsets, const uint32_t* items, uint32_t nnz,
                    double* values, void* stream);

/* Copy every block e with dest[e] != ~0u into the dense domain matrices (FP64, D x D
 * row-major per domain, D = domain_size: 32 (or 0) or 48) at offset dest[e]
 * (numeric_assembly.py:47-61). */
int ysp_mas_domain_scatter(const double* values, int32_t is_full_bsr, uint32_t nnz,
                           const uint32_t* dest, double* domains, uint32_t domain_size,
                           void* stream);

/* Explicit inverse of every domain matrix (identity in the padding), FP64 Gauss-Jordan,
 * stored FP32 column-major (solver/mas/cuda/local_inverse.cu:6-308). With inverses_bf16
 * non-NULL, domains d >= bf16_first_domain are stored BF16 (round to nearest even) at
 * inverses_bf16[(d - bf16_first_domain) * D * D] instead (the BF16 coarse banks of the
 * mixed-precision MAS); the others stay in inverses[d * D * D] (D = domain_size: 32 (or
 * 0) or 48). *status (device int) gets bit 0 set when a pivot is below
 * 1e-12 * max|diag|. */
int ysp_mas_invert_domains(double* domains, const uint32_t* domain_sizes, uint32_t num_domains,
                           float* inverses, uint16_t* inverses_bf16, uint32_t bf16_first_domain,
                           int32_t* status, uint32_t domain_size, void* stream);

#define YSP_MAS_MAX_COARSE_LEVELS 8

typedef struct ysp_mas_args {
  const uint32_t* restrict_offsets;     /* [num_coarse_nodes + 1] */
  const uint32_t* restrict_items;       /* level-0 descendants of the nodes of levels 1
                                           to 3; level-3 descendants (coarse ids) of the
                                           nodes of deeper levels */
  uint32_t num_coarse_nodes;            /* coarse levels concatenated */
  uint32_t num_coarse_levels;           /* <= YSP_MAS_MAX_COARSE_LEVELS */
  uint32_t level_offsets[YSP_MAS_MAX_COARSE_LEVELS + 1];  /* coarse ids of level l + 1:
                                           [level_offsets[l], level_offsets[l + 1]) */
  const uint32_t* prolong_maps;         /* [num_coarse_levels x n/3] coarse ids */
  const float* inverses;                /* [num_domains x D x D] */
  const uint32_t* domain_nodes;         /* level-0 nodes or coarse ids */
  const uint32_t* domain_node_offsets;  /* [num_domains + 1] */
  uint32_t num_domains;
  uint32_t level0_domains;              /* domains [0, level0_domains) are on level 0 */
  const float* weights;                 /* level weights [1 + num_coarse_levels] */
  double* coarse_residual;              /* work [3 x num_coarse_nodes] */
  float* fine_correction;               /* work [n] */
  float* coarse_correction;             /* work [3 x num_coarse_nodes] */
  const uint16_t* inverses_bf16;        /* NULL, or the BF16 bank of domains >= */
  uint32_t bf16_first_domain;           /*   bf16_first_domain (ysp_mas_invert_domains) */
  const float* domain_scale;            /* NULL, or a factor per domain on its correction:
                                           duplicate_level_weight for a domain that repeats
                                           the previous level's (preconditioner_specialized.cu
                                           :415-440), 1 otherwise */
  uint32_t domain_size;                 /* D: 32 (or 0) or 48 scalars per padded domain */
} ysp_mas_args;

/* z = sum_l w_l P_l S_l B_l^-1 R_l r (solver/mas/preconditioner.py:128-154; S_l the
 * per-domain factors of domain_scale, the identity when it is NULL). */
int ysp_mas_apply(const ysp_mas_args* args, const double* r, uint32_t n, double* z,
                  void* stream);

/* --------------------------------------------- atomic-free (staged) assembly */
/* No reference counterpart: the reference accumulates element blocks with FP64 atomics
 * (hessianKernelNoProject.pyx:231-258). These entry points sum staged element output in
 * a fixed order instead (see ysp_scatter_args.block_staging). */

/* Stable inverse of a key map: items lists values[i] (or i when values is NULL) grouped
 * by keys[i] in ascending key order, offsets[num_keys + 1] delimits each key's run.
 * Keys >= num_keys (e.g. ~0u for "no contribution") are dropped. */
int ysp_inverse_map(const uint32_t* keys, const uint32_t* values, uint32_t n, uint32_t num_keys,
                    uint32_t* offsets, uint32_t* items, void* stream);

/* Gradient contribution keys/values of one term for 3-vector segments: for slot
 * (e, i) with index >= 2, key = (index - 2) / 3 and value = staging_base + e * local size +
 * the slot's local offset; other slots get ~0u. keys/values hold N * K entries. */
int ysp_gradient_contributions(const uint32_t* segment_indices, const uint16_t* segment_sizes,
                               uint32_t num_instances, uint32_t K, uint64_t staging_base,
                               uint32_t gradient_local_size, uint32_t* keys, uint32_t* values,
                               void* stream);

/* values[u] = sum of the staged 3x3 blocks of unique block u (offsets/items from
 * ysp_inverse_map over the coordinates' unique ordinals); (v, v) blocks also go to
 * diagonal_blocks (added when accumulate_diagonal is set). */
int ysp_gather_blocks3(const double* block_staging, const uint32_t* offsets,
                       const uint32_t* items, uint32_t num_blocks, const uint32_t* positions,
                       double* values, double* diagonal_blocks,
                       const uint32_t* diagonal_blocks_start,
                       const uint32_t* gradient_segments_start, uint32_t num_attributes,
                       int32_t accumulate_diagonal, void* stream);

/* gradient[3v..3v+2] = sum of the staged gradient segments of vertex v. */
int ysp_gather_gradient3(const double* gradient_staging, const uint32_t* offsets,
                         const uint32_t* items, uint32_t n, double* gradient, void* stream);

/* ---------------------------------------------- full 3x3 block-CSR (solve loop) */
/* The Krylov loop runs on a mirrored full block-CSR built from the upper storage of
 * the static and dynamic parts (all blocks 3x3): row_ptr [n/3 + 1], col [nnz] (block
 * column), src [nnz] (bit 31 = stored transposed, bit 30 = dynamic part, low bits =
 * upper block index). No reference counterpart: it is this build's layout of the
 * operator spmvWithSystem applies (kernel/Solver/solverKernel.pyx:217-274). capacity
 * bounds nnz (2 * (nb_static + nb_dynamic) always suffices). Synchronous. */
int ysp_bsr3_build(const uint32_t* pos_static, uint32_t nb_static,
                   const uint32_t* pos_dynamic, uint32_t nb_dynamic, uint32_t n,
                   uint32_t* row_ptr, uint32_t* col, uint32_t* src, uint32_t capacity,
                   uint32_t* nnz, void* stream);

/* Component-major values of the full block-CSR: out[j * pad64(nnz) + k] is entry j
 * (row-major) of block k; FP64 and/or FP32 copies (either pointer may be NULL). */
int ysp_bsr3_gather_values(const double* static_values, const double* dynamic_values,
                           const uint32_t* src, uint32_t nnz, double* out64, float* out32,
                           void* stream);

/* The same mirrored full 3x3 block-CSR for upper storage whose block dimensions are
 * multiples of 3 (affine-body segments give 9x9, 9x3, 3x9 next to 3x3 blocks): every
 * stored h x w block is split into (h/3)(w/3) 3x3 sub-blocks, each entered at (r, c) and
 * mirrored at (c, r) (the sub-blocks of a diagonal block, which holds both triangles, are
 * entered once). src [nnz] (64-bit): bits 0-39 value offset of the sub-block's first
 * element in its part, bits 40-41 w/3 - 1, bit 42 transposed, bit 43 dynamic part.
 * capacity bounds nnz (the sum of 2 (h/3)(w/3) over all blocks suffices). No reference
 * counterpart: the operator of spmvWithSystem (kernel/Solver/solverKernel.pyx:217-274)
 * laid out for gathers. Synchronous. */
int ysp_bsr3_build_split(const ysp_block_matrix* part, const ysp_block_matrix* dynamic_part,
                         uint32_t n, uint32_t* row_ptr, uint32_t* col, uint64_t* src,
                         uint32_t capacity, uint32_t* nnz, void* stream);

/* Component-major FP64 values of the split block-CSR (layout as ysp_bsr3_gather_values). */
int ysp_bsr3_gather_split(const double* static_values, const double* dynamic_values,
                          const uint64_t* src, uint32_t nnz, double* out64, void* stream);

/* Value sources of the solve loop's operator. Every entry point that takes
 * (values, values_f32) for the full or split block-CSR reads values_f32 as a format:
 * YSP_VALUES_F64 / YSP_VALUES_F32: values is the component-major copy
 * (ysp_bsr3_gather_values / ysp_bsr3_gather_split, or a cast of it);
 * YSP_VALUES_UPPER: values points to a host ysp_upper_values, and the products read the
 * upper block storage through each entry's source word (no mirrored copy: every stored
 * block is read by both of its entries, about half the value bytes of the mirrored copy
 * for operators too large to stay in the 256 MB Infinity Cache). Same arithmetic, same
 * results as the FP64 copy. The batched PCG takes the component-major formats only. */
#define YSP_VALUES_F64 0
#define YSP_VALUES_F32 1
#define YSP_VALUES_UPPER 2

typedef struct ysp_upper_values {
  const void* src;                   /* per entry: uint32 (ysp_bsr3_build) or uint64 (_split) */
  int32_t src_64;                    /* 1: 64-bit split source words */
  const double* static_values;       /* upper storage of the static part (blocks_flattened) */
  const double* dynamic_values;      /* of the dynamic part (NULL: none) */
} ysp_upper_values;

/* out[i] = (float) in[i] (FP32 copies of FP64 device arrays, e.g. inverse blocks). */
int ysp_cast_f32(const double* in, float* out, uint64_t n, void* stream);

/* One step of iterative refinement, no reference counterpart (the mixed-precision form of
 * the solve): solution += correction (skipped when correction is NULL), then
 * residual = rhs - H solution over the full block-CSR with FP64 arithmetic, and the
 * squared norms of residual and rhs returned on the host. Synchronous. */
int ysp_bsr3_refine(const uint32_t* row_ptr, const uint32_t* col, const void* values,
                    int32_t values_f32, uint32_t nnz, uint32_t n, const double* rhs,
                    double* solution, const double* correction, double* residual,
                    double* residual_sq, double* rhs_sq, void* stream);

/* y = H x over the full block-CSR (values FP64 or FP32, arithmetic FP64). */
int ysp_bsr3_spmv(const uint32_t* row_ptr, const uint32_t* col, const void* values,
                  int32_t values_f32, uint32_t nnz, uint32_t n, const double* x, double* y,
                  void* stream);

/* Block-Jacobi PCG on the full block-CSR with 3x3 diagonal blocks (the fast path of
 * ysp_pcg_solve: two kernels per iteration, direction update fused into the SpMV
 * gather). Same recurrence, stopping rules and return codes as ysp_pcg_solve; values
 * and the inverse diagonal blocks may be FP32 (FP64 vectors, dots and scalars).
 * work holds 5 * n doubles, cg_scalars YSP_PCG3_SCALARS doubles. Synchronous. */
#define YSP_PCG3_SCALARS 12288

int ysp_pcg3_solve(uint32_t max_iterations, double threshold, int32_t threshold_mode,
                   const uint32_t* row_ptr, const uint32_t* col, const void* values,
                   int32_t values_f32, uint32_t nnz, const void* diagonal_inverse,
                   int32_t inverse_f32, const double* rhs, uint32_t n, double* solution,
                   const double* initial_guess, int32_t zero_initial_guess, double* work,
                   double* cg_scalars, ysp_pcg_stats* stats, void* stream);

/* The same PCG with the MAS preconditioner (solver/mas/pcg.py:23-102 recurrence on the
 * device): stop on r^T M^-1 r <= tol * r0^T M^-1 r0, or on the relative residual. work
 * holds 5 * n doubles; the MAS inverses must be current (ysp_mas_invert_domains). */
int ysp_pcg3_mas_solve(uint32_t max_iterations, double threshold, int32_t threshold_mode,
                       const uint32_t* row_ptr, const uint32_t* col, const void* values,
                       int32_t values_f32, uint32_t nnz, const ysp_mas_args* mas,
                       const double* rhs, uint32_t n, double* solution,
                       const double* initial_guess, int32_t zero_initial_guess, double* work,
                       double* cg_scalars, ysp_pcg_stats* stats, void* stream);

/* One product pass of the PCG loop on caller state, no reference counterpart: the kernel
 * that dominates every block-Jacobi and MAS iteration of ysp_pcg3_solve /
 * ysp_pcg3_mas_solve (q = A z + beta q, p = z + beta p, partial sums of p.q), exposed so
 * that a benchmark can time the loop's own kernel with events on its stream. With
 * prepare != 0 the scalar state is first set to a running iteration (beta = 1, never
 * converged) and z, p, q to nonzero data; later passes keep updating p and q. Every block
 * row goes through the product kernel (no long-row list). work: 5 n doubles laid out as
 * ysp_pcg3_solve's, cg_scalars: YSP_PCG3_SCALARS doubles. Asynchronous. */
int ysp_pcg3_product_pass(const uint32_t* row_ptr, const uint32_t* col, const void* values,
                          int32_t values_f32, uint32_t nnz, uint32_t n, double* work,
                          double* cg_scalars, int32_t prepare, void* stream);

/* No reference counterpart. ysp_pcg3_solve for callers that cannot allocate memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * work vectors and the scalar state live in a device array of the library, one per GPU
 * (held until the process ends) of YSP_PCG3_SCALARS + 5 * YSP_PCG3_NOALLOC_MAX_N doubles
 * (15.8 MB), looked up on the first call on each device (the stream's device must be the
 * current one). n above YSP_PCG3_NOALLOC_MAX_N is rejected with YSP_ERR_ARG. No device,
 * pinned or library allocation, no events: every block row goes through the one product
 * kernel and the host reads the status with stream synchronisations. All calls on one
 * GPU share the array, so calls that overlap on one GPU (other streams or threads) race.
 * Same recurrence, stopping rules and return codes as ysp_pcg3_solve. Synchronous. */
#define YSP_PCG3_NOALLOC_MAX_N 393216

int ysp_pcg3_solve_noalloc(uint32_t max_iterations, double threshold, int32_t threshold_mode,
                           const uint32_t* row_ptr, const uint32_t* col, const void* values,
                           int32_t values_f32, uint32_t nnz, const void* diagonal_inverse,
                           int32_t inverse_f32, const double* rhs, uint32_t n, double* solution,
                           const double* initial_guess, int32_t zero_initial_guess,
                           ysp_pcg_stats* stats, void* stream);

/* ------------------------------------------------------------ batched PCG (serving) */
/* No reference counterpart: several independent 3x3 systems (e.g. the Newton solves of
 * different requests) advance through the same launches; every workgroup owns a chunk
 * of one system and reductions are segmented per system (each system's own scalar
 * state). Each system is preconditioned by block-Jacobi (diagonal_inverse) or by its own
 * MAS (mas, the apply of solver/mas/pcg.py:23-102 batched over the systems: restriction,
 * local solves and prolongation each one launch for all of them). Zero initial guess,
 * same recurrence, stopping rules and return codes as ysp_pcg3_solve /
 * ysp_pcg3_mas_solve. */
#define YSP_PCG3_RUNNING 0x7fffffff   /* status of a system still iterating */

typedef struct ysp_pcg3_system {
  const uint32_t* row_ptr;            /* full block-CSR (ysp_bsr3_build) */
  const uint32_t* col;
  const void* values;                 /* component-major, FP64 or FP32 (per batch) */
  uint32_t nnz;
  const void* diagonal_inverse;       /* 9 per vertex, FP64 or FP32 (per batch) */
  const double* rhs;
  uint32_t n;
  double* solution;
  double* work;                       /* 5 n doubles */
  double* cg_scalars;                 /* YSP_PCG3_SCALARS doubles */
  uint32_t max_iterations;
  uint32_t passes_done;               /* managed by the library */
  int32_t status;                     /* YSP_PCG3_RUNNING, or the ysp_pcg3_solve code */
  int32_t iterations;
  double delta0, delta, rr;           /* r0^T z0, r^T z, r^T r of the last test */
  const ysp_mas_args* mas;            /* NULL: block-Jacobi (diagonal_inverse); else the MAS
                                         preconditioner of this system (current inverses,
                                         its own coarse and correction buffers) */
} ysp_pcg3_system;

/* Initial residual, preconditioned residual and threshold of one system (asynchronous). */
int ysp_pcg3_batch_start(ysp_pcg3_system* system, double threshold, int32_t threshold_mode,
                         int32_t inverse_f32, void* stream);

/* `passes` PCG iterations of every system (finished systems idle), then their status
 * (synchronous). */
int ysp_pcg3_batch_run(ysp_pcg3_system* systems, uint32_t count, uint32_t passes,
                       int32_t values_f32, int32_t inverse_f32, void* stream);

/* -------------------------------------------------------------- contact sets */
/* The collision module the reference's examples drive (examples/ccd/ccd.py over
 * examples/ccd/mlbvh.cu and examples/ccd/ACCD.cu). Positions x and directions d are
 * [vertices x 3] FP64; faces uint32 [nf x 3] with surface the vertices queried against
 * them (uint32 [ns]); edges uint32 [ne x 2]; mesh (NULL, or uint32 [vertices]) suppresses
 * pairs whose four vertices share one nonzero id. dhat is a squared distance. */

/* Candidates: a linear BVH over the faces and one over the edges (boxes of the vertices
 * at x, and at x - alpha d when d is non-NULL), every surface vertex queried against the
 * faces and every edge against the edges with partner id above its own, the gap
 * sqrt(dhat) (mlbvh.cu:1015-1137; lbvh_{f,e}_construct[_full_ccd] +
 * lbvh_{f,e}_append_{proximity,swept}_candidates, mlbvh.cu:1603-1700,1906-1946). Writes
 * compact int32 pairs (-(p + 1), face) or (edge i, edge j) into
 * candidates[2 * capacity]; *count (host) gets the number found, which may exceed
 * capacity (only the first capacity are stored). Synchronous. */
int ysp_ccd_candidates(const double* x, const double* d, double alpha, uint32_t num_vertices,
                       const uint32_t* faces, uint32_t num_faces, const uint32_t* surface,
                       uint32_t num_surface, const uint32_t* edges, uint32_t num_edges,
                       const uint32_t* mesh, double dhat, int32_t* candidates,
                       uint32_t capacity, uint32_t* count, void* stream);

/* Active pairs of the candidates at positions x: the distance type of each (vertex, edge
 * or face of the triangle; the nine segment cases of two edges), its squared distance and,
 * below dhat, the pair of that type appended to pp [2 ids], pe [3], pt [4] or ee [4], each
 * with room for capacity pairs (process_cached_candidates + scatter_packed_cases,
 * mlbvh.cu:594-725,1139-1300; lbvh_refilter_cached_candidates + lbvh_scatter_packed_cases,
 * mlbvh.cu:1948-2021). counts[4] (host) get the numbers found (may exceed capacity).
 * Order within a case is not fixed. Synchronous. */
int ysp_ccd_active_pairs(const double* x, const uint32_t* faces, const uint32_t* edges,
                         const int32_t* candidates, uint32_t num_candidates, double dhat,
                         uint32_t* pp, uint32_t* pe, uint32_t* pt, uint32_t* ee,
                         uint32_t capacity, uint32_t* counts, void* stream);

/* Largest feasible step along x - t d (self_largestFeasibleStepSizeCompact,
 * ACCD.cu:676-805): the additive CCD time of every candidate with eta = 1 - slackness
 * (times[i] when times is non-NULL), their minimum in *step (host), 1 without
 * candidates. A candidate already at zero distance gets 0. Synchronous. */
int ysp_ccd_step_bound(const double* x, const double* d, const uint32_t* faces,
                       const uint32_t* edges, const int32_t* candidates, uint32_t num_candidates,
                       double slackness, double* times, double* step, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* YASPS_AMD_H */
