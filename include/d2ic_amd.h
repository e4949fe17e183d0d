/*
 * d2ic_amd.h -- C ABI of the MI355X-native global FE-DIC hot path.
 *
 * Every entry point takes plain device pointers (allocated by the caller;
 * the Python layer uses torch tensors for ownership), sizes, and a HIP stream
 * passed as void* (NULL = default stream).  Calls are asynchronous on that
 * stream unless the comment says they return a host value (those synchronise
 * the stream).  Every call returns a status code; d2ic_last_error() gives the
 * thread-local message of the last failure.
 *
 * The reference (D2IC 1.0.1) is pure Python/JAX and has no FFI; each entry
 * point names the reference function whose behaviour it replaces.
 */
#ifndef D2IC_AMD_H
#define D2IC_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define D2IC_OK 0
#define D2IC_ERR_INVALID 1 /* invalid argument (Python: ValueError) */
#define D2IC_ERR_NOMEM 2   /* device allocation / budget exhausted (Python: RuntimeError) */
#define D2IC_ERR_DEVICE 3  /* HIP runtime or kernel failure (Python: RuntimeError) */
#define D2IC_ERR_STATE 4   /* lifecycle error, e.g. no reference set (Python: RuntimeError) */
#define D2IC_ERR_TIMEOUT 5 /* a persistent kernel's bounded wait expired (Python: RuntimeError) */

#define D2IC_INTERP_CUBIC 0
#define D2IC_INTERP_LINEAR 1

/* Message of the last failed call on this thread ("" if none). */
const char* d2ic_last_error(void);
/* ABI version (major*100 + minor). */
int d2ic_version(void);

/* ---------------------------------------------------------------------------
 * Pixel assets (replaces build_pixel_assets, d2ic/pixel_assets.py:21-112)
 * ------------------------------------------------------------------------ */

/* Candidate stage: for every pixel of the bbox rectangle rows [i0,i1) x cols
 * [j0,j1) (row-major, _sample_pixels_in_bbox, pixel_assets.py:115-132) apply
 * the ROI filter (roi: HxW uint8 or NULL, pixel_assets.py:69-73) and the
 * first-match element search (build_pixel_to_element_mapping_numpy,
 * pixel_assets.py:135-237).  cand_elt: (i1-i0)*(j1-j0) int32, -1 when the pixel
 * is dropped.  Writes the number of kept pixels to *n_kept (host; synchronises). */
int d2ic_pixel_map(const float* nodes, int64_t n_nodes, const int32_t* elements, int64_t n_elems, int32_t H,
                   int32_t W, int32_t i0, int32_t i1, int32_t j0, int32_t j1, const uint8_t* roi, int32_t* cand_elt,
                   int64_t* n_kept, void* stream);

/* Fill stage: stable compaction of the kept candidates (row-major order), then
 * pixel_coords_ref (np,2) f32, pixel_nodes (np,4) i32, pixel_elements (np,) i32,
 * roi_mask_flat (np,) i32 (pixel_assets.py:76-98) and pixel_shapeN (np,4) f32
 * from 10 Newton steps (compute_pixel_shape_functions_jax, pixel_assets.py:240-294). */
int d2ic_pixel_fill(const int32_t* cand_elt, int32_t H, int32_t W, int32_t i0, int32_t i1, int32_t j0, int32_t j1,
                    const float* nodes, const int32_t* elements, int64_t n_pixels, float* pixel_coords,
                    int32_t* pixel_nodes, float* pixel_shapeN, int32_t* pixel_elements, int32_t* roi_flat,
                    void* stream);

/* Node->pixel table, degrees: node_pixel_degree (n_nodes,) and the maximum
 * degree (host; synchronises) (build_node_pixel_dense, pixel_assets.py:459-498). */
int d2ic_node_pixel_degree(const int32_t* pixel_nodes, int64_t n_pixels, int64_t n_nodes, int32_t* degree,
                           int32_t* max_degree, void* stream);

/* Node->pixel table, fill: node_pixel_index (n_nodes,max_degree) i32 padded
 * with -1 and node_N_weight f32 padded with 0; each node lists its pixels in
 * ascending pixel order, i.e. the slot order of the sequential scan in
 * _build_node_pixel_dense_jax (pixel_assets.py:501-526). */
int d2ic_node_pixel_table(const int32_t* pixel_nodes, const float* pixel_shapeN, int64_t n_pixels, int64_t n_nodes,
                          int32_t max_degree, int32_t* node_pixel_index, float* node_N_weight, void* stream);

/* ---------------------------------------------------------------------------
 * Global NLCG solver (replaces GlobalCGSolver / _cg_solve,
 * d2ic/solver_global_cg.py:35-469)
 * ------------------------------------------------------------------------ */
typedef struct d2ic_cg_plan d2ic_cg_plan;

/* Build the solver's internal layout from PixelAssets arrays: pixels grouped
 * by element into tiles of <= 256 pixels (structure-of-arrays; elements wider
 * than 24 pixels by column strips of <= 16 pixels; a tile spans few enough
 * pixel rows for its image box to stage in LDS), the
 * node->tile gather table, and the spring-regularisation stencil
 * (node_neighbor_* and node_reg_weight, solver_global_cg.py:261-286).
 * pixel_elements may be NULL (elements are then identified by their node tuple). */
int d2ic_cg_plan_create(const float* pixel_coords, const int32_t* pixel_nodes, const float* pixel_shapeN,
                        const int32_t* pixel_elements, int64_t n_pixels, int64_t n_nodes,
                        const int32_t* nb_index, const int32_t* nb_degree, const float* nb_weight,
                        const float* reg_weight, int32_t nb_max_degree, d2ic_cg_plan** plan, void* stream);

/* The same plan with a chosen tiling (no reference counterpart; what the
 * tile-shape autotuner varies): tiles close at tile_px pixels (64..256; 256
 * in d2ic_cg_plan_create) and elements wider than 24 pixels are cut into
 * column strips of at most strip_w pixels (4..24; 16 in d2ic_cg_plan_create).
 * Every tiling solves the same problem; the tile partials are summed in other
 * groupings, so results differ from the default tiling by rounding only. */
int d2ic_cg_plan_create_tiled(const float* pixel_coords, const int32_t* pixel_nodes, const float* pixel_shapeN,
                              const int32_t* pixel_elements, int64_t n_pixels, int64_t n_nodes,
                              const int32_t* nb_index, const int32_t* nb_degree, const float* nb_weight,
                              const float* reg_weight, int32_t nb_max_degree, int32_t tile_px, int32_t strip_w,
                              d2ic_cg_plan** plan, void* stream);
int d2ic_cg_plan_destroy(d2ic_cg_plan* plan);

/* Sample the reference image f at every pixel once (the reference resamples
 * it in every pass, solver_global_cg.py:221,225).  interp: D2IC_INTERP_*. */
int d2ic_cg_set_reference(d2ic_cg_plan* plan, const float* ref_image, int32_t H, int32_t W, int32_t interp,
                          void* stream);

/* Sizes of the plan's layout (for budgeting and roofline accounting). */
int d2ic_cg_plan_info(const d2ic_cg_plan* plan, int64_t* n_pixels, int64_t* n_nodes, int64_t* n_tiles,
                      int64_t* device_bytes);

/* One solve (one problem per launch): u (n_nodes,2) f32 holds u0 on entry and
 * the solution on exit.  hist (max_iter,3) f32 may be NULL; rows
 * [J_trial, |grad|, alpha], NaN padded.  n_iters: one int32 on the device.
 * pass_count (may be NULL): one int32 on the device, number of pixel passes run. */
int d2ic_cg_solve(d2ic_cg_plan* plan, const float* def_image, int32_t H, int32_t W, float* u, float alpha_reg,
                  int32_t max_iter, float tol, float* hist, int32_t* n_iters, int32_t* pass_count, void* stream);

/* Many independent solves in one persistent launch with per-problem
 * convergence masks (no reference counterpart; the batched form of
 * d2ic_cg_solve).  All arrays have n_problems entries; hist/pass_count
 * entries may be NULL.  A plan must not appear twice in one batch. */
int d2ic_cg_solve_batch(int32_t n_problems, d2ic_cg_plan* const* plans, const float* const* def_images,
                        const int32_t* H, const int32_t* W, float* const* u, const float* alpha_reg,
                        const int32_t* max_iter, const float* tol, float* const* hist, int32_t* const* n_iters,
                        int32_t* const* pass_count, void* stream);

/* d2ic_cg_solve_batch without any allocation (no reference counterpart): for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * kernels and results (bitwise) as d2ic_cg_solve_batch, for at most 8
 * problems: the per-launch descriptors, solver states and workspace live in
 * one static device array of ~0.6 MB per GPU (held until the process ends)
 * instead of a hipMallocAsync block, and descriptors travel as a kernel
 * argument instead of through pinned staging memory.  More than 8 problems
 * is rejected with D2IC_ERR_INVALID.  Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU (other streams or threads) race.  It
 * records no timing events (d2ic_cg_kernel_ms then reports no timed launch). */
int d2ic_cg_solve_batch_noalloc(int32_t n_problems, d2ic_cg_plan* const* plans, const float* const* def_images,
                                const int32_t* H, const int32_t* W, float* const* u, const float* alpha_reg,
                                const int32_t* max_iter, const float* tol, float* const* hist,
                                int32_t* const* n_iters, int32_t* const* pass_count, void* stream);

/* Grid size (blocks, rounded up to a multiple of 8 and capped by what can be
 * co-resident) of the persistent kernel for single-problem solves with this
 * plan; 0 restores the work-based default.  Results do not depend on it
 * (every reduction is order-independent); only speed does.  No reference
 * counterpart (launch tuning, used by CGPlan.autotune). */
int d2ic_cg_plan_set_grid(d2ic_cg_plan* plan, int32_t blocks);

/* Duration in ms of the last persistent CG kernel launch that this plan led
 * (first plan of the batch), from HIP events recorded around that launch on
 * its stream; waits for the launch to finish.  No reference counterpart
 * (measurement hook). */
int d2ic_cg_kernel_ms(const d2ic_cg_plan* plan, float* ms);

/* ---------------------------------------------------------------------------
 * Served sessions: continuous batching of solves in one persistent kernel.
 * No reference counterpart (the reference solves one frame per call,
 * solver_global_cg.py:155-198); each request is the same solve as
 * d2ic_cg_solve and gives bitwise the same result.
 * ------------------------------------------------------------------------ */
typedef struct d2ic_server d2ic_server;

/* Create a session with `slots` (1..64) concurrent requests.  The kernel runs
 * on the session's own stream while requests are in flight and leaves after
 * idle_ms without work (it is relaunched by the next submit or poll). */
int d2ic_serve_create(int32_t slots, double idle_ms, d2ic_server** server);

/* Queue one solve (same arguments as d2ic_cg_solve, plus u0).  u receives
 * the solution; u0 (NULL: u itself) holds the initial guess and is only read,
 * so a stream can start frame k+1 from frame k's result buffer without a
 * device copy.  Inputs must be ready on the device (synchronise their producer
 * first); hist, if given, must already hold NaN.  The request joins the
 * running kernel at its next grid barrier.  *slot = -1 (and *ticket = 0) when
 * every slot is busy: poll, then retry.  A plan may have one request in flight
 * at a time.  Thread-safe.  The request's iteration count (n_iters and the
 * poll's copy) is -2 when the final objective or gradient is not finite and
 * -3 when the objective at u0 is not finite (a NaN/Inf in the frame or in u0
 * where the solve samples it; the request then ends after that first pass);
 * u is not written in either case. */
int d2ic_serve_submit(d2ic_server* server, d2ic_cg_plan* plan, const float* def_image, int32_t H, int32_t W,
                      float* u, const float* u0, float alpha_reg, int32_t max_iter, float tol, float* hist,
                      int32_t* n_iters, int32_t* pass_count, int32_t* slot, int64_t* ticket);

/* Non-blocking completion check of (slot, ticket).  *done = 1 once u, hist,
 * n_iters and pass_count hold the result; the slot is then free.  n_iters
 * and passes (may be NULL) receive host copies.  Thread-safe. */
int d2ic_serve_poll(d2ic_server* server, int32_t slot, int64_t ticket, int32_t* done, int32_t* n_iters,
                    int32_t* passes);

/* Finish outstanding work, stop the kernel and free the session. */
int d2ic_serve_destroy(d2ic_server* server);

/* ---------------------------------------------------------------------------
 * Implicit-function derivatives of a solve (replace jax.grad / jax.jvp
 * through _cg_solve, solver_global_cg.py:326-469, used for uncertainty
 * propagation in the reference's notebooks).  At a solution u* of
 * G(u) = grad_u J = 0: dL/dtheta = -(dG/dtheta)^T lambda with H lambda = dL/du,
 * and forward tangents du = H^-1 (-(dG/dtheta) dtheta); H = dG/du.
 * Buffers are device float32 unless noted; sizes from d2ic_adj_sizes.
 * ------------------------------------------------------------------------ */
/* lin: 3 floats per pixel slot (r, dg/dx, dg/dy); K: per-tile element
 * Hessians; tile: per-tile, per-vector work floats. */
int d2ic_adj_sizes(const d2ic_cg_plan* plan, int64_t* lin_floats, int64_t* k_floats, int64_t* tile_floats);

/* Linearise the image term at u (n_nodes,2) for the deformed image:
 * gauss_newton = 0 keeps the r * hess(g) term (exact Hessian). */
int d2ic_adj_linearize(const d2ic_cg_plan* plan, const float* def_image, int32_t H, int32_t W, const float* u,
                       int32_t gauss_newton, float* lin, float* K, void* stream);

/* out[k] = H v[k] for nv vectors of (n_nodes,2); work: nv * tile_floats. */
int d2ic_adj_hvp(const d2ic_cg_plan* plan, const float* K, float alpha_reg, const float* v, int32_t nv, float* out,
                 float* work, void* stream);

/* b[k] = -(dG/dg) dimg[k] (which = 0, deformed image) or -(dG/df) dimg[k]
 * (which = 1, reference image) for nv perturbation images (H,W). */
int d2ic_adj_tangent(const d2ic_cg_plan* plan, int32_t H, int32_t W, const float* u, const float* lin, int32_t which,
                     const float* dimg, int32_t nv, float* b, float* work, void* stream);

/* grad_f = -(dG/df)^T lambda and grad_g = -(dG/dg)^T lambda as (H,W) images
 * (either may be NULL); work: H*W 64-bit words when grad_g is wanted.  The
 * g-image sums are exact fixed-point sums (deterministic). */
int d2ic_adj_image_grads(const d2ic_cg_plan* plan, int32_t H, int32_t W, const float* u, const float* lin,
                         const float* lambda, float* grad_f, float* grad_g, unsigned long long* work, void* stream);

/* Inverse 2x2 diagonal blocks of H per node, out (n_nodes,3) = (ixx, ixy, iyy):
 * a block-Jacobi preconditioner for the linear solves with H.  No counterpart
 * in the reference (jax.grad differentiates through the unrolled solve). */
int d2ic_adj_diag_inv(const d2ic_cg_plan* plan, const float* K, float alpha_reg, float* out, void* stream);

/* out[k] = L_W v[k], the spring-energy gradient operator without alpha
 * (dG/dalpha = L_W u); work: nv * tile_floats. */
int d2ic_adj_spring(const d2ic_cg_plan* plan, const float* v, int32_t nv, float* out, float* work, void* stream);

/* ---------------------------------------------------------------------------
 * Local Gauss-Newton nodal refinement (replaces LocalGaussNewtonSolver's
 * _local_sweeps, solver_local_gn.py:205-309, with jacobi_nodal_step_spring,
 * :312-414, and the central-difference gradient images, :195-202).
 * Arrays are the PixelAssets fields in the reference layout (device):
 * pixel_coords (np,2) f32, pixel_nodes (np,4) i32, pixel_shapeN (np,4) f32,
 * node_pixel_index / node_N_weight (nn,max_deg_pix), node_pixel_degree (nn),
 * node_neighbor_index / weight (nn,max_deg_nb), degree (nn), node_reg_weight
 * (nn).  u (nn,2) holds u0 and receives the result after n_sweeps sweeps;
 * hist (n_sweeps,2) = (RMS residual, RMS update) per sweep, or NULL.
 * interp: 0 cubic, 1 linear.
 * ------------------------------------------------------------------------ */
int d2ic_lgn_solve(const float* pixel_coords, const int32_t* pixel_nodes, const float* pixel_shapeN, int64_t np,
                   const int32_t* node_pixel_index, const float* node_N_weight, const int32_t* node_pixel_degree,
                   int64_t max_deg_pix, const int32_t* nb_idx, const int32_t* nb_deg, const float* nb_w,
                   const float* node_reg_weight, int64_t max_deg_nb, int64_t nn, const float* ref_image,
                   const float* def_image, int32_t H, int32_t W, int32_t interp, float* u, int32_t n_sweeps, float lam,
                   float max_step, float alpha_reg, float omega, float* hist, void* stream);

/* The same sweeps on a solver plan's layout (d2ic_cg_plan_create from the same
 * PixelAssets): per sweep one block per element tile reduces the tile's pixels
 * to 5 sums per tile node, and each node gathers its tile sums through the
 * plan's node->tile table instead of walking its node->pixel row.  Same
 * arguments as d2ic_lgn_solve minus the pixel arrays and the node->pixel
 * table (the plan holds them); results agree with d2ic_lgn_solve to FP32
 * summation order. */
int d2ic_lgn_solve_plan(const d2ic_cg_plan* plan, const int32_t* nb_idx, const int32_t* nb_deg, const float* nb_w,
                        const float* node_reg_weight, int64_t max_deg_nb, const float* ref_image,
                        const float* def_image, int32_t H, int32_t W, int32_t interp, float* u, int32_t n_sweeps,
                        float lam, float max_step, float alpha_reg, float omega, float* hist, void* stream);

/* ---------------------------------------------------------------------------
 * Translation ZNCC (replaces _match_centers_jax, solver_translation_zncc.py:118-168)
 * ------------------------------------------------------------------------ */
/* centers (n,2) f32; ref (Hr,Wr) and def (Hd,Wd) f32 row-major.  u_out (n,2)
 * integer-valued f32 shifts, score_out (n,) f32 (NaN when the template or the
 * search patch leaves the image). */
int d2ic_zncc_match(const float* centers, int64_t n_centers, const float* ref, int32_t Hr, int32_t Wr,
                    const float* def, int32_t Hd, int32_t Wd, int32_t win, int32_t search, float* u_out,
                    float* score_out, void* stream);

/* Inverse-distance weighting of centre shifts onto nodes, power 2
 * (_project_centers_to_nodes_idw, dic_init_motion.py:64-91). */
int d2ic_idw(const float* nodes, int64_t n_nodes, const float* centers, const float* disp, int64_t n_centers,
             float* u_out, void* stream);

/* Nodal LSQ Green-Lagrange strain (compute_green_lagrange_strain_nodes_lsq +
 * green_lagrange_to_voigt, strain.py:11-114).  F_out/E_out (n,2,2) may be NULL;
 * voigt (n,3). */
int d2ic_strain(const float* u, const float* nodes, int64_t n_nodes, const int32_t* nb_index,
                const int32_t* nb_degree, int32_t max_degree, float gauge_length, float eps, float* F_out,
                float* E_out, float* voigt, void* stream);

/* Reference-domain residual map (compute_discrepancy_map_ref,
 * discrepancy.py:86-127): bilinear with NaN outside, averaged per image
 * pixel (an invalid pixel makes its image pixel NaN; image pixels no valid
 * pixel reached are NaN).  map_out (H*W) f32; rms written to *rms (host;
 * synchronises): NaN when any pixel's sample is invalid, as the reference's
 * sum of (r * valid)^2 is, and 0 for no pixels.  rms may be NULL. */
int d2ic_discrepancy_map(const float* ref_image, const float* def_image, int32_t H, int32_t W, const float* u,
                         const float* pixel_coords, const int32_t* pixel_nodes, const float* pixel_shapeN,
                         const int32_t* roi_flat, int64_t n_pixels, float* map_out, float* rms, void* stream);

/* Synthetic speckle (workload generator, no reference counterpart): sum of
 * isotropic Gaussian spots at pixel centres, each truncated at 4 sigma.
 * spots: (n,4) f64 [cx, cy, sigma, amp]; out (H*W) f64. */
int d2ic_render_speckle(const double* spots, int64_t n_spots, int32_t H, int32_t W, double* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* D2IC_AMD_H */
