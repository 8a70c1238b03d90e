/*
 * r2splat_hip.h — C ABI of the MI355X (gfx950) CT splatting library.
 *
 * Every entry point takes raw device pointers, element counts and a
 * hipStream_t, and returns an R2S_* status code (message: r2s_last_error()).
 * Nothing here allocates device memory: outputs and workspaces belong to the
 * caller, so a sequence of calls can be captured into a hipGraph. Outputs that
 * are accumulated inside a kernel are cleared by the library itself with a fill
 * kernel (never hipMemsetAsync), so callers may pass uninitialised buffers.
 *
 * `dtype` selects the floating-point type of every floating-point pointer of a
 * call (R2S_F32 = float, R2S_F64 = double). Integer buffers have fixed types as
 * declared. Array layouts are the reference operator layouts (row-major,
 * contiguous); each declaration cites the reference interface it replaces, in
 * the reference repository (MedAPT-group/R2Splat), by file:line.
 *
 * Naming of image/volume dimensions follows the reference: for images
 * `width` = dim1 (x, columns), `height` = dim0 (y, rows); for volumes
 * `dim_x` = dim2 (fastest), `dim_y` = dim1, `dim_z` = dim0 (slowest).
 */
#ifndef R2SPLAT_HIP_H
#define R2SPLAT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* r2s_stream_t; /* identical to hipStream_t */

#define R2S_OK 0
#define R2S_ERR_INVALID 1 /* bad argument (shape, size, flag) */
#define R2S_ERR_HIP 2     /* HIP runtime error (launch, device) */

#define R2S_F32 0
#define R2S_F64 1

/* Library identification and error reporting. Version 5: r2s_value_checks
 * (version 4: the fused voxel entry points take the double-precision frame
 * of r2s_voxel_frame; version 3: they take uncertainties and the variance
 * buffers, the TV and L1 loss backward take lambda as a float; version 2:
 * the fused raster entry points take the record kind and the variance
 * buffers, and `packed` holds tile-list entries followed by the projected
 * records). */
int r2s_abi_version(void);
const char* r2s_last_error(void);

/* Value checks of the explicit operators' inputs: flags[i] = 1 when check i
 * finds a bad element of its buffer, else 0 (flags overwritten). One launch
 * for the whole batch replaces the reference's per-argument device
 * reductions (r2splat/_validation.py:14-96, the isfinite / domain checks of
 * r2splat/explicit.py:118-160); the caller copies the flags to the host once.
 * Comparisons follow IEEE semantics as the reference's do: NaN is not <= 0
 * or < 0 but differs from 0 and 1 (the NONFINITE check of the same buffer
 * reports it first). */
#define R2S_CHECK_NONFINITE 0   /* some element is NaN or infinite */
#define R2S_CHECK_NONPOSITIVE 1 /* some element <= 0 */
#define R2S_CHECK_NEGATIVE 2    /* some element < 0 */
#define R2S_CHECK_ZERO_ROW 3    /* some row of `row` consecutive elements is all zero */
#define R2S_CHECK_K_FOCAL 4     /* some 3x3 intrinsics (row 9) has fx <= 0 or fy <= 0 */
#define R2S_CHECK_K_PINHOLE 5   /* some 3x3 intrinsics (row 9) has nonzero skew or last row != [0, 0, 1] */
#define R2S_CHECK_VIEW_ROW 6    /* some 4x4 viewmat (row 16) has last row != [0, 0, 0, 1] */
#define R2S_MAX_VALUE_CHECKS 16
typedef struct r2s_value_check {
    const void* data; /* device buffer of n elements of dtype */
    uint64_t n;
    int32_t dtype;    /* R2S_F32 / R2S_F64 */
    int32_t kind;     /* R2S_CHECK_* */
    uint32_t row;     /* elements per row (divides n): R2S_CHECK_ZERO_ROW; 9 and 16 for the
                         camera checks */
    uint32_t reserved;
} r2s_value_check;
int r2s_value_checks(const r2s_value_check* checks, uint32_t n_checks, uint32_t* flags,
                     r2s_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Per-primitive geometry                                                    */
/* ------------------------------------------------------------------------ */

/* quats [N,4] (w,x,y,z, unnormalised), scales [N,3] -> covars [N,6] packed
 * (00,01,02,11,12,22), max_scales [N].
 * Replaces parameter_transform_fwd_tensor (r2splat/cuda/include/Ops.h:421-424,
 * bound at r2splat/cuda/csrc/ext.cpp:174). */
int r2s_parameter_transform_fwd(int dtype, const void* quats, const void* scales,
                                void* covars, void* max_scales, uint32_t N,
                                r2s_stream_t stream);

/* VJP to raw quats [N,4] and scales [N,3] from v_covars [N,6].
 * Replaces parameter_transform_bwd_tensor (Ops.h:426-430, ext.cpp:176). */
int r2s_parameter_transform_bwd(int dtype, const void* quats, const void* scales,
                                const void* v_covars, void* v_quats, void* v_scales,
                                uint32_t N, r2s_stream_t stream);

/* means [N,3], covars [N,6], viewmats [C,4,4], Ks [C,3,3] ->
 * means2d [C,N,2], cov2d [C,N,6] (projected 3x3, packed), radii int32 [C,N,2].
 * eps and radius_clip are float for every dtype, as in the reference.
 * Replaces world_to_image_projection_fwd_tensor (Ops.h:398-407, ext.cpp:170). */
int r2s_world_to_image_projection_fwd(int dtype, const void* means, const void* covars,
                                      const void* viewmats, const void* Ks,
                                      uint32_t width, uint32_t height, float eps,
                                      float radius_clip, void* means2d, void* cov2d,
                                      int32_t* radii, uint32_t C, uint32_t N,
                                      r2s_stream_t stream);

/* VJP: v_means2d [C,N,2], v_cov2d [C,N,6] -> v_means [N,3], v_covars [N,6]
 * (summed over cameras deterministically; outputs are overwritten).
 * Replaces world_to_image_projection_bwd_tensor (Ops.h:409-418, ext.cpp:172). */
int r2s_world_to_image_projection_bwd(int dtype, const void* means, const void* covars,
                                      const void* viewmats, const void* Ks,
                                      uint32_t width, uint32_t height,
                                      const void* v_means2d, const void* v_cov2d,
                                      void* v_means, void* v_covars, uint32_t C,
                                      uint32_t N, r2s_stream_t stream);

/* cov2d [T,6] -> conics [T,3] = (Sigma_2x2 + eps I)^-1 packed (00,01,11) and
 * mus [T] = sqrt(2 pi (f + eps - [c,e] P [c,e]^T)); the 2x2 inverse and the
 * Schur complement are evaluated in double precision for every dtype.
 * Replaces conics_mu_fwd_tensor (Ops.h:214-217, ext.cpp:166). */
int r2s_conics_mu_fwd(int dtype, const void* cov2d, double eps, void* conics, void* mus,
                      uint32_t total, r2s_stream_t stream);

/* Replaces conics_mu_bwd_tensor (Ops.h:219-225, ext.cpp:168). */
int r2s_conics_mu_bwd(int dtype, const void* cov2d, double eps, const void* conics,
                      const void* mus, const void* v_conics, const void* v_mus,
                      void* v_cov2d, uint32_t total, r2s_stream_t stream);

/* means [N,3], covars [N,6], max_scales [N] (unused metadata), ranges [3],
 * centers [3] -> means3d_v [N,3] voxel coordinates, conics [N,6] (inverse voxel
 * covariance, packed), radii int32 [N,3]. The 3x3 inversion runs in double.
 * Replaces world_to_volume_mapping_fwd_tensor (Ops.h:432-442, ext.cpp:178). */
int r2s_world_to_volume_mapping_fwd(int dtype, const void* means, const void* covars,
                                    const void* max_scales, const void* ranges,
                                    const void* centers, uint32_t dim_x, uint32_t dim_y,
                                    uint32_t dim_z, double eps3d, double radius_clip,
                                    void* means3d_v, void* conics, int32_t* radii,
                                    uint32_t N, r2s_stream_t stream);

/* Replaces world_to_volume_mapping_bwd_tensor (Ops.h:444-454, ext.cpp:180). */
int r2s_world_to_volume_mapping_bwd(int dtype, const void* means, const void* covars,
                                    const void* ranges, const void* centers,
                                    uint32_t dim_x, uint32_t dim_y, uint32_t dim_z,
                                    double eps3d, const void* v_means3d_v,
                                    const void* v_conics, void* v_means, void* v_covars,
                                    uint32_t N, r2s_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Tile binning (2D and 3D)                                                  */
/* ------------------------------------------------------------------------ */

/* Workspace sizes (bytes) for the device scan over `n` counts and for the
 * stable radix sort of `n_isects` (key, value) pairs over `end_bit` key bits. */
size_t r2s_scan_workspace_bytes(uint64_t n);
size_t r2s_sort_workspace_bytes(uint64_t n_isects, uint32_t end_bit);

/* Pass 1 of isect_tiles: per (camera, primitive) tile count (int32 [C,N]) and
 * its exclusive prefix sum `cum` (uint32 [C*N+1]; cum[C*N] = n_isects, also
 * written to *d_total, a device uint32). radii are [C,N,2] if radii_is_rect
 * else [C,N]. Replaces the count/cumsum half of isect_tiles_tensor
 * (r2splat/cuda/csrc/IsectTiles.cpp:31-107, Ops.h:49-55). */
int r2s_isect_tiles_count(int dtype, const void* means2d, const int32_t* radii,
                          int radii_is_rect, uint32_t C, uint32_t N, uint32_t tile_size,
                          uint32_t n_tiles_x, uint32_t n_tiles_y, int32_t* tiles_per_gauss,
                          uint32_t* cum, void* scan_ws, uint32_t* d_total,
                          r2s_stream_t stream);

/* Pass 2: write (key = cam*n_tiles + ty*n_tiles_x + tx, value = cam*N + g)
 * records and sort them stably by key (ties keep ascending value). Outputs
 * isect_ids int64 [n_isects] and flatten_ids int32 [n_isects].
 * Replaces the generate + CUB sort half of isect_tiles_tensor
 * (IsectTiles.cpp:109-170, IsectTilesKernel.cu:137-217,398-445). */
int r2s_isect_tiles_emit(int dtype, const void* means2d, const int32_t* radii,
                         int radii_is_rect, uint32_t C, uint32_t N, uint32_t tile_size,
                         uint32_t n_tiles_x, uint32_t n_tiles_y, const uint32_t* cum,
                         uint32_t n_isects, int64_t* isect_ids, int32_t* flatten_ids,
                         void* sort_ws, size_t sort_ws_bytes, r2s_stream_t stream);

/* 3D analogues over a (z,y,x) tile grid; key = tz*ny*nx + ty*nx + tx, value = g.
 * Replaces isect_tiles_3d_tensor (r2splat/cuda/csrc/IsectTiles3D.cpp:32-206,
 * Ops.h:61-68). */
int r2s_isect_tiles_3d_count(int dtype, const void* means3d, const int32_t* radii,
                             uint32_t N, uint32_t tile_size, uint32_t n_tiles_x,
                             uint32_t n_tiles_y, uint32_t n_tiles_z,
                             int32_t* tiles_per_gauss, uint32_t* cum, void* scan_ws,
                             uint32_t* d_total, r2s_stream_t stream);
int r2s_isect_tiles_3d_emit(int dtype, const void* means3d, const int32_t* radii,
                            uint32_t N, uint32_t tile_size, uint32_t n_tiles_x,
                            uint32_t n_tiles_y, uint32_t n_tiles_z, const uint32_t* cum,
                            uint32_t n_isects, int64_t* isect_ids, int32_t* flatten_ids,
                            void* sort_ws, size_t sort_ws_bytes, r2s_stream_t stream);

/* offsets int32 [n_keys]: offsets[k] = first index i with isect_ids[i] >= k
 * (n_isects if none). n_keys = C*n_tiles (2D) or n_tiles (3D).
 * Replaces isect_offset_encode_tensor / isect_offset_encode_3d_tensor
 * (IsectTiles.cpp:176-215, IsectTiles3D.cpp:~150-206, Ops.h:57-74). */
int r2s_isect_offset_encode(const int64_t* isect_ids, uint32_t n_isects, uint32_t n_keys,
                            int32_t* offsets, r2s_stream_t stream);

/* ------------------------------------------------------------------------ */
/* 2D rasterization (CT line integrals, additive, no alpha compositing)      */
/* ------------------------------------------------------------------------ */
/* Common arguments: means2d [C,N,2], conics [C,N,3], opacities [C,N],
 * mus [C,N], betas [C,N] (projected exponent, Beta ops only), uncertainties
 * [N], tile_offsets int32 [C,n_tiles_y,n_tiles_x], flatten_ids int32
 * [n_isects], isect_ids int64 [n_isects]; images [C,height,width]. */

/* Replaces rasterize_to_pixels_fwd_tensor (Ops.h:227-236, ext.cpp:96). */
int r2s_rasterize_to_pixels_fwd(int dtype, const void* means2d, const void* conics,
                                const void* opacities, const void* mus,
                                const int32_t* tile_offsets, const int32_t* flatten_ids,
                                uint32_t n_isects, uint32_t C, uint32_t N, uint32_t width,
                                uint32_t height, uint32_t tile_size, uint32_t n_tiles_x,
                                uint32_t n_tiles_y, void* render, r2s_stream_t stream);

/* Replaces rasterize_to_pixels_mean_var_fwd_tensor (Ops.h:238-248, ext.cpp:98). */
int r2s_rasterize_to_pixels_mean_var_fwd(int dtype, const void* means2d, const void* conics,
                                         const void* opacities, const void* mus,
                                         const void* uncertainties,
                                         const int32_t* tile_offsets,
                                         const int32_t* flatten_ids, uint32_t n_isects,
                                         uint32_t C, uint32_t N, uint32_t width,
                                         uint32_t height, uint32_t tile_size,
                                         uint32_t n_tiles_x, uint32_t n_tiles_y,
                                         void* render, void* render_var,
                                         r2s_stream_t stream);

/* Replaces beta_rasterize_to_pixels_fwd_tensor (Ops.h:318-328, ext.cpp:133). */
int r2s_beta_rasterize_to_pixels_fwd(int dtype, const void* means2d, const void* conics,
                                     const void* opacities, const void* mus,
                                     const void* betas, const int32_t* tile_offsets,
                                     const int32_t* flatten_ids, uint32_t n_isects,
                                     uint32_t C, uint32_t N, uint32_t width,
                                     uint32_t height, uint32_t tile_size,
                                     uint32_t n_tiles_x, uint32_t n_tiles_y, void* render,
                                     r2s_stream_t stream);

/* Replaces beta_rasterize_to_pixels_mean_var_fwd_tensor (Ops.h:330-341, ext.cpp:135). */
int r2s_beta_rasterize_to_pixels_mean_var_fwd(
    int dtype, const void* means2d, const void* conics, const void* opacities,
    const void* mus, const void* betas, const void* uncertainties,
    const int32_t* tile_offsets, const int32_t* flatten_ids, uint32_t n_isects, uint32_t C,
    uint32_t N, uint32_t width, uint32_t height, uint32_t tile_size, uint32_t n_tiles_x,
    uint32_t n_tiles_y, void* render, void* render_var, r2s_stream_t stream);

/* Gradient of sum(v_render_var * var) w.r.t. uncertainties [N] (overwritten).
 * betas == NULL selects the Gaussian profile.
 * Replaces rasterize_to_pixels_uncertainty_bwd_tensor and its Beta variant
 * (Ops.h:250-260,343-354; ext.cpp:100,137). */
int r2s_rasterize_to_pixels_uncertainty_bwd(
    int dtype, const void* means2d, const void* conics, const void* opacities,
    const void* mus, const void* betas, const int32_t* tile_offsets,
    const int32_t* flatten_ids, uint32_t n_isects, uint32_t C, uint32_t N, uint32_t width,
    uint32_t height, uint32_t tile_size, uint32_t n_tiles_x, uint32_t n_tiles_y,
    const void* v_render_var, void* v_uncertainties, r2s_stream_t stream);

/* Tile-parallel ("pixel" mode) VJP. betas/v_betas NULL selects the Gaussian
 * profile; v_means2d_abs may be NULL. All outputs are overwritten.
 * Replaces rasterize_to_pixels_bwd_tensor / beta_rasterize_to_pixels_bwd_tensor
 * (Ops.h:262-274,356-369; ext.cpp:102,139). */
int r2s_rasterize_to_pixels_bwd(
    int dtype, const void* means2d, const void* conics, const void* opacities,
    const void* mus, const void* betas, const int32_t* tile_offsets,
    const int32_t* flatten_ids, uint32_t n_isects, uint32_t C, uint32_t N, uint32_t width,
    uint32_t height, uint32_t tile_size, uint32_t n_tiles_x, uint32_t n_tiles_y,
    const void* v_render, void* v_means2d_abs, void* v_means2d, void* v_conics,
    void* v_opacities, void* v_mus, void* v_betas, r2s_stream_t stream);

/* Intersection-parallel ("primitive" mode) VJP: one wave64 per intersection,
 * lanes stride the tile's pixels, wave reduction, then one atomic per output
 * field per wave. v_render_var == NULL: mean-only VJP. With v_render_var and
 * uncertainties != NULL the variance path also feeds the mean adjoint
 * (v += 2 q alpha v_var); uncertainties == NULL is the detached form. The
 * uncertainty gradient is written to v_uncertainties when it is non-NULL.
 * Replaces rasterize_to_pixels_bwd_per_gaussian_tensor,
 * rasterize_to_pixels_mean_var_bwd_per_gaussian_tensor and the Beta variants
 * (Ops.h:276-316,371-395; ext.cpp:104-128,141-162). */
int r2s_rasterize_to_pixels_bwd_per_gaussian(
    int dtype, const void* means2d, const void* conics, const void* opacities,
    const void* mus, const void* betas, const int64_t* isect_ids,
    const int32_t* flatten_ids, uint32_t n_isects, uint32_t C, uint32_t N, uint32_t width,
    uint32_t height, uint32_t tile_size, uint32_t n_tiles_x, uint32_t n_tiles_y,
    const void* v_render, const void* v_render_var, const void* uncertainties,
    void* v_means2d_abs, void* v_means2d, void* v_conics, void* v_opacities, void* v_mus,
    void* v_betas, void* v_uncertainties, r2s_stream_t stream);

/* ------------------------------------------------------------------------ */
/* 3D voxelization                                                           */
/* ------------------------------------------------------------------------ */
/* means3d [N,3] voxel coordinates, conics [N,6], opacities [N], betas [N]
 * (physical exponent; Beta ops only), uncertainties [N], tile_offsets int32
 * [n_tiles_z,n_tiles_y,n_tiles_x]; volumes [dim_z,dim_y,dim_x].
 * render_var == NULL -> mean only; betas == NULL -> Gaussian profile.
 * Replaces voxelize_to_voxels_fwd_tensor, voxelize_to_voxels_mean_var_fwd_tensor
 * and the Beta variants (Ops.h:76-99,138-161; ext.cpp:26-29,66-69). */
int r2s_voxelize_to_voxels_fwd(int dtype, const void* means3d, const void* conics,
                               const void* opacities, const void* betas,
                               const void* uncertainties, const int32_t* tile_offsets,
                               const int32_t* flatten_ids, uint32_t n_isects, uint32_t N,
                               uint32_t dim_x, uint32_t dim_y, uint32_t dim_z,
                               uint32_t tile_size, uint32_t n_tiles_x, uint32_t n_tiles_y,
                               uint32_t n_tiles_z, void* render, void* render_var,
                               r2s_stream_t stream);

/* Tile-parallel VJP (mean only). Outputs overwritten; v_betas used iff betas.
 * Replaces voxelize_to_voxels_bwd_tensor / beta_voxelize_to_voxels_bwd_tensor
 * (Ops.h:101-112,163-175; ext.cpp:30,70). */
int r2s_voxelize_to_voxels_bwd(int dtype, const void* means3d, const void* conics,
                               const void* opacities, const void* betas,
                               const int32_t* tile_offsets, const int32_t* flatten_ids,
                               uint32_t n_isects, uint32_t N, uint32_t dim_x,
                               uint32_t dim_y, uint32_t dim_z, uint32_t tile_size,
                               uint32_t n_tiles_x, uint32_t n_tiles_y, uint32_t n_tiles_z,
                               const void* v_render, void* v_means3d, void* v_conics,
                               void* v_opacities, void* v_betas, r2s_stream_t stream);

/* Intersection-parallel VJP with optional variance path (see the 2D form).
 * Replaces voxelize_to_voxels_bwd_per_gaussian_tensor,
 * voxelize_to_voxels_mean_var_bwd_per_gaussian_tensor and the Beta variants
 * (Ops.h:114-136,177-212; ext.cpp:32-54,72-94). */
int r2s_voxelize_to_voxels_bwd_per_gaussian(
    int dtype, const void* means3d, const void* conics, const void* opacities,
    const void* betas, const int64_t* isect_ids, const int32_t* flatten_ids,
    uint32_t n_isects, uint32_t N, uint32_t dim_x, uint32_t dim_y, uint32_t dim_z,
    uint32_t tile_size, uint32_t n_tiles_x, uint32_t n_tiles_y, uint32_t n_tiles_z,
    const void* v_render, const void* v_render_var, const void* uncertainties,
    void* v_means3d, void* v_conics, void* v_opacities, void* v_betas,
    void* v_uncertainties, r2s_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Fused float32 Gaussian path (8x8, 16x16 or 32x32 tiles)                  */
/* ------------------------------------------------------------------------ */
/* One call runs the whole rasterize_fwd sequence of the reference
 * (r2splat/explicit.py:160-183: parameter transform, projection, conics/mu,
 * isect_tiles + sort, offsets, rasterize_to_pixels_fwd) from physical
 * parameters, and one call the whole rasterize_bwd sequence
 * (explicit.py:208-241), with identical semantics. Intersection-dependent
 * sizes stay on the device: the sort covers min(counters[0], capacity)
 * records, so the pair can be captured into a hipGraph.
 *
 * tile_size is 8, 16 or 32 (the reference's tile_size argument); waves is
 * the number of wave64s sharing one tile in the raster kernels (4, 8 or 16;
 * 0 = 8). The backward's results do not depend on waves; the forward's
 * per-pixel summation order does (rounding only).
 *
 * Buffers: scratch (r2s_fused_scratch_bytes, dead after the forward);
 * packed = r2s_fused_packed_bytes(C, N, capacity, kind) bytes: `capacity`
 * list entries {record index, emission index} (uint2) in camera-major tile
 * lists (the reference's list layout; within a list the entries run by
 * decreasing height class, a deterministic order that only changes the
 * per-pixel summation order), then the C*N projected records of
 * r2s_fused_record_bytes(kind) each (offset: capacity * 8 rounded up to 256
 * bytes; Beta records are 96 bytes: the 48 the raster kernels gather through
 * the entries, then the record's mean, conic and cut-off bound in double, which
 * the backward reads to decide pairs within float32 rounding of the cut-off as
 * float64 does); tile_offsets int32
 * [C, ceil(H/tile), ceil(W/tile)] (the reference's layout); tile_order
 * uint32 [same count]
 * (tiles by decreasing intersection count, the order the raster kernels
 * walk them in; may be null, then natural order), record_offsets uint32
 * [C*N + 1 + N] and counters uint32 [2] persist from the forward to the
 * backward. The forward processes primitives in the order of a 15-bit
 * Morton code of their means (its top 8 bits for up to 2^17 primitives;
 * ties in input order): record_offsets[C*N + 1
 * + p] is the input index of position p, and record_offsets[c*N + p] the
 * first intersection of (camera c, position p) in emission order. The
 * gradients come back in the input order. counters[0]
 * = intersections, counters[1] = 1 when capacity was exceeded (that
 * forward's outputs are then incomplete and the call must be repeated with
 * a larger capacity). isect_grads: float [capacity * 8], one VJP partial per
 * intersection (8 floats: v_mx, v_my, v_a, v_b, v_c, v_w, v_b2d, v_q),
 * summed per primitive in a fixed order (the backward is deterministic and
 * uses no atomics).
 *
 * Primitive kind: betas == NULL selects the Gaussian profile (explicit.py
 * with beta=None). Otherwise betas [N] is the per-primitive beta (projected
 * exponent beta + 1/2), beta_ratio [N] = B(1/2, beta + 1) / sqrt(2 pi)
 * (the line-integral factor, mus = mu * ratio; explicit.py:166-170) and
 * beta_log_grad [N] = d log ratio / d beta (explicit.py:226-230), both as
 * r2splat/_beta.py computes them. uncertainties [N] (may be NULL) adds the
 * variance output var = sum q alpha^2 (explicit.py:176-183 with uncertainty;
 * RasterizeToPixelsFwd.cu mean_var kernels) and its VJP: the pixel adjoint
 * becomes v_image + 2 q alpha v_image_var unless `detach` (the reference's
 * detach_variance), and v_uncertainties = sum v_image_var alpha^2.
 * `kind` (bit 0 Beta, bit 1 uncertainty) selects the record layout in
 * r2s_fused_scratch_bytes / r2s_fused_record_bytes / r2s_fused_render /
 * r2s_fused_raster_vjp and must match the forward. */
size_t r2s_fused_scratch_bytes(uint32_t C, uint32_t N, uint32_t width, uint32_t height,
                               uint32_t tile_size, uint32_t capacity, int kind);
size_t r2s_fused_record_bytes(int kind);
size_t r2s_fused_packed_bytes(uint32_t C, uint32_t N, uint32_t capacity, int kind);

int r2s_fused_forward(const float* means, const float* quats, const float* scales,
                      const float* densities, const float* betas, const float* beta_ratio,
                      const float* uncertainties, const float* viewmats, const float* Ks, uint32_t C,
                      uint32_t N, uint32_t width, uint32_t height, uint32_t tile_size, float eps,
                      float radius_clip, uint32_t capacity, void* scratch, size_t scratch_bytes,
                      void* packed, int32_t* tile_offsets, uint32_t* tile_order, uint32_t* record_offsets,
                      uint32_t* counters, float* image, float* image_var, r2s_stream_t stream);

/* Outputs: v_means [N,3], v_quats [N,4], v_scales [N,3], v_densities [N],
 * v_betas [N] (Beta only), v_uncertainties [N] (with uncertainties; all
 * overwritten; camera sums in a fixed order). v_image_var may be NULL. */
int r2s_fused_backward(const float* means, const float* quats, const float* scales,
                       const float* densities, const float* betas, const float* beta_ratio,
                       const float* beta_log_grad, const float* uncertainties, const float* viewmats,
                       const float* Ks, uint32_t C, uint32_t N, uint32_t width, uint32_t height,
                       uint32_t tile_size, float eps, uint32_t capacity, const void* packed,
                       const int32_t* tile_offsets, const uint32_t* tile_order, const uint32_t* record_offsets,
                       const uint32_t* counters, float* isect_grads, void* workspace, size_t workspace_bytes,
                       const float* v_image, const float* v_image_var, int detach, float* v_means,
                       float* v_quats, float* v_scales, float* v_densities, float* v_betas,
                       float* v_uncertainties, r2s_stream_t stream);

/* The same two calls split into stages: forward = prepare (primitive order,
 * projection, binning into tile lists, offsets) + render (rasterization);
 * backward = raster_vjp (rasterization VJP, one partial per intersection) +
 * param_vjp (per-primitive sums, conic/mu, projection and parameter-transform
 * VJPs, summed over cameras). */
int r2s_fused_prepare(const float* means, const float* quats, const float* scales,
                      const float* densities, const float* betas, const float* beta_ratio,
                      const float* uncertainties, const float* viewmats, const float* Ks, uint32_t C,
                      uint32_t N, uint32_t width, uint32_t height, uint32_t tile_size, float eps,
                      float radius_clip, uint32_t capacity, void* scratch, size_t scratch_bytes,
                      void* packed, int32_t* tile_offsets, uint32_t* tile_order, uint32_t* record_offsets,
                      uint32_t* counters, r2s_stream_t stream);
/* image_var: [C, H, W] (required iff kind has the uncertainty bit). */
int r2s_fused_render(const void* packed, const int32_t* tile_offsets, const uint32_t* tile_order,
                     const uint32_t* counters, uint32_t capacity, uint32_t C, uint32_t width, uint32_t height,
                     uint32_t tile_size, uint32_t waves, int kind, float* image, float* image_var,
                     r2s_stream_t stream);
int r2s_fused_raster_vjp(const void* packed, const int32_t* tile_offsets, const uint32_t* tile_order,
                         const uint32_t* counters, uint32_t capacity, uint32_t C, uint32_t width, uint32_t height,
                         uint32_t tile_size, uint32_t waves, int kind, const float* v_image,
                         const float* v_image_var, int detach, float* isect_grads, r2s_stream_t stream);
/* Diagnostic, no counterpart in the reference: the (list entry, pixel) pairs
 * that r2s_fused_render and r2s_fused_raster_vjp keep at the cut-off, decided
 * by the same operations as those kernels. keep_fwd and keep_bwd are uint32
 * [counters[0]][tile_size^2 / 32]: bit b of word w of entry i is pixel
 * p = 32 w + b = row * tile_size + column of entry i's tile (pixels past the
 * image edge included). For tests that evaluate exactly the kept pairs in
 * float64, separating arithmetic error from pairs at the threshold. */
int r2s_fused_keep_masks(const void* packed, const int32_t* tile_offsets, const uint32_t* counters,
                         uint32_t capacity, uint32_t C, uint32_t width, uint32_t height, uint32_t tile_size,
                         int kind, uint32_t* keep_fwd, uint32_t* keep_bwd, r2s_stream_t stream);

/* The parameter VJP sums each primitive's cameras inside one kernel in a
 * fixed order and needs no workspace: r2s_fused_param_workspace_bytes returns
 * 0 and the workspace arguments may be NULL / 0 (kept for callers that pass
 * a buffer). */
size_t r2s_fused_param_workspace_bytes(uint32_t C, uint32_t N);

/* Packed launches, no counterpart in the reference (which fits one scene per
 * call): the render or the raster VJP of several independent jobs (prepared
 * fused forwards of any C, N and image size, with one tile edge, wave count
 * and record kind) as one launch over the concatenation of their tiles, in
 * job order (job j's C * ceil(H/tile) * ceil(W/tile) tiles from first_tile
 * on: first_tile is the exclusive prefix of those counts). Each job's results
 * are bitwise those of its own r2s_fused_render / r2s_fused_raster_vjp.
 * Each of the 8 XCDs takes the band of every job's tile order that the job's
 * own launch would give it, so jobs of very different per-tile cost spread
 * evenly. jobs_host and jobs_device hold the same table (n_jobs entries, 1 to 1024):
 * the host copy sizes and checks the launch, the kernels read the device
 * copy, which must stay valid while the launch (or a graph that captured
 * it) runs. Render reads packed, tile_offsets, tile_order, counters,
 * capacity, C, width, height and writes image (and image_var with the
 * uncertainty kind); the VJP also reads v_image (v_image_var, detach) and
 * writes isect_grads. Unused fields may be NULL / 0. */
typedef struct r2s_raster_job {
    const void* packed;
    const int32_t* tile_offsets;
    const uint32_t* tile_order;
    const uint32_t* counters;
    float* image;
    float* image_var;
    const float* v_image;
    const float* v_image_var;
    float* isect_grads;
    uint32_t capacity, C, width, height;
    int32_t detach;
    uint32_t first_tile;
    uint32_t reserved[2];
} r2s_raster_job;
int r2s_fused_render_packed(const r2s_raster_job* jobs_host, const r2s_raster_job* jobs_device, uint32_t n_jobs,
                            uint32_t tile_size, uint32_t waves, int kind, r2s_stream_t stream);
int r2s_fused_raster_vjp_packed(const r2s_raster_job* jobs_host, const r2s_raster_job* jobs_device,
                                uint32_t n_jobs, uint32_t tile_size, uint32_t waves, int kind, r2s_stream_t stream);
int r2s_fused_param_vjp(const float* means, const float* quats, const float* scales,
                        const float* densities, const float* betas, const float* beta_ratio,
                        const float* beta_log_grad, const float* viewmats, const float* Ks, uint32_t C,
                        uint32_t N, uint32_t width, uint32_t height, float eps,
                        const uint32_t* record_offsets, const uint32_t* counters, uint32_t capacity,
                        const float* isect_grads, void* workspace, size_t workspace_bytes,
                        float* v_means, float* v_quats, float* v_scales, float* v_densities,
                        float* v_betas, float* v_uncertainties, r2s_stream_t stream);

/* Double-precision voxel frame of the fused voxel kernels: frame [N,9]
 * float64 = per primitive the voxel-frame mean (x, y, z) and the inverse of
 * the voxel-frame covariance (xx, xy, xz, yy, yz, zz), computed in double
 * from float32 means [N,3], quats [N,4], scales [N,3] and float64 ranges [3]
 * and centers [3] (the volume frame as given) by the chain
 * r2s_parameter_transform_fwd ->
 * r2s_world_to_volume_mapping_fwd (ParameterTransformFwd.cu,
 * WorldToVolumeMappingFwd.cu; eps added to the voxel-frame diagonal; zeros
 * for a singular covariance). No counterpart in the reference: the fused
 * float32 kernels take their means and conics from it in place of the
 * float32 mapping's, because the Beta profile (1 - d^2)^beta with beta < 1 is
 * steep at its support edge (a change of 1e-7 in d^2 moves a term by up to
 * ~3e-4 there). */
int r2s_voxel_frame(const float* means, const float* quats, const float* scales, const double* ranges,
                    const double* centers, uint32_t dim_x, uint32_t dim_y, uint32_t dim_z, double eps,
                    uint32_t N, double* frame, r2s_stream_t stream);

/* Fused float32 voxel forward (8^3 tiles) over the 3D binning: render
 * [dim_z, dim_y, dim_x] and, with uncertainties [N] and render_var, the
 * variance sum q alpha^2 (both overwritten). Replaces voxelize_to_voxels_fwd
 * and voxelize_to_voxels_mean_var_fwd (VoxelizeToVoxelsFwd.cu:181-219, :297
 * and the Beta kernels) for float32; the summation order within a voxel
 * differs from the operator (rounding only). means3d [N,3] are the float32
 * voxel-frame means the binning used (r2s_world_to_volume_mapping_fwd); frame
 * (required) is r2s_voxel_frame's. Gaussian terms take the mean from the frame
 * as a float pair hi + lo (voxel offsets (v - hi) - lo) and the conic rounded
 * to float32; Beta terms form d^2 and 1 - d^2 in double from the frame and
 * decide the cut-off (0 <= d^2, |alpha| >= 1e-5) in double. */
int r2s_voxelize_fused_fwd(const float* means3d, const double* frame, const float* conics,
                           const float* opacities,
                           const float* betas, const float* uncertainties, const int32_t* tile_offsets,
                           const int32_t* flatten_ids, uint32_t n_isects, uint32_t N, uint32_t dim_x,
                           uint32_t dim_y, uint32_t dim_z, uint32_t tile_size, uint32_t n_tiles_x,
                           uint32_t n_tiles_y, uint32_t n_tiles_z, float* render, float* render_var,
                           r2s_stream_t stream);

/* Fused float32 voxel VJP (8^3 tiles): the whole per-primitive voxelize
 * backward (VoxelizeToVoxelsBwd.cu per-Gaussian kernels, mean and mean_var, as
 * called by explicit.py:327-350) without atomics. radii int32 [N,3] and
 * tile_cum uint32 [N+1] (exclusive scan of the per-primitive tile counts) are
 * the binning's own (r2s_isect_tiles_3d_count); isect_ids / flatten_ids its
 * sorted output; means3d and frame as for r2s_voxelize_fused_fwd (the same
 * terms and cut-off decisions).
 * With v_render_var the voxel adjoint becomes v_render + 2 q
 * alpha v_render_var (v_render alone when `detach`, the reference's
 * detach_variance) and v_uncertainties [N] = sum v_render_var alpha^2. One
 * partial per intersection is written at its emission index (partials:
 * r2s_voxel_partials_bytes(n_isects) bytes) and summed per primitive in a
 * fixed order, so the result is bitwise reproducible. Outputs overwritten:
 * v_means3d [N,3], v_conics [N,6], v_opacities [N], v_betas [N] (iff betas),
 * v_uncertainties [N] (iff v_render_var). */
size_t r2s_voxel_partials_bytes(uint32_t n_isects);
int r2s_voxelize_fused_vjp(const float* means3d, const double* frame, const float* conics,
                           const float* opacities,
                           const float* betas, const float* uncertainties, const int32_t* radii,
                           const uint32_t* tile_cum, const int64_t* isect_ids, const int32_t* flatten_ids,
                           uint32_t n_isects, uint32_t N, uint32_t dim_x, uint32_t dim_y, uint32_t dim_z,
                           uint32_t tile_size, uint32_t n_tiles_x, uint32_t n_tiles_y, uint32_t n_tiles_z,
                           const float* v_render, const float* v_render_var, int detach, float* partials,
                           size_t partials_bytes, float* v_means3d, float* v_conics, float* v_opacities,
                           float* v_betas, float* v_uncertainties, r2s_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Training-step kernels (no counterpart in the reference, which leaves image */
/* losses beyond L1 and the optimizer to the caller; used by graph-captured   */
/* fits that must not depend on host-side scalars)                            */
/* ------------------------------------------------------------------------ */
/* Number of per-block loss partials r2s_mse_adjoint writes for n pixels. */
size_t r2s_mse_partials(uint64_t n);

/* v_image = 2 (image - target) / n; when partials (r2s_mse_partials(n)
 * floats) and loss are given, loss[0] = mean((image - target)^2) summed in a
 * fixed order. float32. */
int r2s_mse_adjoint(const float* image, const float* target, uint64_t n, float* v_image,
                    float* partials, float* loss, r2s_stream_t stream);

/* r2s_mse_adjoint with partials and loss, for callers that cannot allocate
 * (code under hipGraph capture, tight serving loops, harnesses that forbid
 * device allocation): no counterpart in the reference. The loss partials live
 * in a static device array of the library (256 KiB per GPU, held until the
 * process ends) instead of a caller workspace; n is limited to 134,217,728
 * pixels (larger n returns R2S_ERR_INVALID). Results are bitwise those of
 * r2s_mse_adjoint. Every call on one GPU shares that GPU's array, so calls
 * that overlap on one GPU (concurrent streams or threads) race. */
int r2s_mse_adjoint_noalloc(const float* image, const float* target, uint64_t n, float* v_image,
                            float* loss, r2s_stream_t stream);

/* Adam over a flat float32 buffer (torch.optim.Adam semantics, no weight
 * decay). state uint32 [2]: applied steps (bias correction uses state[0]+1)
 * and skipped steps. When skip_flag is non-null and *skip_flag != 0 (the
 * fused forward's overflow flag, counters + 1) the update is skipped and
 * state[1] is incremented instead. */
int r2s_adam_step(float* params, const float* grads, float* exp_avg, float* exp_avg_sq, uint64_t n,
                  float lr, float beta1, float beta2, float eps, uint32_t* state,
                  const uint32_t* skip_flag, r2s_stream_t stream);

/* x[i] = max(x[i], lo): keeps optimized parameters inside the explicit API's
 * domain (positive scales, nonnegative densities) after an update. */
int r2s_clamp_min(float* x, uint64_t n, float lo, r2s_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Losses                                                                    */
/* ------------------------------------------------------------------------ */

/* Anisotropic TV over volume [B,D,H,W]: tv_map (per-voxel sum of |forward
 * differences|) and tv_sum[0] = sum of tv_map (overwritten).
 * Replaces the kernel half of volume_loss_fwd (r2splat/cuda/csrc/Loss.cpp:115-153). */
int r2s_volume_tv_fwd(int dtype, const void* volume, uint32_t B, uint32_t D, uint32_t H,
                      uint32_t W, void* tv_map, void* tv_sum, r2s_stream_t stream);

/* grad_volume = (sign-sum of adjacent differences * (grad_output[0] /
 * edge_count)) * lambda_tv, lambda_tv a float as in the reference, in the
 * volume's precision. Replaces volume_loss_bwd (Ops.h:41-44, Loss.cpp:156-188,
 * VolumeLossBwd.cu:46-115). */
int r2s_volume_tv_bwd(int dtype, const void* volume, const void* grad_output, uint32_t B,
                      uint32_t D, uint32_t H, uint32_t W, float lambda_tv,
                      void* grad_volume, r2s_stream_t stream);

/* Isotropic TV over volume [B,D,H,W]: tv_map[v] = sqrt(dz^2 + dy^2 + dx^2 + eps)
 * with forward differences (zero past the last index of an axis) and
 * tv_sum[0] = sum of tv_map (overwritten); the loss is lambda * tv_sum / (B D H W).
 * No counterpart in the reference (whose TV is anisotropic, Loss.cpp:115-153);
 * the isotropic option of the CT objective. eps > 0. */
int r2s_volume_tv_iso_fwd(int dtype, const void* volume, uint32_t B, uint32_t D, uint32_t H,
                          uint32_t W, double eps, void* tv_map, void* tv_sum, r2s_stream_t stream);

/* grad_volume = lambda * grad_output[0] / (B D H W) * d(sum of tv_map)/d volume. */
int r2s_volume_tv_iso_bwd(int dtype, const void* volume, const void* grad_output, uint32_t B,
                          uint32_t D, uint32_t H, uint32_t W, double eps, double lambda_tv,
                          void* grad_volume, r2s_stream_t stream);

/* Mean-L1 image loss map |p - t| and its sum (overwritten); gradient
 * g[0] * lambda * sign(p - t) / n (lambda a float, as in the reference).
 * Replaces image_loss_l1_fwd/bwd
 * (Ops.h:14-34, r2splat/cuda/csrc/ImageLossFwd.cu:20-31, ImageLossBwd.cu:14-30). */
int r2s_image_l1_fwd(int dtype, const void* prediction, const void* target, uint64_t n,
                     void* l1_map, void* l1_sum, r2s_stream_t stream);
int r2s_image_l1_bwd(int dtype, const void* prediction, const void* target,
                     const void* grad_output, uint64_t n, float lambda_l1,
                     void* grad_prediction, r2s_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* R2SPLAT_HIP_H */
