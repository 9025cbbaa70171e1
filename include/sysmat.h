/* sysmat.h -- C ABI of libsysmat_hip.so, the MI355X (gfx950) system-matrix
 * generator.  Plain pointers and sizes only; every entry point returns a
 * status (>= 0 success, < 0 an SM_E* code; sm_last_error() gives the text).
 *
 * Reference interfaces replaced (Tridev-github/gpu-spect-pet-sysmat):
 *   PESysMatGen(...)          PEGen_RayTracing_CircularHole/PESysMatGen.h:4-9,
 *                             PESysMatGen.cu:766-833
 *   scatter(...) combined     ScatterGen_RayTracing_CircularHole/scatter.h:4-5,
 *                             scatter.cu:2013-2378
 *   scatter(...) collimator   ScatterGen_Collimator/scatter.h:4-5, scatter.cu:1511-1747
 *   scatter(...) crystal      ScatterGen_Crystal/scatter.h:4, scatter.cu:1528-1738
 * The C++-linkage originals are also exported, declared in sysmat_ref.h, so
 * the reference driver sources link against this library unchanged.
 *
 * Parameter arrays are the reference's raw float32 Params_*.dat contents
 * (layout: PEGen_RayTracing_CircularHole/ReadMe.txt).  Counts are read from
 * the headers (Col[0], Col[(m+1)*10], Det[0]); buffers must hold at least
 * 100 + 9*holes (Col), 1 + 12*N (Det) and 32 (Img) floats.
 */
#ifndef SYSMAT_H_
#define SYSMAT_H_

#ifdef __cplusplus
extern "C" {
#endif

#define SM_OK 0
#define SM_E_ARG (-1)       /* invalid argument / parameter array */
#define SM_E_DEVICE (-2)    /* HIP device missing or call failed */
#define SM_E_NOMEM (-3)     /* device allocation failed */
#define SM_E_OVERFLOW (-4)  /* cull list overflow even with 1x1x1 tiles */
#define SM_E_IO (-5)        /* file read / write failed */

/* Library version (major*10000 + minor*100 + patch). */
int sm_version(void);
/* Text of the last error on this thread ("" if none). */
const char* sm_last_error(void);
/* Number of visible HIP devices (0 when none). */
int sm_device_count(void);

/* ---- photoelectric system matrix -------------------------------------- */

typedef struct SmPeOptions {
    int voxel_sub;             /* voxel sub-points per axis, averaged (1 = voxel centre, reference) */
    int sub_x, sub_y, sub_z;   /* crystal sub-cells, summed (8, 16, 8 = reference, PESysMatGen.cu:435) */
    int det_begin, det_count;  /* detector rows to compute (det_count <= 0: all from det_begin) */
    int z_begin, z_count;      /* voxel z range of the slab (z_count <= 0: to the end) */
    int rotation;              /* rotation index (reference: Img[20]) */
    int tile_x, tile_y, tile_z;/* voxel tile per workgroup (0 = automatic) */
    int bundle_filter;         /* 1 = skip attenuators no sub-ray of an element can touch */
    int device;                /* HIP device id */
} SmPeOptions;

/* Reference defaults: voxel centre, 8x16x8 crystal sub-cells, all detectors,
 * whole volume, rotation 0, automatic tiles, filter on, device 0. */
int sm_pe_default_options(SmPeOptions* o);

/* Drop-in for PESysMatGen (PESysMatGen.cu:766-833): one Z-slab described
 * in-band by Img[2] (slab Z), Img[30] (full Z), Img[31] (z0), Img[20]
 * (rotation); dst (host) receives N x (X*Y*Img[2]) floats, det-major.
 * Returns X*Y*Img[2] (numImagebin) like the reference, or an SM_E* code. */
int sm_PESysMatGen(const float* col, const float* det, const float* img, float* dst, int cuda_id);

/* PE rows [det_begin, det_begin+det_count) x voxels [z_begin, z_begin+z_count)
 * into host memory dst (det_count x X*Y*z_count floats, det-major).
 * Img[0..11] are read as in Params_Image.dat; the slab fields are taken from o. */
int sm_pe_compute(const float* col, const float* det, const float* img, const SmPeOptions* o, float* dst);

/* Same into device memory dst_dev on HIP stream `stream` (NULL = default);
 * blocks until done.  row_stride = floats between detector rows of dst_dev
 * (0 = X*Y*z_count). */
int sm_pe_compute_device(const float* col, const float* det, const float* img, const SmPeOptions* o,
                         float* dst_dev, long long row_stride, void* stream);

/* Allocation-free form of sm_pe_compute_device.  No counterpart in the
 * reference (whose PESysMatGen allocates per call, PESysMatGen.cu:792-804); it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The caller
 * passes inputs and dst_dev; the geometry tables, per-detector candidate lists
 * and status word live in a file-scope device array of
 * sm_pe_noalloc_scratch_bytes() (64 MiB) per GPU, held until the process ends.
 * A geometry whose tables and lists exceed it returns SM_E_ARG.  Calls share
 * that array: calls that overlap on one GPU (several host threads) race.
 * Results equal sm_pe_compute_device's bit for bit.  Blocks until done. */
long long sm_pe_noalloc_scratch_bytes(void);
int sm_pe_compute_device_noalloc(const float* col, const float* det, const float* img, const SmPeOptions* o,
                                 float* dst_dev, long long row_stride, void* stream);

/* Tile-shape autotuner (no counterpart in the reference): times candidate voxel
 * tiles on ~sample_rows detector rows spread over the detector (sampling and
 * image as in `base`) and returns `base` with the fastest tile in `best`.
 * cache_path (nullable): text cache keyed by geometry, image, sampling and GPU
 * architecture; a cached answer returns 1 without timing, a fresh one 0. */
int sm_pe_autotune(const float* col, const float* det, const float* img, const SmPeOptions* base, int sample_rows,
                   const char* cache_path, SmPeOptions* best, double* best_ns_per_element);

/* Plans keep the geometry tables resident in HBM for repeated launches. */
typedef struct SmPePlan SmPePlan;
int sm_pe_plan_create(const float* col, const float* det, const float* img, const SmPeOptions* o, SmPePlan** plan);
/* Asynchronous launch of detector rows [det_begin, det_begin+det_count) of the
 * plan into dst_dev (row stride as above) on `stream`. */
int sm_pe_plan_launch(SmPePlan* plan, int det_begin, int det_count, float* dst_dev, long long row_stride,
                      void* stream);
/* Synchronise the plan's last stream; SM_E_OVERFLOW if a launch overflowed. */
int sm_pe_plan_sync(SmPePlan* plan);
int sm_pe_plan_destroy(SmPePlan* plan);
/* Algorithmic sub-ray count per element of the plan (voxel_sub^3 * sub_x*sub_y*sub_z). */
long long sm_pe_plan_subrays(const SmPePlan* plan);

/* ---- primary Compton scatter ------------------------------------------- */

/* GeometryRelationShip bitmaps (raw uint32 words, LSB-first, the reference's
 * file layout): crystal bit ((pair(min,max) of scatter/detector) * N + k),
 * pair(i,j) = i*N - i(i+1)/2 + j (scatter.cu:1040-1058, builder 1440-1540);
 * collimator bit (hole*N*N + det*N + k) (scatter.cu:1061-1074, builder 1939-2009). */
long long sm_crystal_bitmap_words(int n_det);
long long sm_collimator_bitmap_words(int n_holes_layer0, int n_det);
int sm_bitmap_crystal(const float* det, unsigned* words, long long n_words, int device);
int sm_bitmap_collimator(const float* col, const float* det, unsigned* words, long long n_words, int device);

/* Drop-in for the combined scatter() (ScatterGen_RayTracing_CircularHole/
 * scatter.cu:2013-2378): crystal + collimator scatter of rotation Img[20] for the
 * N x V PE matrix `pe`; bitmaps computed and written when Phys[8]/[9] == 1, else
 * read from the files.  dst: N x V floats.  Returns V or an SM_E* code. */
int sm_scatter(const float* col, const float* det, const float* img, const float* phy, const float* pe,
               const char* fname_geo_crystal, const char* fname_geo_collimator, float* dst, int cuda_id);
/* Drop-in for ScatterGen_Collimator's scatter() (ScatterGen_Collimator/scatter.cu:1511-1747). */
int sm_scatter_collimator_only(const float* col, const float* det, const float* img, const float* phy,
                               const float* pe, const char* fname_geo, float* dst, int cuda_id);
/* Drop-in for ScatterGen_Crystal's scatter() (ScatterGen_Crystal/scatter.cu:1528-1738). */
int sm_scatter_crystal_only(const float* det, const float* img, const float* phy, const float* pe,
                            const char* fname_geo, float* dst, int cuda_id);

/* Host-memory row blocks: crystal scatter (crystalScatterSysMatCuda, scatter.cu:1077-1437)
 * and collimator scatter (variant 0: collimatorScatterSysMatCuda scatter.cu:1544-1936,
 * variant 1: ScatterGen_Collimator/scatter.cu:1056-1436).  Bitmap words may be NULL
 * (then computed on the GPU).  dst: det_count x V floats. */
int sm_scatter_crystal(const float* det, const float* img, const float* phy, const float* pe,
                       const unsigned* crystal_words, int rotation, int det_begin, int det_count, float* dst,
                       int device);
int sm_scatter_collimator(const float* col, const float* det, const float* img, const float* phy,
                          const unsigned* collimator_words, int rotation, int variant, int det_begin, int det_count,
                          float* dst, int device);

/* Plans: geometry, bitmaps and per-row tables resident in HBM.  col may be NULL
 * (crystal scatter only).  Launches are synchronous on `stream`. */
typedef struct SmScatterPlan SmScatterPlan;
int sm_scatter_plan_create(const float* col, const float* det, const float* img, const float* phy, int rotation,
                           int collimator_variant, const unsigned* crystal_words, const unsigned* collimator_words,
                           int device, SmScatterPlan** plan);
/* flags: without a bitmap the pair preparation evaluates the same attenuator
 * criterion inline (identical results; no O(N^3)-bit table, e.g. 275 GB at N = 16384). */
#define SM_PLAN_NO_CRYSTAL_BITMAP 1
#define SM_PLAN_NO_COLLIMATOR_BITMAP 2
int sm_scatter_plan_create_ex(const float* col, const float* det, const float* img, const float* phy, int rotation,
                              int collimator_variant, int flags, const unsigned* crystal_words,
                              const unsigned* collimator_words, int device, SmScatterPlan** plan);
/* Crystal scatter of the listed detector rows into dst_dev rows 0..n_rows-1. */
int sm_scatter_plan_crystal_rows(SmScatterPlan* plan, const float* pe_dev, const int* rows, int n_rows,
                                 float* dst_dev, long long row_stride, void* stream);
int sm_scatter_plan_crystal(SmScatterPlan* plan, const float* pe_dev, int det_begin, int det_count, float* dst_dev,
                            long long row_stride, void* stream);
int sm_scatter_plan_collimator(SmScatterPlan* plan, int det_begin, int det_count, float* dst_dev,
                               long long row_stride, void* stream);
/* Collimator scatter with the plates cut into cells of cell_size mm, each cell a
 * Compton site.  No counterpart in the reference, whose collimator scatter puts
 * one site at each layer-0 hole with a virtual-disc weight and no collimator
 * attenuation (scatter.cu:1544-1936, replaced by sm_scatter_plan_collimator).
 * Cell weight mu_c vol / (4 pi d^2) exp(-plate path in), plate path out scaled to
 * E' like the crystal attenuators; the rest is the hole-site physics.  keep_dist < 0
 * keeps every material cell, else cells whose centre lies within keep_dist mm of a
 * hole wall of their plate.  cells_per_chunk 0: automatic (cell-weight table
 * <= 4 GiB).  Synchronous on `stream`.  The plan must have a collimator. */
int sm_scatter_plan_cells(SmScatterPlan* plan, float cell_size, float keep_dist, int det_begin, int det_count,
                          float* dst_dev, long long row_stride, int cells_per_chunk, void* stream);
/* The kept cells of (cell_size, keep_dist): up to `cap` records of 4 doubles
 * (x, y with FOV2C, z, mu_c vol / (4 pi)) into `out` (may be NULL); returns the count. */
long long sm_scatter_cell_table(SmScatterPlan* plan, float cell_size, float keep_dist, double* out, long long cap);
/* Host-memory form of sm_scatter_plan_cells: rows [det_begin, det_begin + det_count)
 * (det_count <= 0: to the last detector) of rotation `rotation` into dst (rows x V). */
int sm_scatter_cells(const float* col, const float* det, const float* img, const float* phy, int rotation,
                     float cell_size, float keep_dist, int det_begin, int det_count, float* dst, int device);
int sm_scatter_plan_destroy(SmScatterPlan* plan);

/* ---- sparse output stage (no counterpart in the reference, whose only output
 * is the dense writer PEGen_CircularHole.cpp:201-216; the dense layout stays the
 * drop-in default).  An element v of a row is kept when v > 0 and
 * (double)v >= rel * (double)max(0, row max).  Columns ascend within a row. */

/* Row maxima (row_max_dev: rows floats) and row pointers (row_ptr_dev: rows+1
 * int64, exclusive prefix of the kept counts) of `rows` dense device rows of
 * `cols` floats, row_stride floats apart (0 = cols).  Synchronises `stream` and
 * returns the total kept count in *nnz. */
int sm_csr_count_device(const float* dense_dev, int rows, long long cols, long long row_stride, double rel,
                        float* row_max_dev, long long* row_ptr_dev, long long* nnz, void* stream);
/* Fill column indices (int32) and values (float32), nnz entries each, from the
 * outputs of sm_csr_count_device on the same rows.  Asynchronous on `stream`. */
int sm_csr_fill_device(const float* dense_dev, int rows, long long cols, long long row_stride, double rel,
                       const float* row_max_dev, const long long* row_ptr_dev, int* col_dev, float* val_dev,
                       void* stream);

/* Host-side CSR result of PE rows [det_begin, det_begin+det_count) (options as
 * sm_pe_compute): computed in dense device blocks, compacted on the GPU. */
typedef struct SmCsr SmCsr;
int sm_pe_compute_csr(const float* col, const float* det, const float* img, const SmPeOptions* o, double rel,
                      SmCsr** out);
int sm_csr_info(const SmCsr* c, long long* rows, long long* cols, long long* nnz);
/* Copy out row_ptr (rows+1 int64), col_idx (nnz int32), val (nnz float32); NULL skips one. */
int sm_csr_copy(const SmCsr* c, long long* row_ptr, int* col_idx, float* val);
/* Write parts stacked row-wise to a .csr file: 64-byte header ("SMCSR\0\0\1",
 * int64 rows, int64 cols, int64 nnz, float64 rel, 24 zero bytes), then
 * row_ptr[rows+1] int64, col[nnz] int32, val[nnz] float32, little endian. */
int sm_csr_write(const char* path, const SmCsr* const* parts, int n_parts);
int sm_csr_destroy(SmCsr* c);

/* Beam masks and multiplexing index of device CSR rows (row_ptr rows+1 int64,
 * col int32, val float32, as from sm_csr_count_device / sm_csr_fill_device).  The
 * reference computes these from dense .sysmat files on the host
 * (PE/extract_beam_prop_masks.py, PE/compute_mpxi_3d.py); here they read the CSR
 * output on the GPU.  An element is in its row's beam when (double)v > frac *
 * (double)(row max); mpxi_dev (uint32, one per column) gains 1 per element in a
 * beam (it accumulates: clear it first); beam_size_dev (int32 per row, nullable)
 * receives each row's beam size.  Asynchronous on `stream`. */
int sm_csr_beams_device(const long long* row_ptr_dev, const int* col_dev, const float* val_dev, int rows,
                        double frac, unsigned* mpxi_dev, int* beam_size_dev, void* stream);

/* ---- scatter-surrogate MLP (no counterpart in the reference) ------------
 * 12 -> 256 -> 256 -> 256 -> 256 -> 1, SiLU; bf16 operands on the matrix cores,
 * FP32 accumulation.  Parameters are sm_mlp_param_count() FP32 values in device
 * memory (layout: W1[256][32] (inputs 12..31 zero), b1, W2[256][256], b2, W3,
 * b3, W4, b4, w5[256], b5).  Rows of x are 12 FP32 features.  All calls are
 * asynchronous on `stream`. */
long long sm_mlp_param_count(void);
long long sm_mlp_packed_bytes(void);
/* Device workspace for a training batch (batch: positive multiple of 64). */
long long sm_mlp_workspace_bytes(int batch);
/* Uniform +-sqrt(6/fan_in) weights, zero biases, from a counter-based RNG. */
int sm_mlp_init(float* params_dev, unsigned long long seed, void* stream);
/* Training-path forward of n <= batch rows: y_dev[n]; keeps activations in ws. */
int sm_mlp_forward(const float* params_dev, const float* x_dev, int n, int batch, void* ws, long long ws_size,
                   float* y_dev, void* stream);
/* Gradient of sum(dy * y) for the rows of the last sm_mlp_forward on ws. */
int sm_mlp_backward(const float* params_dev, const float* dy_dev, int n, int batch, void* ws, long long ws_size,
                    float* grad_dev, void* stream);
/* One Adam step (beta 0.9/0.999, eps 1e-8, step counts from 1) on the loss
 * sum_i w_i (y_i - t_i)^2 (w_dev NULL: w_i = 1/n); loss_dev (nullable) gets it. */
int sm_mlp_train_step(float* params_dev, float* adam_m, float* adam_v, const float* x_dev, const float* t_dev,
                      const float* w_dev, int n, int batch, float lr, int step, void* ws, long long ws_size,
                      float* loss_dev, void* stream);
/* `steps` Adam steps for each of n_members (1..8) networks (member m: its
 * params/adam_m/adam_v/ws, batch seed seeds[m]) on batches drawn on the device:
 * step s = step0 + 1 .. step0 + steps takes `batch` rows of the n training pairs
 * (x_dev [n][12], t_dev [n], w_dev [n] or NULL for equal weights) at indices
 * hashed from (seeds[m], s), the weights normalised to sum 1 per batch;
 * learning rate max(lr (1 + cos(pi i / steps)) / 2, lr_floor) at step i of the
 * run.  No host round trip per step; the members run concurrently on their own
 * streams, joined back to `stream` before return.  loss_dev (nullable, or
 * per-member entries NULL) gets each member's last-batch loss. */
int sm_mlp_train_run(int n_members, float* const* params_dev, float* const* adam_m, float* const* adam_v,
                     const float* x_dev, const float* t_dev, const float* w_dev, long long n, int batch, int steps,
                     float lr, float lr_floor, const unsigned long long* seeds, int step0, void* const* ws,
                     long long ws_size, float* const* loss_dev, void* stream);
/* Repack parameters for inference (sm_mlp_packed_bytes() bytes per member). */
int sm_mlp_pack(const float* params_dev, void* packed_dev, void* stream);
/* Register-resident inference of n rows for n_members consecutive packed
 * members: y_dev[m*n + i]. */
int sm_mlp_infer(const void* packed_dev, int n_members, const float* x_dev, long long n, float* y_dev, void* stream);

/* ---- batched scatter over collimator variants (no counterpart in the reference,
 * which handles one geometry per process) ------------------------------------
 * Variants share the detector, image and physics; cols[v] are their collimator
 * parameter arrays.  run: dst_devs[v] = collimator term (combined model,
 * scatter.cu:1544-1936) + crystal term (scatter.cu:1077-1437) of variant v for
 * the N x V PE matrix pe_devs[v], attenuator criterion evaluated inline.  With
 * use_graph != 0 the launch sequence is captured once into a hipGraph (on
 * `stream`, or an internal non-blocking stream when NULL) and replayed;
 * synchronous. */
typedef struct SmScatterBatch SmScatterBatch;
int sm_scatter_batch_create(int n_variants, const float* const* cols, const float* det, const float* img,
                            const float* phy, int rotation, int device, SmScatterBatch** out);
int sm_scatter_batch_run(SmScatterBatch* b, const float* const* pe_devs, float* const* dst_devs, void* stream,
                         int use_graph);
/* Nodes of the captured graph (0 before the first graph run). */
int sm_scatter_batch_graph_nodes(const SmScatterBatch* b);
int sm_scatter_batch_destroy(SmScatterBatch* b);

/* ---- crystal-scatter surrogate (no counterpart in the reference) ----------
 * Exact near field + MLP far field.  Per detector the n_near + n_ring sites with
 * the largest single-site weight at the image centre are ranked once; the
 * n_near best are summed exactly for every voxel (S_near), the MLP predicts the
 * rest from 12 features (one uses the fixed-weight ring of the next n_ring
 * sites), ensemble members that disagree send their 4 x 4 x 4 voxel brick to the
 * exact kernel.  pe_dev: the full N x V PE matrix of the rotation (row stride V). */
typedef struct SmSurrogate SmSurrogate;
int sm_surrogate_create(const float* det, const float* img, const float* phy, int rotation,
                        const unsigned* crystal_words, int use_bitmap, int n_near, int n_ring, int device,
                        SmSurrogate** out);
/* Feature / target normalisation: PE scale, scatter scale, target mean and
 * deviation of log(S_far + phi), phi. */
int sm_surrogate_set_norm(SmSurrogate* s, float pe_scale, float s_scale, float mu_y, float sig_y, float phi);
/* Ranked sites of one row (site index, weight); returns the count. */
int sm_surrogate_row_sites(const SmSurrogate* s, int row, int* sites, float* weights, int cap);
/* S_near of rows [det_begin, +det_count) into dst_dev (asynchronous). */
int sm_surrogate_near(SmSurrogate* s, const float* pe_dev, int det_begin, int det_count, float* dst_dev,
                      long long row_stride, void* stream);
/* Exact crystal scatter of listed rows into dst_dev rows 0..n_rows-1 (synchronous). */
int sm_surrogate_exact_rows(SmSurrogate* s, const float* pe_dev, const int* rows, int n_rows, float* dst_dev,
                            long long row_stride, void* stream);
/* 12 features per (row, voxel): rows [det_begin, +det_count), voxels vox_dev[0..n_vox)
 * (NULL: 0..n_vox-1), S_near of those rows in near_dev; x_dev[(row*n_vox + j)*12 + f]. */
int sm_surrogate_features(SmSurrogate* s, const float* pe_dev, const float* near_dev, long long near_stride,
                          int det_begin, int det_count, const int* vox_dev, int n_vox, float* x_dev, void* stream);
/* Voxel tiles per row: 4 x 4 x 4 bricks, ceil(X/4) ceil(Y/4) ceil(Z/4). */
long long sm_surrogate_tile_count(const SmSurrogate* s);
/* dst_dev holds S_near on entry and S_near + predicted far field on exit;
 * tiles where members differ by more than tol x value (value > floor_abs) are
 * marked in tile_flags_dev (det_count x sm_surrogate_tile_count bytes, zeroed by
 * the caller). */
int sm_surrogate_apply(SmSurrogate* s, const void* packed_dev, int n_members, const float* pe_dev, int det_begin,
                       int det_count, float* dst_dev, long long row_stride, float tol, float floor_abs,
                       unsigned char* tile_flags_dev, void* stream);
/* Exact recomputation of the marked tiles; returns the number of tiles. */
long long sm_surrogate_fallback(SmSurrogate* s, const float* pe_dev, int det_begin, int det_count,
                                const unsigned char* tile_flags_dev, float* dst_dev, long long row_stride,
                                void* stream);
int sm_surrogate_destroy(SmSurrogate* s);

#ifdef __cplusplus
}
#endif
#endif /* SYSMAT_H_ */
