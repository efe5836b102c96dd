/* gvdb_amd.h -- C ABI of the MI355X point-to-sparse-volume library.
 *
 * Each entry point replaces one VolumeGVDB method of NVIDIA GVDB Voxels 1.1.1
 * (class nvdb::VolumeGVDB, source/gvdb_library/src/gvdb_volume_gvdb.h:319-901
 * of the reference); the comment above each names the method it replaces as
 * file:line of the reference.  Arguments keep the reference's meaning.
 *
 * Conventions
 *   - One handle per grid; a handle owns one HIP stream on one device.  Calls
 *     on one handle are not thread-safe (as the reference).  Work is enqueued
 *     on the handle's stream; calls that return sizes to the host (bounding
 *     box, node counts, subcell entry count) synchronise that stream, as the
 *     reference's blocking copies do.
 *   - Every function returns a status: GVDB_OK (0) or a negative code; the
 *     message of the last failure is available from gvdb_last_error().  The
 *     reference prints and returns (soft errors) or exits (hard errors).
 *   - Device pointers are plain `const void*` / `void*` (HIP device memory);
 *     the library never frees memory it did not allocate.
 *   - Data layout kept from the reference: 64-byte Node records in one pool
 *     per level (group 0), dense child lists of res^3 uint64 references per
 *     interior node (group 1), references packed grp | lev<<8 | idx<<16,
 *     atlas slots placed x-fastest in a cnt.x*cnt.y*cnt.z grid of padded
 *     bricks with mValue = slot corner + apron.
 */
#ifndef GVDB_AMD_H
#define GVDB_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GVDB_OK 0
#define GVDB_ERR_INVALID (-1)     /* bad argument */
#define GVDB_ERR_HIP (-2)         /* HIP runtime failure */
#define GVDB_ERR_STATE (-3)       /* call order (e.g. Configure before Initialize) */
#define GVDB_ERR_DEPTH (-4)       /* point extents below one brick or above the tree */
#define GVDB_ERR_NOBRICKS (-5)    /* InsertPointsSubcell with no bricks (reference: warning) */
#define GVDB_ERR_RANGE (-6)       /* coordinates outside what the topology build supports */
#define GVDB_ERR_UNSUPPORTED (-7) /* feature of the reference not provided */
#define GVDB_ERR_OOM (-8)         /* device allocation failed */
#define GVDB_ERR_IO (-9)          /* file missing, unreadable or malformed (VBX) */

/* channel types (src/gvdb_types.h of the reference); T_HALF is an addition */
#define GVDB_T_UCHAR4 2
#define GVDB_T_FLOAT 3
#define GVDB_T_HALF 100
#define GVDB_CHAN_UNDEF 255 /* no colour channel */

/* shading modes (kernels/cuda_gvdb_scene.cuh:25-32 of the reference) */
#define GVDB_SHADE_TRILINEAR 4
#define GVDB_SHADE_VOLUME 7

/* pool group / level node record (src/gvdb_node.h:27-40 of the reference) */
typedef struct gvdb_node {
    uint8_t lev, flags, priority, pad;
    int32_t pos[3];
    int32_t value[3];
    float vrange[3];
    uint64_t parent;
    uint64_t childlist;
    uint64_t mask;
} gvdb_node;

typedef struct gvdb_volume* gvdb_handle;

/* VolumeGVDB() + SetCudaDevice + Initialize: gvdb_volume_gvdb.cpp:60-172, :240-334, :2320-2361 */
int gvdb_create(gvdb_handle* out, int device);
int gvdb_destroy(gvdb_handle h);
const char* gvdb_last_error(gvdb_handle h);

/* Configure(levs, r, numcnt): gvdb_volume_gvdb.cpp:2380-2434.  log2dims leaf
 * first; initcnt may be NULL for the Configure(q4..q0) defaults (:2364-2377). */
int gvdb_configure(gvdb_handle h, int levels, const int* log2dims, const int* initcnt);
/* SetChannelDefault: gvdb_volume_gvdb.h (Initialize sets 16,16,1) */
int gvdb_set_channel_default(gvdb_handle h, int cx, int cy, int cz);
/* AddChannel(chan, dt, apron, filter, border, axiscnt): gvdb_volume_gvdb.cpp:2437-2453 */
int gvdb_add_channel(gvdb_handle h, int chan, int dtype, int apron, int filter, int border, int ax, int ay, int az);
/* SetEpsilon(eps, maxiter): gvdb_volume_gvdb.h:334 */
int gvdb_set_epsilon(gvdb_handle h, float eps, int maxiter);

/* SetDataGPU + SetPoints(pos, vel, clr): gvdb_volume_gvdb.cpp:4872-4880, :4900-4908.
 * Non-owning float3 positions at dptr + i*stride + offset. */
int gvdb_set_points_gpu(gvdb_handle h, const void* dptr, int n, int offset, int stride);
/* AllocData + CommitData: gvdb_volume_gvdb.cpp:4821-4862 (owning copy of host float3 positions) */
int gvdb_upload_points(gvdb_handle h, const float* host_xyz, int n);
/* SetPoints' colour DataPtr (gvdb_volume_gvdb.cpp:4900-4908): a uint32 (uchar4 r, g, b, a) per point at
 * dptr + i*stride + offset, not owned; null (or n = 0) clears it.  InsertPointsSubcell carries the colours with the
 * points; GatherDensity / GatherLevelSet then write chanClr (cuda_gvdb_particles.cuh:717-731). */
int gvdb_set_colors_gpu(gvdb_handle h, const void* dptr, int n, int offset, int stride);
/* the same from host memory (copied into a buffer the handle owns) */
int gvdb_upload_colors(gvdb_handle h, const uint32_t* host_rgba, int n);
/* SetColorChannel(chan): gvdb_volume_gvdb.cpp:686-689; the march reads that uchar4 channel's colours
 * (getColorF, cuda_gvdb_raycast.cuh:206-209); GVDB_CHAN_UNDEF or -1: none */
int gvdb_set_color_channel(gvdb_handle h, int chan);

/* RequestFullRebuild: gvdb_volume_gvdb.h:432 */
int gvdb_request_full_rebuild(gvdb_handle h, int on);
/* RebuildTopology(num_pnts, radius, orig): gvdb_volume_gvdb.cpp:1181-1218.  A full rebuild after Configure or
 * RequestFullRebuild, or when the depth changes (the second call after Configure is one); otherwise the incremental
 * build of ActivateIncreBricksGPU (:1020-1156) without accumulation. */
int gvdb_rebuild_topology(gvdb_handle h, int n, float radius, const float orig[3]);
/* AccumulateTopology(num_pnts, radius, orig, iDepth): gvdb_volume_gvdb.cpp:1158-1178, the incremental build keeping
 * the existing leaves; GVDB_ERR_DEPTH when the points need a deeper tree than the current one. */
int gvdb_accumulate_topology(gvdb_handle h, int n, float radius, const float orig[3], int min_depth);
/* RebuildTopologyCPU(num_pnts, orig, pos): gvdb_volume_gvdb.cpp:1220-1227 (host build) */
int gvdb_rebuild_topology_cpu(gvdb_handle h, int n, const float orig[3], const float* host_xyz);
/* ActivateSpace(pos): gvdb_volume_gvdb.cpp:2766-2774 (host build) */
int gvdb_activate_space(gvdb_handle h, const float pos[3]);
/* Clear: gvdb_volume_gvdb.cpp:2487-2503 */
int gvdb_clear(gvdb_handle h);
/* FinishTopology(commit, bounds): gvdb_volume_gvdb.cpp:1579-1593 */
int gvdb_finish_topology(gvdb_handle h, int commit_pool, int compute_bounds);
/* UpdateAtlas: gvdb_volume_gvdb.cpp:2630-2704 */
int gvdb_update_atlas(gvdb_handle h);
/* ClearChannel: gvdb_volume_gvdb.cpp:1596-1606 */
int gvdb_clear_channel(gvdb_handle h, int chan);

/* InsertPointsSubcell(subcell_size, num_pnts, radius, trans, &len): gvdb_volume_gvdb.cpp:4971-5092 */
int gvdb_insert_points_subcell(gvdb_handle h, int subcell_size, int n, float radius, const float trans[3], int* out_len);
/* GatherDensity(subcell_size, num_pnts, radius, trans, len, chanDensity, chanClr, bAccum): gvdb_volume_gvdb.cpp:5318-5344 */
int gvdb_gather_density(gvdb_handle h, int subcell_size, int n, float radius, const float trans[3], int sc_len,
                        int chan_density, int chan_clr, int accumulate);
/* GatherLevelSet(subcell_size, num_pnts, radius, trans, len, chanLevelset, chanClr, bAccum): gvdb_volume_gvdb.cpp:5348-5371,
 * kernel gvdbGatherLevelSet (kernels/cuda_gvdb_particles.cuh:736-774): min(3, min over the subcell's entries of
 * |p - v| - r), from the atlas value when accumulating */
int gvdb_gather_level_set(gvdb_handle h, int subcell_size, int n, float radius, const float trans[3], int sc_len,
                          int chan_levelset, int chan_clr, int accumulate);
/* InsertPointsSubcell_FP16(subcell_size, num_pnts, radius, trans, &len): gvdb_volume_gvdb.cpp:5094-5189,
 * kernel gvdbInsertSubcell_fp16 (kernels/cuda_gvdb_particles.cuh:154-216): the subcell tables, each entry stored as
 * u16((p + trans - pos_min) / pos_range * 65535) per axis against the bounds of the last RebuildTopology
 * (saturated outside them).  Read the u16 entries with gvdb_read_subcells(h, 6, ...). */
int gvdb_insert_points_subcell_fp16(gvdb_handle h, int subcell_size, int n, float radius, const float trans[3],
                                    int* out_len);
/* GatherLevelSet_FP16(subcell_size, num_pnts, radius, trans, len, chanLevelset, chanClr): gvdb_volume_gvdb.cpp:5374-5401,
 * kernel gvdbGatherLevelSet_fp16 (kernels/cuda_gvdb_particles.cuh:776-811): min(3, min (|q - v| - r)) over the
 * subcell's entries q = u / 65535 * pos_range + pos_min; no accumulation */
int gvdb_gather_level_set_fp16(gvdb_handle h, int subcell_size, int n, float radius, const float trans[3], int sc_len,
                               int chan_levelset, int chan_clr);
/* Compute(effect, channel, num_iterations, parameters, bUpdateApron, skipOverAprons, boundval):
 * gvdb_volume_gvdb.cpp:4527-4552, kernels cuda_gvdb_operators.cuh:148-455.  Float and half channels:
 * GVDB_FUNC_FILL_F (p1, or a sine pattern when p3 < 0), GVDB_FUNC_SMOOTH ((p1 v + six face neighbours) / (p1 + 6)
 * + p2), GVDB_FUNC_NOISE (v + random(voxel) p1 where v > 0.01), GVDB_FUNC_GROW (v + 10 p1 where v != 0, then 0 below
 * 0.01); the uchar effects return GVDB_ERR_UNSUPPORTED.  UpdateApron(channel, boundval) after every iteration when
 * update_apron is set. */
#define GVDB_FUNC_FILL_F 110
#define GVDB_FUNC_SMOOTH 113
#define GVDB_FUNC_NOISE 114
#define GVDB_FUNC_GROW 115
int gvdb_compute(gvdb_handle h, int effect, int chan, int iterations, const float params[3], int update_apron,
                 int skip_aprons, float boundval);
/* ComputeKernel(user_module, user_kernel, channel, bUpdateApron, skipOverAprons): gvdb_volume_gvdb.cpp:4499-4525.
 * hip_function is a hipFunction_t; the kernel's arguments and launch shape are in gvdb_amd_device.h.  With
 * update_apron every channel's apron is refreshed afterwards (UpdateApron()). */
int gvdb_compute_kernel(gvdb_handle h, void* hip_function, int chan, int update_apron, int skip_aprons);
/* RenderKernel(user_kernel, in_channel, outbuf): gvdb_volume_gvdb.cpp:4309-4333; 8x8 blocks over the render buffer. */
int gvdb_render_kernel(gvdb_handle h, void* hip_function, int chan, int rbuf);
/* Loading a code object (no counterpart: the reference's applications call cuModuleLoad themselves):
 * hipModuleLoad / hipModuleGetFunction / hipModuleUnload on the current device. */
int gvdb_module_load(const char* code_object_path, void** hip_module);
int gvdb_module_function(void* hip_module, const char* kernel_name, void** hip_function);
int gvdb_module_unload(void* hip_module);
/* SaveVBX(fname): gvdb_volume_gvdb.cpp:1626-1768.  VBX 1.11 with one grid: header, GVDB topology pools, each
 * channel's atlas in the reference's 3D layout (GVDB_FILESPEC.txt). */
int gvdb_save_vbx(gvdb_handle h, const char* path);
/* LoadVBX(fname, force_maj, force_min): gvdb_volume_gvdb.cpp:507-685.  Configure from the file, pools, FinishTopology,
 * channels, UpdateAtlas.  Files with bitmask nodes (1.0, or 2.x with masks) and non-float grids return
 * GVDB_ERR_UNSUPPORTED; unreadable or malformed files GVDB_ERR_IO. */
int gvdb_load_vbx(gvdb_handle h, const char* path, int force_major, int force_minor);
/* The grid transform (SetTransform's fields, as a VBX file carries them): pretrans, angs, scale, trans. */
int gvdb_get_transform(gvdb_handle h, float out12[12]);
/* SetTransform(pretrans, scale, angs, trans): gvdb_volume_gvdb.cpp:5770-5793; places the grid in application space as
 * T R S PT (R = Rz Ry Rx, angles in degrees); Render then takes the camera rays and the light into index space
 * through its inverse. */
int gvdb_set_transform(gvdb_handle h, const float pretrans[3], const float scale[3], const float angs[3],
                       const float trans[3]);
/* UpdateApron(chan, boundval): gvdb_volume_gvdb.cpp:4418-4457 */
int gvdb_update_apron(gvdb_handle h, int chan, float boundval);
/* UpdateApron without allocation.  No counterpart in the reference: an addition for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same result as
 * gvdb_update_apron; the neighbour table it builds lives in a fixed device array of GVDB_NOALLOC_SCRATCH_BYTES that
 * each device holds from the first gvdb_create on it until the process ends, so grids of more than
 * GVDB_NOALLOC_MAX_BRICKS bricks are refused (GVDB_ERR_OOM), as are pools built on the host and not yet committed
 * (GVDB_ERR_STATE: call gvdb_update_atlas first).  Calls no hipMalloc/hipFree.  Every call on one device shares that
 * array: calls that overlap on one device (other handles' streams included) race; order them on the host. */
#define GVDB_NOALLOC_SCRATCH_BYTES (32u << 20)
#define GVDB_NOALLOC_MAX_BRICKS 103563 /* GVDB_NOALLOC_SCRATCH_BYTES / (27 * 12) */
int gvdb_update_apron_noalloc(gvdb_handle h, int chan, float boundval);

/* Scene: Camera3D::setOrbit + fov/near/far (gvdb_camera.cpp:90-105, :175-228),
 * Light position, Scene setters SetBackgroundClr/SetExtinct/SetSteps/SetCutoff/
 * SetVolumeRange/SetShadowParams (gvdb_scene.h:142-153, gvdb_scene.cpp:492-495). */
int gvdb_set_camera_orbit(gvdb_handle h, const float angs[3], const float to[3], float dist, float dolly, float fov,
                          float znear, float zfar);
int gvdb_set_light(gvdb_handle h, const float pos[3]);
int gvdb_set_background(gvdb_handle h, const float rgba[4]);
int gvdb_set_extinct(gvdb_handle h, float a, float b, float c);
int gvdb_set_steps(gvdb_handle h, float direct, float shadow, float fine);
int gvdb_set_cutoff(gvdb_handle h, float a, float b, float c);
int gvdb_set_volume_range(gvdb_handle h, float thresh, float vmin, float vmax);
int gvdb_set_shadow_params(gvdb_handle h, float amount, float bias, float c);
/* LinearTransferFunc(t0, t1, a, b) + CommitTransferFunc: gvdb_scene.cpp:66-88, gvdb_volume_gvdb.cpp:4892-4898 */
int gvdb_linear_transfer_func(gvdb_handle h, float t0, float t1, const float a[4], const float b[4]);
int gvdb_commit_transfer_func(gvdb_handle h);
/* AddRenderBuf(buf, w, h, bytes_per_pixel): gvdb_volume_gvdb.h; bytes 4 = uchar4 (reference), 16 = float4 */
int gvdb_add_render_buf(gvdb_handle h, int buf, int w, int hgt, int bytes_per_pixel);
/* Render(shading, chan, buf): gvdb_volume_gvdb.cpp:4336-4381 */
int gvdb_render(gvdb_handle h, int shading, int chan, int buf);
/* ReadRenderBuf(buf, dst): gvdb_volume_gvdb.cpp:4241-4251 */
int gvdb_read_render_buf(gvdb_handle h, int buf, void* host_dst);
/* Autotuner (no counterpart in the reference; for callers that run one workload shape many times).  The launch-shape
 * knobs below change no result, only time; each tuner runs the caller's own workload once per value (a warm-up, then
 * the median of three runs timed with HIP events on the handle's or batch's stream) and keeps the fastest,
 * process-wide.  Times go to ms[] in the order of the knob's values (may be NULL). */
#define GVDB_KNOB_RENDER_UNROLL_EA 0  /* samples fetched ahead per brick step, SHADE_VOLUME: 1, 2, 4 */
#define GVDB_KNOB_RENDER_UNROLL_ISO 1 /* the same, SHADE_TRILINEAR */
#define GVDB_KNOB_RENDER_TILE_EA 2    /* ray-packet width of a 64-lane workgroup (8x8, 16x4, 32x2, 64x1): 8..64 */
#define GVDB_KNOB_RENDER_TILE_ISO 3
#define GVDB_KNOB_GATHER_THREADS 4    /* threads per brick of the scatter gather, radius <= 3: 128, 256 */
#define GVDB_KNOB_APRON_RUN 5         /* bricks per workgroup of the multi-grid apron: 1, 2, 4, 8 */
#define GVDB_KNOB_XCD_ORDER 6         /* per-brick kernels: 0 launch order, 1 each XCD a contiguous range of bricks */
#define GVDB_KNOB_APRON_QUEUE 7       /* multi-grid apron: 0 static runs, 1 persistent workgroups over a brick queue */
#define GVDB_KNOB_MARK_PATH 8         /* topology marks: 0 by domain size, 1 byte per brick, 2 direction masks, 3 bits,
                                         4 sorted brick keys (taken above 2^34 level-0 bricks) */
#define GVDB_KNOB_APRON_RUN_ONE 9     /* bricks per workgroup of the one-grid apron: 1, 2, 4, 8 */
#define GVDB_KNOB_COUNT 10
int gvdb_set_tuning(int knob, int value);
/* current value, and the values it can take (choices8: room for 8, may be NULL; nchoices of them) */
int gvdb_get_tuning(int knob, int* value, int* choices8, int* nchoices);
/* Render: samples fetched ahead (1, 2, 4) at the current ray-packet shape; ms3 per value */
int gvdb_tune_render(gvdb_handle h, int shading, int chan, int buf, int* best_unroll, float* ms3);
/* Render: ray-packet shape x samples fetched ahead; best2 = {unroll, tile width}; ms12[tile][unroll] */
int gvdb_tune_render_shape(gvdb_handle h, int shading, int chan, int buf, int* best2, float* ms12);
/* GatherDensity (arguments as gvdb_gather_density, no accumulate): threads per brick; ms2 per value */
int gvdb_tune_gather(gvdb_handle h, int subcell_size, int n, float radius, const float trans[3], int sc_len, int chan,
                     int* best_threads, float* ms2);
/* Sets the samples-ahead knob of a shading mode directly (1, 2 or 4). */
int gvdb_set_render_unroll(int shading, int unroll);
/* The 68-word scene block PrepareRender sends to the kernels (gvdb_volume_gvdb.cpp:4254-4306) */
int gvdb_get_scene_block(gvdb_handle h, int w, int hgt, float* out68);

/* Node and pool access: getNumTotalNodes/getNodeAtLevel/PoolFetch (gvdb_volume_gvdb.cpp:5799, gvdb_allocator.cpp:127-131).
 * host_dst may be NULL to query count and stride (bytes per element). */
int gvdb_get_pool(gvdb_handle h, int grp, int lev, void* host_dst, uint64_t* count, uint64_t* stride);
/* Tree facts: top level (PrepareVDB :3955-3979), bounds (ComputeBounds :1792-1816), bbox (GetBoundingBox :5404-5467).
 * out: [0] top_lev, [1] depth; outf: objmin[3], objmax[3], posmin[3], posmax[3]. */
int gvdb_get_info(gvdb_handle h, int* out2, float* outf12);
/* Atlas geometry: cnt[3], res[3] (voxels), used slots. */
int gvdb_get_atlas_info(gvdb_handle h, int chan, int* out7);
/* Atlas contents in GVDB's 3D layout (x fastest, res.x*res.y*res.z values of the channel type) */
int gvdb_read_atlas(gvdb_handle h, int chan, void* host_dst);
/* Atlas map (AtlasNode {int3 pos; int leaf}) per slot */
int gvdb_read_atlas_map(gvdb_handle h, void* host_dst);
/* InsertPointsSubcell outputs: which = 0 cnt, 1 off, 2 pos (int3), 3 nid, 4 positions (float3), 5 mapping,
 * 6 InsertPointsSubcell_FP16 positions (ushort3) */
int gvdb_read_subcells(gvdb_handle h, int which, void* host_dst, uint64_t* count);

/* Device-side handles for callers that stay on the GPU */
int gvdb_get_stream(gvdb_handle h, void** stream);
int gvdb_synchronize(gvdb_handle h);
/* Milliseconds of the last call of each stage, measured with HIP events on
 * the handle's stream: 0 topology, 1 atlas, 2 insert, 3 gather, 4 apron,
 * 5 render, 6 clear.  Synchronises the stream. */
int gvdb_stage_ms(gvdb_handle h, float* out8);
/* Kernel-only time of the most recent gather / apron / render kernel (ms). */
int gvdb_kernel_ms(gvdb_handle h, float* out4);

/* ---------------------------------------------------------------- batches
 * The batched form (no counterpart in the reference, which runs one grid per
 * call): each call runs one stage of the point path over `count` volumes --
 * grids of any size, each keeping its own pools and atlas -- with one kernel
 * launch per kernel of the stage and one host round trip per stage for the
 * whole set.  Per-grid arguments are arrays of `count` entries (orig3/trans3:
 * 3 floats per grid, may be NULL for zero).  On failure *failed (may be NULL)
 * names the grid at fault; its gvdb_last_error and gvdb_batch_last_error
 * hold the message.  Volumes run on the batch's stream from their first
 * batched call on; gvdb_release_stream returns one to its own stream. */
typedef struct gvdb_batch_s* gvdb_batch;
int gvdb_batch_create(gvdb_batch* out, int device, int max_grids);
int gvdb_batch_destroy(gvdb_batch b);
const char* gvdb_batch_last_error(gvdb_batch b);
/* RebuildTopology over many grids (gvdb_volume_gvdb.cpp:1181-1218 per grid) */
int gvdb_batch_rebuild_topology(gvdb_batch b, int count, const gvdb_handle* hs, const int* n, const float* radius,
                                const float* orig3, int* failed);
/* FinishTopology (gvdb_volume_gvdb.cpp:1579-1593 per grid) */
int gvdb_batch_finish_topology(gvdb_batch b, int count, const gvdb_handle* hs, int commit_pool, int compute_bounds,
                               int* failed);
/* UpdateAtlas (gvdb_volume_gvdb.cpp:2630-2704 per grid) */
int gvdb_batch_update_atlas(gvdb_batch b, int count, const gvdb_handle* hs, int* failed);
/* ClearChannel (gvdb_volume_gvdb.cpp:1596-1606 per grid) */
int gvdb_batch_clear_channel(gvdb_batch b, int count, const gvdb_handle* hs, int chan, int* failed);
/* InsertPointsSubcell (gvdb_volume_gvdb.cpp:4971-5092 per grid); out_len: count entries */
int gvdb_batch_insert_points_subcell(gvdb_batch b, int count, const gvdb_handle* hs, int subcell_size, const int* n,
                                     const float* radius, const float* trans3, int* out_len, int* failed);
/* GatherDensity (gvdb_volume_gvdb.cpp:5318-5344 per grid) */
int gvdb_batch_gather_density(gvdb_batch b, int count, const gvdb_handle* hs, int subcell_size, const int* sc_len,
                              const float* radius, int chan_density, int accumulate, int* failed);
/* GatherLevelSet (gvdb_volume_gvdb.cpp:5348-5371 per grid) */
int gvdb_batch_gather_level_set(gvdb_batch b, int count, const gvdb_handle* hs, int subcell_size, const int* sc_len,
                                const float* radius, int chan_levelset, int accumulate, int* failed);
/* UpdateApron (gvdb_volume_gvdb.cpp:4418-4457 per grid) */
int gvdb_batch_update_apron(gvdb_batch b, int count, const gvdb_handle* hs, int chan, float boundval, int* failed);
/* Render (gvdb_volume_gvdb.cpp:4336-4381 per grid), each into its own buffer */
int gvdb_batch_render(gvdb_batch b, int count, const gvdb_handle* hs, const int* shading, const int* chan,
                      const int* buf, int* failed);
/* Batched autotuner (no counterpart in the reference): stage 0 = gvdb_batch_gather_density (sc_len, radius, chan;
 * threads per brick, ms[2]), 1 = gvdb_batch_update_apron (chan, boundval; best = {bricks per static run, queue 0/1}, ms[5]: runs of 1, 2, 4, 8,
 * then the queue), 2 =
 * gvdb_batch_render (shading, chans, bufs; best = {unroll, tile width}, ms[12]), each timed over the whole batch. */
int gvdb_batch_tune(gvdb_batch b, int count, const gvdb_handle* hs, int stage, int subcell_size, const int* sc_len,
                    const float* radius, int chan, float boundval, const int* shading, const int* chans,
                    const int* bufs, int* best, float* ms, int* failed);
int gvdb_batch_synchronize(gvdb_batch b);
/* Device time from the first stage call after the last synchronize/elapsed to the last one (ms); synchronises. */
int gvdb_batch_elapsed_ms(gvdb_batch b, float* ms);
int gvdb_release_stream(gvdb_handle h);

/* HIP graphs of the per-frame chain (no counterpart in the reference; for callers that repeat ClearChannel ->
 * InsertPointsSubcell -> GatherDensity/GatherLevelSet -> UpdateApron -> Render over a fixed topology, e.g. new
 * point positions written into the same device buffer every frame).  Between capture_begin and capture_end the calls
 * on that grid (or batch) are recorded, not run; any other call on it fails with GVDB_ERR_STATE and fails the
 * recording.  Run the chain once without capture first: buffers that grow while recording fail the capture.
 * gvdb_graph_launch replays the recorded launches, with their recorded arguments (point count, radius, camera), over
 * the current device buffers, with no host round trip; gvdb_graph_synchronize waits for it and brings the insert's
 * entry count back to the grid.  A launch is refused once a grid's topology changed or a library buffer was
 * reallocated since the capture; a replay whose insert met points outside every brick reports GVDB_ERR_STATE at
 * synchronize.  Destroy graphs before their grids or batch. */
typedef struct gvdb_graph_s* gvdb_graph;
int gvdb_capture_begin(gvdb_handle h, gvdb_graph* out);
int gvdb_capture_end(gvdb_handle h, gvdb_graph g);
int gvdb_batch_capture_begin(gvdb_batch b, gvdb_graph* out);
int gvdb_batch_capture_end(gvdb_batch b, gvdb_graph g);
int gvdb_graph_launch(gvdb_graph g);
int gvdb_graph_synchronize(gvdb_graph g);
int gvdb_graph_destroy(gvdb_graph g);
const char* gvdb_graph_last_error(gvdb_graph g);

/* ---------------------------------------------------------------- serving
 * A serving loop over resident grids (no counterpart in the reference).
 * Grids are added with their initial points and splat radius; the server
 * owns a growing copy of each grid's points.  Requests carry an arrival time
 * in seconds from the start of gvdb_server_run:
 *   render: w x hgt RGBA8 (<= 1024 per side) into out_rgba, camera orbiting
 *           the grid's centre at `angs` (degrees) and `dist` (<= 0: 1.5x the
 *           grid's extent), with the volume's other scene settings;
 *   insert: appends n points (host_xyz must stay valid until the run ends)
 *           and rebuilds the grid (topology, atlas, insert, gather, apron).
 * gvdb_server_run mode 0 serves one request at a time in arrival order
 * through the one-grid calls; mode 1 batches continuously (see serve.h):
 * each pass admits arrived requests whose grid is idle, within max_batch and
 * the HBM budget (an insert that can never fit is refused with
 * GVDB_ERR_OOM in its status), renders as one launch overlapping the
 * inserts' batched pipeline.  latency/status get one entry per request in
 * submission order; wall = first arrival to last completion (s).
 * Returns the request index from the submit calls, the grid index from
 * gvdb_server_add_grid. */
typedef struct gvdb_server_s* gvdb_server;
int gvdb_server_create(gvdb_server* out, int device, unsigned long long hbm_budget, int max_batch);
int gvdb_server_destroy(gvdb_server s);
const char* gvdb_server_last_error(gvdb_server s);
int gvdb_server_add_grid(gvdb_server s, gvdb_handle h, const float* host_xyz, int n, float radius);
/* (re)builds every grid from its initial points (batched) */
int gvdb_server_reset(gvdb_server s);
int gvdb_server_submit_render(gvdb_server s, int grid, double arrival, int w, int hgt, int shading, const float angs[3],
                              float dist, unsigned char* out_rgba);
int gvdb_server_submit_insert(gvdb_server s, int grid, double arrival, const float* host_xyz, int n);
int gvdb_server_clear_requests(gvdb_server s);
int gvdb_server_run(gvdb_server s, int mode, double* latency, int* status, double* wall);
unsigned long long gvdb_server_device_bytes(gvdb_server s);

#ifdef __cplusplus
}
#endif
#endif /* GVDB_AMD_H */
