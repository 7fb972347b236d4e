/*
 * mapx.h - C ABI of the MI355X map-algebra backend (libmapx.so).
 *
 * The ma_* entry points are the drop-in boundary: the same names, argument
 * order, by-value struct layouts and enum values as the reference's C ABI
 * (front/bindings.hpp:21-94), which python/map.py binds with ctypes
 * (python/map.py:723-830). Each declaration cites the reference function it
 * replaces. Entry points marked "extension" have no reference counterpart.
 *
 * Differences in error behaviour: the reference asserts and aborts on every
 * error (bindings.cpp, OclEnv.cpp:1264-1271). Here a failing call returns
 * NULL (node factories), a non-zero int, or an empty variant (dtype 0), and
 * ma_lastError() describes the failure; the process keeps running.
 */
#ifndef MAPX_H_
#define MAPX_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef MAPX_NODE_TYPE
#define MAPX_NODE_TYPE struct ma_node
#endif
typedef MAPX_NODE_TYPE ma_node_t; /* opaque DAG node, owned by the runtime */

/* VariantType: 8-byte value union + dtype tag (python/map.py:91-109). */
typedef struct {
  union {
    unsigned char b8;
    signed char s8;
    short s16;
    int s32;
    long long s64;
    unsigned char u8;
    unsigned short u16;
    unsigned int u32;
    unsigned long long u64;
    float f32;
    double f64;
  } u;
  int dtype;
} ma_variant_t;

/* Array4<int> with N=2: {int arr[2]; int n} (util/Array4.hpp:37-44; python/map.py:157-161). */
typedef struct {
  int arr[2];
  int n;
} ma_dim2_t;

/* Enum values: python/map.py:35-83 (DataTypeEnum F32=1 ... S64=11, D0=2, D2=6,
 * ROW+BLK=3, DEV_GPU=4, UnaryEnum, BinaryEnum, ReductionEnum). */
typedef int ma_datatype_t;
typedef int ma_numdim_t;
typedef int ma_memorder_t;
typedef int ma_devicetype_t;
typedef int ma_unary_t;
typedef int ma_binary_t;
typedef int ma_reduction_t;
typedef int ma_percent_t;    /* PercentEnum (util/PercentType.hpp:18): AGE = 1, ILE = 2 */
typedef int ma_diversity_t;  /* DiversityEnum (util/DiversityType.hpp:19-21): VARI = 1, MAJO, MINO, MEAN */

/* ---- devices and evaluation ------------------------------------------------ */

/* front/bindings.hpp:27 ma_setupDevices. Selects the GPU (MAPX_DEVICE, default 0). */
void ma_setupDevices(const char *plat_name, ma_devicetype_t dev, const char *dev_name);
/* front/bindings.hpp:29-30 */
void ma_increaseRef(ma_node_t *node);
void ma_decreaseRef(ma_node_t *node);
/* front/bindings.hpp:32 ma_eval: evaluates the nodes (fuse, compile, run). */
void ma_eval(ma_node_t **vec, int num);
/* front/bindings.hpp:33 ma_value: value of an evaluated D0 node. */
ma_variant_t ma_value(ma_node_t *node);
/* front/bindings.hpp:36-42 node attributes. */
int ma_nodeid(ma_node_t *node);
ma_datatype_t ma_datatype(ma_node_t *node);
ma_numdim_t ma_numdim(ma_node_t *node);
ma_memorder_t ma_memorder(ma_node_t *node);
ma_dim2_t ma_datasize(ma_node_t *node);
ma_dim2_t ma_blocksize(ma_node_t *node);

/* ---- I/O --------------------------------------------------------------------- */

/* front/bindings.hpp:48 ma_read: a raster file by extension (file/File.cpp:25-45): .bin ROW+BLK or
 * tiled .tif/.tiff (TIFF or BigTIFF, one uncompressed tile per block); lazy. */
ma_node_t *ma_read(const char *file_path);
/* front/bindings.hpp:49 ma_write: evaluates and writes a .bin or tiled BigTIFF (.tif/.tiff) file;
 * 0 on success. */
int ma_write(ma_node_t *node, const char *file_path);

/* ---- operations -------------------------------------------------------------- */

/* front/bindings.hpp:55-60 */
ma_node_t *ma_constant(ma_variant_t var, ma_dim2_t ds, ma_datatype_t dt, ma_memorder_t mo, ma_dim2_t bs);
ma_node_t *ma_cast(ma_node_t *node, ma_datatype_t type);
ma_node_t *ma_index(ma_node_t *node, ma_numdim_t dim);
ma_node_t *ma_conditional(ma_node_t *cond, ma_node_t *lhs, ma_node_t *rhs);
/* front/bindings.hpp:62-63 */
ma_node_t *ma_unary(ma_node_t *node, ma_unary_t type);
ma_node_t *ma_binary(ma_node_t *lhs, ma_node_t *rhs, ma_binary_t type);
/* front/bindings.hpp:69 ma_neighbor */
ma_node_t *ma_neighbor(ma_node_t *node, ma_dim2_t coord);
/* front/bindings.hpp:72 ma_convolution: mask_s values, mask_ds = {ncols, nrows}, row-major. */
ma_node_t *ma_convolution(ma_node_t *node, ma_variant_t *mask_v, int mask_s, ma_dim2_t mask_ds);
/* front/bindings.hpp:73 ma_focalFunc */
ma_node_t *ma_focalFunc(ma_node_t *node, ma_variant_t *mask_v, int mask_s, ma_dim2_t mask_ds, ma_reduction_t type);
/* front/bindings.hpp:73 ma_focalPercent (front/bindings.cpp:203-207): per cell, the window's count of
 * cells equal to the centre (AGE) or below it (ILE), in the input's type, divided by the reference's
 * neighbourhood size (2*H0+1)*H1*2+1 (the mask's values are not used, only its size). */
ma_node_t *ma_focalPercent(ma_node_t *node, ma_variant_t *mask_v, int mask_s, ma_dim2_t mask_ds, ma_percent_t type);
/* front/bindings.hpp:63 ma_diversity (front/bindings.cpp:145-151): per cell over 2..16 rasters of one
 * shape and type: distinct values (VARI, S8), most / least frequent value (MAJO / MINO, first on
 * ties), or MEAN (sum of values over the number of distinct values). */
ma_node_t *ma_diversity(ma_node_t **vec, int size, ma_diversity_t type);
/* front/bindings.hpp:76 ma_radialScan (front/bindings.cpp:218-221): radial scan from the cell
 * coord = {x, y}: each cell gets type (SUM/PROD/MAX/MIN) of its input and the value its ray
 * from coord carries in (interpolated between the two cells of the previous column); F32/F64 only. */
ma_node_t *ma_radialScan(ma_node_t *node, ma_reduction_t type, ma_dim2_t coord);
/* front/bindings.hpp:75 ma_zonalReduc: whole-raster reduction to a D0 value. */
ma_node_t *ma_zonalReduc(ma_node_t *node, ma_reduction_t type);
/* front/bindings.hpp:70 ma_spreadNeighbor (front/bindings.cpp:177-180): per cell, the reduction
 * (from its neutral value) of the 3x3 neighbours whose D8 code in `dir` (1 E, 2 SE, 4 S, 8 SW,
 * 16 W, 32 NW, 64 N, 128 NE) points at the cell, y outer / x inner, with the focal border rule;
 * the input's type. */
ma_node_t *ma_spreadNeighbor(ma_node_t *node, ma_node_t *dir, ma_reduction_t type);
/* front/bindings.hpp:77 ma_spreadScan (front/bindings.cpp:225-228): accumulation of `node` along
 * the U8 D8 directions `dir` (SUM only): each cell gets the total that leaves it towards its
 * downstream cell (its value plus all its upstream totals); cells with no downstream cell in
 * the raster or on a direction cycle get 0. */
ma_node_t *ma_spreadScan(ma_node_t *node, ma_node_t *dir, ma_reduction_t type);
/* front/bindings.hpp:79 ma_stats (front/bindings.cpp:232-235): the input's values, with its
 * maximum and minimum computed in the same evaluation (read them with ma_statsRange). */
ma_node_t *ma_stats(ma_node_t *node);
/* front/bindings.hpp:80 ma_barrier (front/bindings.cpp:237-240): the input's values; the
 * groups before and after it are never fused. */
ma_node_t *ma_barrier(ma_node_t *node);
/* front/bindings.hpp:65 ma_access (front/bindings.cpp:155-158): the value of cell coord = {x, y}
 * as a D0 node. */
ma_node_t *ma_access(ma_node_t *node, ma_dim2_t coord);
/* front/bindings.hpp:66 ma_lhsAccess (front/bindings.cpp:160-163): `lhs` with cell coord = {x, y}
 * replaced by the D0 value `rhs` (of lhs's type). */
ma_node_t *ma_lhsAccess(ma_node_t *lhs, ma_node_t *rhs, ma_dim2_t coord);

/* front/bindings.hpp:56 ma_rand (front/bindings.cpp:113-116, runtime/dag/Rand.cpp): a raster of
 * the seed raster's shape (type dt, or the seed's when dt = NONE_DATATYPE) whose cells are
 * Philox2x32-10 outputs (Random123) with counter {seed cell, 0} and key {cell index}. */
ma_node_t *ma_rand(ma_node_t *seed, ma_datatype_t dt, ma_memorder_t mo);
/* front/bindings.hpp:69 ma_boundedNbh (front/bindings.cpp:172-175): declared so that callers
 * bind; the reference generates no kernel statement for it, and here the call fails (NULL,
 * ma_lastError). */
ma_node_t *ma_boundedNbh(ma_node_t *node, ma_node_t *cx, ma_node_t *cy);
/* front/bindings.hpp:86-92 symbolic-loop digestion (front/bindings.cpp:244-275,
 * Runtime.cpp:156-311), used by python/map.py's Python 2 __nonzero__. Declared so that
 * callers bind; here a `while` over a D0 raster runs as ordinary Python 3 (the D0
 * condition is evaluated on the device each iteration), and these calls fail: the void
 * ones leave ma_lastError set, ma_loopAssemble returns NULL, ma_loopAgainTail sets *num to 0. */
void ma_loopStart(void);
void ma_loopCond(ma_node_t *cond);
void ma_loopBody(void);
void ma_loopAgain(void);
ma_node_t *ma_loopAssemble(void);
void ma_loopEnd(void);
void ma_loopAgainTail(ma_node_t *loop, ma_node_t ***agains, ma_node_t ***tails, int *num);

/* ---- extensions (no reference counterpart) ---------------------------------- */

/* Description of the last failure of a ma_* call on this thread ("" if none). */
const char *ma_lastError(void);
/* Per-zone statistics of `data` over the integer `zone` raster, zones 0..nzones-1:
 * a D1 node of nzones entries; after evaluation ma_zonalTable returns them. */
ma_node_t *ma_zonalStats(ma_node_t *data, ma_node_t *zone, int nzones);
/* Copies an evaluated per-zone table: FP64 sums, min, max (as double) and counts. */
int ma_zonalTable(ma_node_t *node, double *sum, double *mn, double *mx, long long *count, int nzones);
/* In-memory input raster: `host` holds the block-tiled payload (the .bin
 * layout) of nbytes; the runtime copies it. */
ma_node_t *ma_input(const void *host, size_t nbytes, ma_dim2_t ds, ma_datatype_t dt, ma_dim2_t bs);
/* Copies the resident raster of an evaluated D2 node (block-tiled) to host. */
int ma_fetch(ma_node_t *node, void *host, size_t nbytes);
/* Host-only file access (no device needed): reads the block-tiled payload of a .bin or tiled
 * .tif/.tiff raster into `host` (nbytes = the whole payload; ma_read gives its sizes), or writes
 * one as a .bin or tiled BigTIFF file. */
int ma_fileLoad(const char *file_path, void *host, size_t nbytes);
int ma_fileStore(const char *file_path, const void *host, size_t nbytes, ma_dim2_t ds, ma_datatype_t dt,
                 ma_dim2_t bs);
/* Releases the resident raster of a node (it is recomputed if evaluated again). */
int ma_release(ma_node_t *node);
/* Runtime option: "form" (0 plain: one launch per group per block, 1 batched),
 * "graph" (replay launch sequences from hipGraphs), "focal_zonal" (fuse a focal
 * group into the zonal group that consumes it), "tile_tx"/"tile_ty"/"tile_vx",
 * "wg_per_cu", "profile" (time every launch), "batch_tasks" (independent groups
 * of one kernel in one evaluation share one launch through a device task
 * table), "queue" (device chunk queue: 0 batched launches only, 1 always,
 * 2 never), "chunk_tiles" (tiles a work-group takes at a time in groups
 * without reductions; 0: 16), "tile_pf" (focal groups load the next tile
 * while computing one), "tile_st" (focal groups whose halo inputs are rasters
 * hold their rows in registers instead of staging tiles in LDS; default 0),
 * "libm" (FP32 sqrt, atan2 and sin/cos pairs call the device library's
 * functions instead of the fast paths, which are within 4 ulp; default 0).
 * Returns 0 on success. */
int ma_setOption(const char *key, long long value);
long long ma_getOption(const char *key);
/* Fusion groups and generated kernel sources of an evaluation, as JSON, without
 * running it (valid until the next call on this thread). */
const char *ma_plan(ma_node_t **vec, int num);
/* Compiles every kernel of the plan with hiprtc for gfx950 (no device needed). */
int ma_compilePlan(ma_node_t **vec, int num);
/* Counters and device-memory figures of the runtime, as JSON. */
const char *ma_runStats(void);
/* Maximum and minimum of an evaluated Stats node (ma_stats), in its data type. */
int ma_statsRange(ma_node_t *node, ma_variant_t *max, ma_variant_t *min);
/* Synthetic DEM generated in HBM: ng Gaussians given as rows (cx, cy, sigma,
 * amplitude) plus value-noise fBm with the given seed; F32, block-tiled. */
ma_node_t *ma_genDem(ma_dim2_t ds, ma_dim2_t bs, const double *gauss, int ng, unsigned long long noise_seed);
/* ma_genDem into caller memory, allocating nothing (no counterpart in the reference; for callers
 * that cannot allocate device memory: code under graph capture, serving loops, harnesses that
 * forbid hipMalloc). `out` is device memory of at least out_bytes >= the block-tiled raster
 * (ceil(ds/bs) blocks of bs cells, 4 bytes each); at most 256 Gaussians, whose table goes through
 * a 10240-byte device array of the library that each GPU holds until the process ends. Calls
 * that overlap on one GPU race on that array. Returns once the raster is written; 0 on success. */
int ma_genDemNoAlloc(void *out, size_t out_bytes, ma_dim2_t ds, ma_dim2_t bs, const double *gauss, int ng,
                     unsigned long long noise_seed);
/* Voronoi zone raster of np integer points (px, py), type U8/U16/S32/U32. */
ma_node_t *ma_genZones(ma_dim2_t ds, ma_dim2_t bs, ma_datatype_t dt, const int *px, const int *py, int np);
/* Device time of the last evaluation in milliseconds (HIP events). */
double ma_lastEvalMs(void);
/* Durations of the kernels launched since the last call while the "profile"
 * option is set (HIP events around each launch), as JSON by kernel name. */
const char *ma_kernelTimes(void);
/* ---- serving (extension) ----
 * Requests (sets of target nodes) are evaluated as they arrive: admitted while
 * their HBM footprint fits hbm_budget bytes (0: 80% of free HBM) and fewer than
 * max_inflight run; each step enqueues the next level of every admitted
 * request, batching the fused groups of all requests by kernel, and retires
 * the requests the GPU has finished (waiting up to wait_us for one). While a
 * server is open, ma_eval is refused. */
int ma_serveOpen(long long hbm_budget, int max_inflight);
/* Queues a request; returns its id (>= 0) or -1. */
long long ma_serveSubmit(ma_node_t **targets, int num);
/* One step; returns the number of requests retired in it, or -1. */
int ma_serveStep(int wait_us);
/* Copies up to cap retired requests (oldest first): id and steady-clock
 * (CLOCK_MONOTONIC) seconds of submission, admission and completion. */
int ma_serveTake(long long *ids, double *t_submit, double *t_admit, double *t_done, int cap);
/* Server counters as JSON. */
const char *ma_serveStats(void);
/* Runs every submitted request to completion and closes the server. */
int ma_serveClose(void);
/* ---- resident cache (extension) ----
 * Rasters kept in HBM under caller keys (e.g. DEM ids shared by the requests of a server):
 * the cache holds a reference on each node; past the byte budget (0: no limit) the least
 * recently used entries are released, and their rasters leave HBM once no request or caller
 * holds them. ma_cacheGet returns NULL on a miss (not an error); a hit becomes the most recent. */
int ma_cacheBudget(long long bytes);
ma_node_t *ma_cacheGet(long long key);
int ma_cachePut(long long key, ma_node_t *node);
int ma_cacheDrop(long long key);
const char *ma_cacheStats(void);
/* Generated HIP source of a kernel this process has built (by kernel name), or NULL. */
const char *ma_kernelSource(const char *name);
/* Autotuner: evaluates the nodes once per candidate tile shape (width,
 * height, cells per thread: also the layout of the LDS halo tile), times each
 * launch with HIP events, and keeps the fastest shape per fused-group
 * signature and block size for later evaluations. Returns a JSON report. */
const char *ma_autotune(ma_node_t **vec, int num, int reps);
/* Loads (save = 0) or writes (save = 1) the tuning results; returns their count. */
int ma_tuneFile(const char *path, int save);
/* Fusion legality of a pattern pair (PatternEnum bit sets): bit 0 pipe, bit 1
 * flat (runtime/Pattern.cpp:53-182 as applied by the fuser). */
int ma_fusionRules(int top_or_left, int bot_or_right);
/* Accuracy of the FP32 fast paths of generated kernels against FP64 on the device: fn "sqrt",
 * "sin", "cos", "sincos" (every FP32 bit pattern in [lo, hi), hi <= 2^32), "atan2" / "atan2_near"
 * (each pattern x paired with a y of random bits, or of x's sign and exponent). Gives the largest
 * error in ulps (of the FP32 value nearest the FP64 result), the count of results above 4 ulp,
 * the count checked and the input bits of the largest error; 0 on success. */
int ma_mathCheck(const char *fn, unsigned long long lo, unsigned long long hi, double *max_ulp,
                 unsigned long long *n_over, unsigned long long *n, unsigned *worst_in);

#ifdef __cplusplus
}
#endif

#endif /* MAPX_H_ */
