/*
 * pma.h -- C ABI of the MI355X pore-throat + geodesic pipeline.
 *
 * One context owns one HIP stream, one device-resident volume set and the scratch of
 * every stage.  Volumes are laid out [z][x][y]: voxel (x, y, z) is element
 * (z * X + x) * Y + y, with X = TIFF rows, Y = TIFF columns, Z = slice files, exactly the
 * reference's DataArray layout (ImageProcessing/Model/Headers/DataArray.h:85-94).
 * Calls are blocking, a context is not shared between threads, and no C++ exception
 * crosses this boundary: every entry point returns a pma_status and leaves a message for
 * pma_last_error().
 *
 * The stage entry points replace the reference's service calls, which exchange data
 * through folders of TIFF stacks; here the data stays in HBM between stages and can be
 * loaded or read back per volume (pma_load_volume / pma_read_volume) and per text file
 * (pma_set_text / pma_get_text) in the reference's own text formats.
 */
#ifndef PMA_PMA_H
#define PMA_PMA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pma_ctx pma_ctx;

typedef enum pma_status {
    PMA_OK = 0,
    PMA_ERR_INVALID_ARGUMENT = 1, /* bad pointer, size or enum value */
    PMA_ERR_HIP = 2,              /* a HIP runtime call or kernel launch failed */
    PMA_ERR_OUT_OF_MEMORY = 3,    /* device allocation failed */
    PMA_ERR_STATE = 4,            /* a stage's input volume or table has not been produced */
    PMA_ERR_CAPACITY = 5,         /* caller buffer too small; the needed size is returned */
    PMA_ERR_INTERNAL = 6          /* an internal bound was exceeded (message says which) */
} pma_status;

typedef struct pma_dims {
    int32_t x, y, z; /* image size along X (rows), Y (columns), Z (slices) */
} pma_dims;

/* Stencil of the pore, watershed and geodesic stages: "Use D3Q7 stencil" of config.txt
 * (Serializer.cpp:40-54; ModelUtilities.h:22-29). */
typedef enum pma_stencil { PMA_STENCIL_D3Q27 = 0, PMA_STENCIL_D3Q7 = 1 } pma_stencil;

/* The volumes of the reference's output folders (Constants.h:28-39). */
typedef enum pma_volume {
    PMA_VOL_SOLID = 0,      /* u8, non-zero = solid              Images/                      */
    PMA_VOL_EDT2 = 1,       /* u32 squared EDT                    EuclideanDistanceSquares/    */
    PMA_VOL_CR2 = 2,        /* u32 containing-ball radius square  ContainingBallRadiiSquares/  */
    PMA_VOL_CX = 3,         /* u32 containing-ball centre x       ContainingBallCoordinatesX/  */
    PMA_VOL_CY = 4,         /* u32 containing-ball centre y       ContainingBallCoordinatesY/  */
    PMA_VOL_CZ = 5,         /* u32 containing-ball centre z       ContainingBallCoordinatesZ/  */
    PMA_VOL_MASK = 6,       /* u8 0 solid, 127 inner, 255 centre  IsMaximumBall/               */
    PMA_VOL_PT = 7,         /* u32 pore / throat IDs              PoreThroatIds/               */
    PMA_VOL_PTW = 8,        /* u32 watershed pore / throat IDs    PoreThroatIdsWatershed/      */
    PMA_VOL_SPD = 9,        /* f32 geodesic distance              ShortestPathDistances/       */
    PMA_VOL_CAVITY = 10,    /* u32 cavity IDs                     CavityIds/                   */
    PMA_VOL_COUNT = 11
} pma_volume;

/* The reference's text outputs (Constants.h:15-25; formats of Serializer.cpp:139-221). */
typedef enum pma_text {
    PMA_TEXT_INTERMEDIATE_STATISTICS = 0,  /* intermediateStatistics.txt        */
    PMA_TEXT_EDT_COUNTS = 1,               /* euclideanDistanceSquareCounts.txt */
    PMA_TEXT_PORES = 2,                    /* pores.txt                         */
    PMA_TEXT_THROATS = 3,                  /* sharedPixelTypes.txt              */
    PMA_TEXT_PORES_WATERSHED = 4,          /* poresWatershed.txt                */
    PMA_TEXT_THROATS_WATERSHED = 5,        /* sharedPixelTypesWatershed.txt     */
    PMA_TEXT_CAVITIES = 6,                 /* cavityDescriptions.txt            */
    PMA_TEXT_COUNT = 7
} pma_text;

/* Per-stage device time of the last call of each stage, in milliseconds (HIP events on
 * the context's stream), and the counters the roofline figures are computed from. */
typedef enum pma_stage {
    PMA_STAGE_EDT = 0,
    PMA_STAGE_CONTAINING_BALLS = 1,
    PMA_STAGE_PORES = 2,
    PMA_STAGE_REMOVE_EMPTY_PORES = 3,
    PMA_STAGE_WATERSHED = 4,
    PMA_STAGE_GEODESIC = 5,
    PMA_STAGE_COUNT = 6
} pma_stage;

typedef struct pma_stage_info {
    double device_ms;     /* events around the stage on the context's stream            */
    double host_ms;       /* wall time of the entry point                                */
    uint64_t launches;    /* kernel launches issued by the stage                          */
    uint64_t work_items;  /* voxels (EDT), candidate centres (containing balls), DFS replay
                             frames (pores), folded IDs (removal), flood rounds (watershed),
                             frontier relaxations (geodesic)                                */
} pma_stage_info;

/* ---- context --------------------------------------------------------------------- */
pma_status pma_create(pma_ctx** ctx, int device);
void pma_destroy(pma_ctx* ctx);
/* Message of the last failure on this thread ("" if none). */
const char* pma_last_error(void);

/* Stencil and the reference's "Number of threads"-independent knobs.  Replaces the
 * Config fields read by Serializer::ReadConfig (Serializer.cpp:31-64) that the path uses. */
pma_status pma_set_stencil(pma_ctx* ctx, pma_stencil stencil);

/* ---- volumes ------------------------------------------------------------------------ */
/* Sets the image size (allocating nothing yet) and invalidates every held volume. */
pma_status pma_set_dims(pma_ctx* ctx, pma_dims dims);
pma_status pma_get_dims(pma_ctx* ctx, pma_dims* dims);
/* Copies a whole volume in (src_on_device != 0: src is a device pointer on the context's
 * device).  Loading a stage input lets a stage run on data produced elsewhere, like the
 * reference's -rib / -apt modes reading earlier folders (CommandLineDispatcher.cpp:91-106). */
pma_status pma_load_volume(pma_ctx* ctx, pma_volume which, const void* src, int src_on_device);
/* Copies a whole volume out; bytes must be the volume's size. */
pma_status pma_read_volume(pma_ctx* ctx, pma_volume which, void* dst, size_t bytes, int dst_on_device);
/* Device address of a held volume (valid until pma_set_dims or pma_destroy). */
pma_status pma_volume_device_ptr(pma_ctx* ctx, pma_volume which, void** dptr);

/* ---- stages (each replaces one reference service entry point) ------------------------ */
/* EuclideanDistanceComputer::ComputeEuclideanDistances (EuclideanDistanceComputer.h:42):
 * SOLID -> EDT2, max EDT square, histogram of non-boundary-crossing squares. */
pma_status pma_edt(pma_ctx* ctx);
/* ContainingBallsAssigner::SetContainingBalls (ContainingBallsAssigner.h:20):
 * EDT2 -> CR2, CX, CY, CZ, MASK and the maximum-ball count. */
pma_status pma_containing_balls(pma_ctx* ctx);
/* PoreAndThroatComputer::AssignPoreAndThroatIds (PoreAndThroatComputer.h:21):
 * MASK, EDT2, CR2 -> PT, pores.txt, sharedPixelTypes.txt. */
pma_status pma_pores(pma_ctx* ctx);
/* EmptyPoresRemover::RemoveEmptyPores (EmptyPoresRemover.h:27): PT + tables in place. */
pma_status pma_remove_empty_pores(pma_ctx* ctx);
/* WatershedComputer::AssignPoreAndThroatIds (WatershedComputer.h:22):
 * CR2 + pores.txt -> PTW, poresWatershed.txt, sharedPixelTypesWatershed.txt. */
pma_status pma_watershed(pma_ctx* ctx);
/* ShortestPathComputer::ComputeShortestPathDistances (ShortestPathComputer.h:35):
 * SOLID -> SPD, CAVITY, cavityDescriptions.txt. */
pma_status pma_geodesic(pma_ctx* ctx);
/* The reference's -all mode followed by -spd (ImageProcessingManager.cpp:63-108). */
pma_status pma_run_all(pma_ctx* ctx);

/* ---- the reference's folders ------------------------------------------------------------ */
/* Console modes (CommandLineDispatcher.cpp:80-106). */
typedef enum pma_mode {
    PMA_MODE_ALL = 0, /* -all: EDT, containing balls, pores, empty-pore removal, watershed */
    PMA_MODE_EDT = 1, /* -edt: (resampling and) the Euclidean distance transform */
    PMA_MODE_RIB = 2, /* -rib: containing balls from EuclideanDistanceSquares/ */
    PMA_MODE_APT = 3, /* -apt: pores, removal, watershed from the -rib outputs */
    PMA_MODE_SPD = 4  /* -spd: (resampling and) geodesic distances */
} pma_mode;
/* One console pass over a folder holding config.txt and Images/ (ImageProcessingManager.cpp:
 * 44-108 with Serializer.cpp:31-64, 223-281 and ImageResampler.cpp:28-206): reads the binary TIFF
 * stack (8-bit, non-zero = solid; files ordered by name), resamples it when the resolution in
 * config.txt is anisotropic, runs the mode's stages on the GPU and writes the same sub-folders
 * (one TIFF per slice, named and tagged like Images/; 32-bit volumes as 4-sample CMYK images of
 * the values' little-endian bytes, LZW) and text files as the reference.  -rib and -apt read the
 * earlier modes' folders and intermediateStatistics.txt.  Errors return a status (the reference
 * printed them and exited 0, Main.cpp:62-98). */
pma_status pma_run_folder(pma_ctx* ctx, const char* folder, pma_mode mode);
/* Reads a folder of 8-bit or 32-bit (4-sample) TIFF slices, ordered by file name, into a host
 * buffer in [z][x][y] order (DataArray.h:85-94; Image.h:189-214).  With dst == NULL it only
 * reports the size (x = rows, y = columns, z = files) and the bytes per voxel (1 or 4). */
pma_status pma_read_tiff_stack(const char* folder, void* dst, size_t bytes, pma_dims* dims, int* bytes_per_voxel);
/* Writes a host volume ([z][x][y], 1 or 4 bytes per voxel) as a folder of TIFF slices named and
 * tagged like the 8-bit slices in `like` (an Images/ folder), in the reference's output form
 * (DataArray.h:226-252, Image.h:72-83): 8-bit with the slices' photometric interpretation,
 * 32-bit as 4-sample CMYK of the little-endian bytes; LZW with the horizontal predictor. */
pma_status pma_write_tiff_stack(const char* folder, const char* like, const void* src, pma_dims dims,
                                int bytes_per_voxel);

/* ---- tuning knobs and batches of subvolumes ---------------------------------------------- */
/* No counterpart in the reference (it runs one volume per console pass).  Knobs that change how
 * the work is cut, never a result: */
typedef enum pma_tunable {
    PMA_TUNE_GEODESIC_BUCKET = 0,  /* distance width of a geodesic round (1 .. 1024; default 16)   */
    PMA_TUNE_FLOOD_GRID = 1,       /* workgroups of a watershed round (64 .. 65536; default 4096)  */
    PMA_TUNE_REPLAY_MEMORY_MB = 2  /* device memory the pore stage's DFS replays may hold, in MiB
                                      (256 .. 1048576; default 131072)                             */
} pma_tunable;
pma_status pma_set_tuning(pma_ctx* ctx, pma_tunable which, double value);
pma_status pma_get_tuning(pma_ctx* ctx, pma_tunable which, double* value);

/* One subvolume of a batch: a host solid volume ([z][x][y], non-zero = solid) and its size. */
typedef struct pma_batch_volume {
    pma_dims dims;
    const uint8_t* solid;
} pma_batch_volume;
/* How a batch runs: `workers` host threads, each with its own context and stream, take the
 * subvolumes largest first; the other fields are the knobs above (0 = default; replay memory
 * defaults to the free device memory / (2 * workers), at most 128 GiB). */
typedef struct pma_batch_config {
    uint32_t workers;
    double geodesic_bucket;
    uint32_t flood_grid;
    uint32_t replay_memory_mb;
} pma_batch_config;
/* Called on a worker's thread when subvolume `index` is done (status PMA_OK) or failed; `ctx`
 * holds that subvolume's volumes and tables until the callback returns, for pma_read_volume and
 * pma_get_text.  Calls for different subvolumes may run at the same time. */
typedef void (*pma_batch_done)(void* user, size_t index, pma_ctx* ctx, pma_status status);
/* Runs the whole pipeline (pma_run_all) on every subvolume.  Every subvolume is attempted; the
 * return value is PMA_OK or the status of the first failure (pma_last_error names its index).
 * `done` may be NULL. */
pma_status pma_run_batch(int device, const pma_batch_volume* volumes, size_t count, pma_stencil stencil,
                         const pma_batch_config* config, pma_batch_done done, void* user);
/* The autotuner: times pma_run_batch on `sample` (after one untimed pass that allocates) over
 * worker counts 1, 2, 4, 8, 12, 16, then geodesic buckets 8, 16, 32, then flood grids 1024, 4096,
 * 16384, keeping the fastest of each in turn, and writes the fastest configuration and its
 * voxels per second. */
pma_status pma_tune_batch(int device, const pma_batch_volume* sample, size_t count, pma_stencil stencil,
                          pma_batch_config* best, double* best_voxels_per_second);

/* ---- learned screening ----------------------------------------------------------------------
 * No counterpart in the reference.  A 3-D U-Net (3 levels, 16/32/64 channels, 3x3x3
 * convolutions, ~0.36 M parameters, bf16 weights and activations, fp32 accumulation on MFMA)
 * that predicts, per void voxel, the containing-ball radius and the ratio of the geodesic to the
 * Euclidean distance from the voxel's cavity start, from the solid volume and its EDT.  Its
 * fields are estimates for screening many subvolumes: they never replace a volume or table of the
 * exact pipeline, which makes the training labels and is rerun where the estimates are off. */
typedef struct pma_surrogate pma_surrogate;
/* He-normal weights from `seed`; its own stream on `device`. */
pma_status pma_surrogate_create(pma_surrogate** s, int device, uint64_t seed);
void pma_surrogate_destroy(pma_surrogate* s);
pma_status pma_surrogate_param_count(pma_surrogate* s, size_t* count);
/* Copies the fp32 weights (which = 0) or the last step's gradients (which = 1) out; layers in
 * order, each as weights [tap][out][in] then bias [out], with the padded channels. */
pma_status pma_surrogate_params(pma_surrogate* s, int which, float* dst, size_t count);
pma_status pma_surrogate_set_params(pma_surrogate* s, const float* src, size_t count);
/* One step on the volume of `ctx`, which must hold SOLID, EDT2, CR2, SPD and CAVITY (pma_edt,
 * pma_containing_balls and pma_geodesic run), every side a multiple of 4: forward pass, mean
 * squared error against the exact fields (*loss), backward pass; with update != 0 an Adam step. */
pma_status pma_surrogate_train_step(pma_surrogate* s, pma_ctx* ctx, float learning_rate, int update, float* loss);
/* Predicted radius (voxels) and distance ratio of every voxel of `ctx` (0 on solid voxels), from
 * SOLID and EDT2 only; host buffers of one float per voxel, or device buffers. */
pma_status pma_surrogate_predict(pma_surrogate* s, pma_ctx* ctx, float* radius, float* ratio, int dst_on_device);

/* ---- allocation-free geodesic ---------------------------------------------------------- */
/* One row of cavityDescriptions.txt (ShortestPathComputer.cpp:143-164, Serializer.cpp:205-221). */
typedef struct pma_cavity {
    uint32_t id;
    int32_t x, y, z;           /* start voxel */
    double distance_to_center; /* to the image centre voxel */
    uint64_t void_count;       /* void voxels reached */
    int32_t reached_walls;     /* distinct boundary planes reached */
    int32_t reaches_all;       /* 1 if every boundary plane was reached */
} pma_cavity;
/* Bytes of the static per-device scratch of pma_geodesic_noalloc: room for up to
 * (48 MiB - 0.2 MiB) / 20 B = 2.5 M blocks of 8^3 voxels (a 1024^3 volume has 2 097 152). */
#define PMA_GEODESIC_NOALLOC_SCRATCH_BYTES (48u << 20)
/* No counterpart in the reference: pma_geodesic for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid it).  It
 * computes the same distances, cavity ids and cavity rows as pma_geodesic, from and into the
 * caller's device buffers (solid u8, dist f32, cavity u32, [z][x][y] layout), on `stream` (a
 * hipStream_t; NULL = the default stream), and blocks until done.  It calls no hipMalloc or
 * hipFree: its scratch is one static device array per GPU of
 * PMA_GEODESIC_NOALLOC_SCRATCH_BYTES, held until the process ends.  Volumes whose scratch would
 * not fit are rejected with PMA_ERR_CAPACITY before any device work; more than `cap` cavities
 * end the call with PMA_ERR_CAPACITY (*count rows written).  Calls that overlap on one GPU share
 * that array and race: serialise them.  Makes `device` current. */
pma_status pma_geodesic_noalloc(int device, const uint8_t* solid, pma_dims dims, pma_stencil stencil, float* dist,
                                uint32_t* cavity, pma_cavity* cavities, size_t cap, size_t* count, void* stream);

/* ---- statistics and tables ------------------------------------------------------------ */
pma_status pma_get_statistics(pma_ctx* ctx, uint32_t* max_edt2, uint64_t* maximum_balls_count);
/* counts[0..maxE] of non-boundary-crossing EDT squares; *len receives maxE + 1. */
pma_status pma_get_histogram(pma_ctx* ctx, uint64_t* dst, size_t cap, size_t* len);
/* Text of one of the reference's files, NUL-terminated; *len receives strlen + 1. */
pma_status pma_get_text(pma_ctx* ctx, pma_text which, char* dst, size_t cap, size_t* len);
/* Replaces a held table by parsing the reference's text (pores.txt, sharedPixelTypes.txt,
 * intermediateStatistics.txt), as the reference's later stages read earlier files. */
pma_status pma_set_text(pma_ctx* ctx, pma_text which, const char* text);
pma_status pma_get_stage_info(pma_ctx* ctx, pma_stage stage, pma_stage_info* info);
/* Per-kernel device time: while enabled, every launch of an instrumented kernel is bracketed
 * by HIP events on the context's stream; totals accumulate until reset (enabled = 0 resets).
 * For roofline figures (achieved bytes / average launch duration). */
pma_status pma_set_kernel_timing(pma_ctx* ctx, int enabled);
pma_status pma_get_kernel_time(pma_ctx* ctx, const char* kernel, double* total_ms, uint64_t* launches);
/* Names of the instrumented kernels, separated by '\n' (NUL-terminated; *len = strlen + 1). */
pma_status pma_list_kernel_times(pma_ctx* ctx, char* dst, size_t cap, size_t* len);

/* ---- synthetic inputs (written into SOLID; 0 = void, 255 = solid) --------------------- */
/* Union of n spheres with centres uniform in the box and radii uniform in [rmin, rmax]. */
pma_status pma_generate_spheres(pma_ctx* ctx, int32_t n, float rmin, float rmax, uint64_t seed);
/* i.i.d. N(0,1) noise, separable Gaussian blur (sigma, truncated at 4 sigma, reflected
 * boundary), void = the lowest `porosity` fraction of the field. */
pma_status pma_generate_gaussian(pma_ctx* ctx, float sigma, float porosity, uint64_t seed);

#ifdef __cplusplus
}
#endif

#endif /* PMA_PMA_H */
