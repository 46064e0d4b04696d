/*
 * fullwave_hip.h -- C ABI of the MI355X Fullwave 2.5 time-stepping engine.
 *
 * The reference runs its engine as a separate executable: Launcher.run() chdirs into a
 * directory of raw .dat inputs, runs the binary and reads genout.dat back
 * (fullwave/solver/launcher.py:160-254, called from fullwave/solver/solver.py:756-759).
 * This library replaces that process boundary in two ways:
 *
 *   - fw_run_dir():  same contract as the executable -- reads the input files the reference's
 *                    InputFileWriter writes (fullwave/solver/input_file_writer.py:563-881) and
 *                    writes genout.dat (read by fullwave/solver/utils.py:12-45);
 *   - fw_sim_*():    the same inputs handed over in memory, one handle per simulation.
 *
 * Conventions: plain C types only; every function returns an fw_status (0 = FW_OK) and never
 * throws; on failure fw_last_error() returns a message for the calling thread. Host buffers
 * belong to the caller and are only read during the call. Arrays are C order, the last axis
 * fastest; coordinate lists are int32 [n, ndim] as in icc.dat / outc.dat / icczero.dat.
 * Calls on one handle must not overlap; different handles may run from different threads.
 */
#ifndef FULLWAVE_HIP_H
#define FULLWAVE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FW_ABI_VERSION 4

typedef enum fw_status {
    FW_OK = 0,
    FW_ERR_INVALID = 1,     /* bad argument, missing input, inconsistent sizes */
    FW_ERR_DEVICE = 2,      /* HIP runtime error */
    FW_ERR_NOMEM = 3,       /* device allocation failed */
    FW_ERR_IO = 4,          /* file missing or short (fw_run_dir) */
    FW_ERR_STATE = 5,       /* call not valid in the handle's current state */
    FW_ERR_UNSUPPORTED = 6  /* valid request this build does not implement */
} fw_status;

/* Simulation description: the scalar .dat files of the contract
 * (fullwave/solver/input_file_writer.py:764-813). */
typedef struct fw_desc {
    int32_t abi_version; /* FW_ABI_VERSION */
    int32_t ndim;        /* 2 or 3 */
    int32_t isotropic;   /* 1: kappau/kappax, apml/bpml{u,x}{nu}; 0: per-axis maps */
    int32_t n_relax;     /* relaxation mechanisms; 2 */
    int64_t nx, ny, nz;  /* grid (nX, nY, nZ); nz = 1 in 2D */
    int64_t nt;          /* time steps of a full run (nT) */
    double dx, dy, dz, dt, c0;
    int64_t ncoords;     /* source points (ncoords) */
    int64_t ntic;        /* samples per source signal (nTic) */
    int64_t ncoordsout;  /* sensors (ncoordsout) */
    int64_t modt;        /* record every modt steps (modT) */
    int64_t ncoordszero; /* pressure-release points (ncoordszero) */
    int64_t ndmap;       /* coefficient sets in dmap (ndmap) */
    /* 0: multi-relaxation power-law attenuation (kappa, apml/bpml maps);
     * 1: exponential attenuation, one per-step decay factor per voxel ("a_exp", the reference's
     *    fullwave2_*_exponential_attenuation engines, fullwave/solver/solver.py:193-218) */
    int32_t attenuation_model;
    /* Transmits on this one medium advanced together (0 or 1: a single simulation). Every
     * wave field, source matrix and record buffer holds n_shots copies; the medium maps are
     * stored once and each launch covers every shot. The reference runs one transmit per
     * engine call (fullwave/solver/solver.py:620-745, is_static_map reusing the maps on
     * disk); this is that loop in one handle. */
    int32_t n_shots;
} fw_desc;

typedef struct fw_sim fw_sim;

/* ---- library ------------------------------------------------------------------------ */
int fw_abi_version(void);
const char* fw_version_string(void);
/* Replaces the reference's CUDA device probing (fullwave/solver/cuda_utils.py:317-395,
 * fullwave/solver/launcher.py:107-139). */
int fw_device_count(int* count);
/* Message of the last failed call on this thread ("" if none). */
const char* fw_last_error(void);
/* Device bandwidth self-check; it has no counterpart in the reference. Copies `bytes` bytes
 * (rounded down to 16) between two device buffers with a plain vectorized copy kernel
 * `repeats` times and returns the achieved read + write rate in GB/s. Used to calibrate
 * hardware-counter byte counts against a kernel of known traffic. */
int fw_device_copy_bandwidth(int device, size_t bytes, int repeats, double* gbs);
/* Streaming self-check; no counterpart in the reference. Reads `nread` (0..16) and writes
 * `nwrite` (1..16) device arrays of `cells` floats, one float per lane at the same index in
 * every array (the per-cell operand pattern of the stencil kernels, without their stencils),
 * `repeats` times, and returns the achieved read + write rate in GB/s: the rate a kernel with
 * that many per-cell streams can reach on this device. */
int fw_device_stream_bandwidth(int device, int64_t cells, int32_t nread, int32_t nwrite, int32_t repeats,
                               double* gbs);
/* Free and total device memory in bytes (hipMemGetInfo); no counterpart in the reference.
 * Used to admit jobs by memory budget. */
int fw_device_memory(int device, size_t* free_bytes, size_t* total_bytes);

/* ---- directory contract: drop-in for running the engine executable ------------------- */
/* Reads every input of sim_dir, runs nT steps on `device` and writes sim_dir/genout.dat
 * (float32 [n_out, ncoordsout], n_out = ceil(nT / modT)). Replaces
 * fullwave/solver/launcher.py:196-215. */
int fw_run_dir(const char* sim_dir, int device);

/* Device bytes that fw_run_dir_noalloc works in: one static device array per GPU, held from the
 * first such call on that GPU until the process ends. */
#define FW_NOALLOC_SCRATCH_BYTES (256ull << 20)
/* fw_run_dir without any device allocation. It has no counterpart in the reference; it is for
 * callers that cannot allocate device memory (a serving loop with a fixed device footprint, code
 * that must not call hipMalloc or hipFree). Same files, same genout.dat, same kernels as
 * fw_run_dir, with two differences: the kernel shapes are FULLWAVE_HIP_KERNELS's or the defaults
 * (not timed), and the relaxation coefficients are read from the per-cell maps (no coding; the
 * results are the same). Every device array of the run (input maps, wave fields, memory
 * variables, records, index lists, coefficient table, staging) is placed in the static array;
 * nothing is allocated on the device, no pinned host memory, stream or event is created, and the
 * run uses the calling thread's per-thread stream. Host memory for the files is allocated as
 * usual. A run that needs more than FW_NOALLOC_SCRATCH_BYTES (2D isotropic: about 2.5 M cells)
 * fails with FW_ERR_UNSUPPORTED before any device work. Each call clears the wave state and the
 * records it uses; calls on one GPU share the array, so this process serializes them (a second
 * call waits for the first), while another process has its own array. */
int fw_run_dir_noalloc(const char* sim_dir, int device);

/* ---- in-memory handle ------------------------------------------------------------------ */
int fw_sim_create(const fw_desc* desc, int device, fw_sim** out);
/* Upload one input by its file stem: "K", "rho", "c", "beta", "dcmap" (int32 per voxel),
 * "dmap" (float32 [9, 2, ndmap]), "icc", "outc", "icczero" (int32 [n, ndim]),
 * "icmat" (float32 [n_shots, ncoords, ntic]; may be uploaded again between runs, e.g. after
 * fw_sim_reset for the next transmits on the same medium), and the relaxation maps "kappau", "kappax",
 * "apmlu1", "bpmlu1", ... (isotropic) or additionally "kappaw", "kappay", "kappav", "kappaz",
 * "apmlw1", ... (anisotropic), or "a_exp" (exponential model). Per-voxel maps are float32
 * [nx, ny(, nz)]. `bytes` must match. */
int fw_sim_upload(fw_sim* sim, const char* name, const void* host, size_t bytes);
/* Advance n_steps time steps (synchronous). The first call checks that every input is there. */
int fw_sim_run(fw_sim* sim, int64_t n_steps);
/* Zero the wave state, memory variables, recorded traces and the step counter. */
int fw_sim_reset(fw_sim* sim);
int fw_sim_step_count(const fw_sim* sim, int64_t* steps);
/* Number of records written so far (ceil(steps / modt)). */
int fw_sim_record_count(const fw_sim* sim, int64_t* n_records);
/* Recorded traces, float32 [n_shots, n_records, ncoordsout], each shot time-major (the
 * genout.dat layout). */
int fw_sim_download_sensors(fw_sim* sim, float* out, size_t bytes);
/* Current wave field by name: "p", "vx", "vy", "vz"; float32 [nx, ny(, nz)] of the first
 * shot, or [n_shots, nx, ny(, nz)] when `bytes` covers every shot. */
int fw_sim_download_field(fw_sim* sim, const char* name, float* out, size_t bytes);
/* Overwrite a wave field by name ("p", "vx", "vy", "vz") from float32 [nx, ny(, nz)] (first
 * shot) or [n_shots, nx, ny(, nz)]; ghost cells are left as given. Used to start from an
 * initial condition or a checkpoint. */
int fw_sim_set_field(fw_sim* sim, const char* name, const float* values, size_t bytes);
/* Kernel family and shape the handle runs for `kind` (0 = velocity kernel, 1 = pressure
 * kernel). It has no counterpart in the reference, whose binary fixes its kernels.
 * variant 0 = per-cell kernels, 1 = x-marching kernels with
 * `width` lanes per workgroup along the contiguous axis and `chunk` planes per workgroup,
 * 2 = 2D tiled kernels with 4-wide vector access on tiles of `chunk` rows x 256 columns,
 * 3 / 4 = 3D software-pipelined x-marching kernels (workgroups of `width` = 512 or 1024 lanes;
 * per-cell operands loaded one iteration ahead / at the top of their iteration).
 * Chosen on the first run by timing the candidates on the handle's own arrays, separately for
 * the two kernels (or forced for both with
 * the FULLWAVE_HIP_KERNELS environment variable: "percell", "stream[:width[:chunk]]",
 * "tile[:rows]", "pipe[:width[:chunk]]" or "pipelite[:width[:chunk]]", or
 * "<velocity>/<pressure>" to set each on its own).
 * Calling this before the first run makes that choice. */
int fw_sim_kernel_config(fw_sim* sim, int32_t kind, int32_t* variant, int32_t* width, int32_t* chunk);
/* Fix the kernel family and shape of `kind` (values as fw_sim_kernel_config reports them)
 * before the first run, instead of timing the candidates; it overrides FULLWAVE_HIP_KERNELS.
 * It has no counterpart in the reference. The first run fails with FW_ERR_UNSUPPORTED or
 * FW_ERR_INVALID if the shape cannot run this grid. Lets a caller reuse a configuration tuned
 * once for many handles of one grid shape (a server), and keeps results independent of
 * timing noise. */
int fw_sim_set_kernel_config(fw_sim* sim, int32_t kind, int32_t variant, int32_t width, int32_t chunk);
/* Run steady steps from a captured HIP graph of `steps` time steps (0: plain launches, the
 * default). It has no counterpart in the reference. fw_sim_run then replays the graph for
 * whole multiples of `steps` and launches the remainder one step at a time; results are
 * identical. The step number the source/sensor pass uses is then kept in device memory. Cuts
 * the per-launch host cost, which dominates for small grids. fw_sim_run_timed never uses it. */
int fw_sim_set_graph_steps(fw_sim* sim, int64_t steps);
/* Coded relaxation coefficients (isotropic relaxation model). No counterpart in the reference,
 * which reads its ten coefficient maps (kappau/x, apml/bpml u/x 1-2) and dcmap for every cell
 * and step. On the first run the handle counts the distinct tuples of those eleven maps on the
 * device and gives the (at most 65535) most frequent a 16-bit code; the x-marching and tiled
 * kernels then read a code per cell and that code's row of a small table instead of the maps
 * (2 bytes per cell and kernel instead of 24). Cells of the remaining tuples (PML edges and
 * corners of 3D grids) read the maps as before. The values read are the maps' own, bit for
 * bit, so results do not change. Codes are not used when they would cover less than half of
 * the cells. fw_sim_set_relaxation_codes(sim, 0) keeps the per-cell maps (as does
 * FULLWAVE_HIP_CODES=0 in the environment); called after the first run it re-codes at once.
 * Uploading a relaxation map or dcmap after the first run re-codes too.
 * fw_sim_relaxation_codes returns the number of codes in use (0: per-cell maps) and, if
 * coded_fraction is not NULL, the fraction of device cells that carry a code (3D device rows
 * are padded to 32 cells); it runs the first-run set-up if it has not happened yet. */
int fw_sim_set_relaxation_codes(fw_sim* sim, int32_t enable);
int fw_sim_relaxation_codes(fw_sim* sim, int32_t* n_codes, double* coded_fraction);
/* Run n_steps with a HIP event pair around every launch on the handle's stream and add the
 * summed milliseconds per kernel kind to ms[0..n_kinds) (kinds: 0 velocity, 1 pressure,
 * 2 source/air/sensor); also returns the number of kinds in *n_kinds. */
int fw_sim_run_timed(fw_sim* sim, int64_t n_steps, double* ms, int32_t* n_kinds);
int fw_sim_destroy(fw_sim* sim);

/* ---- groups: several simulations advanced by shared launches ----------------------------- */
/* No counterpart in the reference, which runs one simulation per process. A group holds
 * handles of 2D isotropic relaxation simulations on one device, of any grid sizes and media
 * (and any n_shots), each at its own step. fw_group_run advances every member by n_steps; each
 * step is one velocity launch, one pressure launch and one source/air/sensor launch over the
 * tiles of all members (x-marching kernels, `width` lanes of 64, 128 or 256 and `chunk` rows
 * per workgroup; a member too small to give about 128 workgroups takes fewer rows, down to 2),
 * so many small or uneven jobs fill the GPU together. width = 0 or chunk = 0: timed on the
 * group when every member's wave fields are zero (created or reset, not run since), else 256
 * and 16; fw_group_config reports the shape in use. Each member's records are those of
 * fw_sim_run with the same kernel shape, bit for bit; read them with
 * fw_sim_download_sensors. Creating the group runs each member's first-run set-up. Between
 * group runs a member may be run alone, reset or read; new maps make fw_group_run fail with
 * FW_ERR_STATE. Destroy the group before destroying its members. */
typedef struct fw_group fw_group;
int fw_group_create(fw_sim* const* sims, int32_t n, int32_t width, int32_t chunk, fw_group** out);
int fw_group_config(const fw_group* group, int32_t* width, int32_t* chunk);
/* The same with the kernel family chosen: variant 1 is the x-marching kernels above
 * (fw_group_create); variant 2 the tiled float4 kernels, 256 columns x `chunk` rows (8, 16 or
 * 32; 0: timed) per workgroup, for members with ny % 4 == 0 and at most 768 sound speeds
 * (FW_ERR_UNSUPPORTED otherwise); variant 0 times both families (the tiled one when every
 * member takes it) with width and chunk 0. Each member's records equal fw_sim_run with the same family and shape bit for bit
 * ("tile:<chunk>" for variant 2). fw_group_shape reports the family and shape in use. */
int fw_group_create_shaped(fw_sim* const* sims, int32_t n, int32_t variant, int32_t width, int32_t chunk,
                           fw_group** out);
int fw_group_shape(const fw_group* group, int32_t* variant, int32_t* width, int32_t* chunk);
int fw_group_run(fw_group* group, int64_t n_steps);
/* Run steady group steps from a captured HIP graph of `steps` steps (0: plain launches, the
 * default); no counterpart in the reference. fw_group_run then replays the graph for whole
 * multiples of `steps` and launches the remainder step by step; every member's records are the
 * same either way. Cuts the host cost of a step's three launches for groups of small grids. */
int fw_group_set_graph_steps(fw_group* group, int64_t steps);
int fw_group_destroy(fw_group* group);

/* ---- adjoint: gradients of a functional of the traces ----------------------------------- */
/* None of these has a counterpart in the reference, which runs forward simulations only.
 * For J = sum_{n,j} w[n][j] * rec[n][j] (rec: the traces of a full run from rest, w: e.g. the
 * residual of an L2 misfit), the sweep returns dJ/dK and dJ/drho per voxel, the coefficient
 * sets (dcmap, dmap) held fixed. The sweep reads a tape of the forward run (ndim + 1 grids a
 * step: each velocity update's corrected gradient per axis and the pressure update's corrected
 * divergence). The run is cut into segments of `checkpoint_every` steps: the forward run keeps
 * its state before each and tapes the last; the backward sweep recomputes the tape of every
 * other segment from its checkpoint. checkpoint_every >= nT tapes the whole run, with no
 * recomputation. */
typedef struct fw_adj fw_adj;
/* Bind to a single-transmit handle whose inputs are uploaded (n_shots 1); allocates the
 * checkpoints (ceil(nT / checkpoint_every) states), the tape of one segment and scratch.
 * checkpoint_every 0: the whole run when its tape fits in half of the device's free memory,
 * else the segment length that minimises checkpoints plus tape. The handle's velocity and
 * pressure kernels become the tiled 2D or per-cell ones if it had chosen others. */
int fw_adj_create(fw_sim* sim, int64_t checkpoint_every, fw_adj** out);
/* The segment length in use (as chosen for checkpoint_every 0). */
int fw_adj_checkpoint_every(const fw_adj* adj, int64_t* every);
/* Reset `sim`, run its nT steps from rest keeping checkpoints; the traces are then available
 * through fw_sim_download_sensors. */
int fw_adj_forward(fw_adj* adj);
/* grad_rec: float32 [n_records, ncoordsout] (the weights w); grad_k, grad_rho: float32 grids
 * [nx, ny(, nz)] (grid_bytes each) receiving dJ/dK and dJ/drho. */
int fw_adj_backward(fw_adj* adj, const float* grad_rec, size_t grad_bytes, float* grad_k, float* grad_rho,
                    size_t grid_bytes);
int fw_adj_destroy(fw_adj* adj);

#ifdef __cplusplus
}
#endif

#endif /* FULLWAVE_HIP_H */
