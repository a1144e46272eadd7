/*
 * nbfmm_amd: C ABI of the MI355X-native nbfmm force-evaluation and
 * time-step path.
 *
 * It replaces the C++ class nbfmm::Solver (include/nbfmm/core/solver.hpp:22-209)
 * and the Verlet step of nbfmm::Stars::update (source/nbfmm/display/stars/update.cu:26-55).
 * Each entry point below names the reference interface it stands for.
 *
 * Conventions
 *  - Every function returns an int status: NBFMM_OK (0) or one of the codes
 *    below.  Nothing aborts; nbfmm_last_error() gives a message for the
 *    calling thread.
 *  - Array arguments named d_* are device pointers (hipMalloc'd or torch
 *    tensors).  Particle layouts are the reference's: positions and effects
 *    are float2[N] (x,y interleaved, passed as float*), weights float[N].
 *  - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *    Solver entry points only enqueue work on it: no host synchronisation, no
 *    allocation, so a call can be captured into a hipGraph.
 *  - A solver is not thread-safe: use one per stream/thread, as the
 *    reference's Solver owns its scratch buffers.
 *  - num_level keeps the reference meaning: the leaf grid is
 *    base_dim x base_dim with base_dim = 2^(num_level+1)
 *    (source/nbfmm/core/solver.cu:19), 0 <= num_level <= 8.
 */
#ifndef NBFMM_AMD_H_
#define NBFMM_AMD_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NBFMM_OK 0
#define NBFMM_ERR_INVALID_ARGUMENT 1 /* bad size, level, limits, epsilon or null pointer */
#define NBFMM_ERR_OUT_OF_MEMORY 2
#define NBFMM_ERR_HIP 3              /* a HIP runtime call failed */
#define NBFMM_ERR_CAPACITY 4         /* num_particle exceeds the solver's capacity */
#define NBFMM_ERR_NO_DEVICE 5
#define NBFMM_ERR_IO 6                /* a file could not be opened or written */

/* Status flags raised by the device (read with nbfmm_solver_flags). */
#define NBFMM_FLAG_OUT_OF_BOX 1u /* a particle was outside the limits or NaN; its cell index was clamped */
#define NBFMM_FLAG_P2P_WAIT 2u   /* a P2P wave gave up a wait on another wave of its block (after
                                  * nbfmm_set_p2p_wait_polls polls): for the segments before its own in
                                  * a split run, or for the block's split-run slot.  That run was then
                                  * summed by one wave, segment after segment: its results are the same
                                  * bits as the split sum's, only later.  Not expected at the default
                                  * limit. */
#define NBFMM_FLAG_P2P_LOST 4u   /* a P2P wave found the split-run protocol broken (a lock held past every
                                  * bound) and dropped a segment: that run's results are missing.  Cannot
                                  * happen unless the protocol has a defect; reported rather than hung on. */

/* Buffer ids for nbfmm_buffer_read / nbfmm_buffer_write.  They name the
 * reference Solver members that NBFMM_CHECK makes public
 * (include/nbfmm/core/solver.hpp:24-76) and use their layouts:
 *   POSITION float2[N] sorted, EFFECT float2[N] sorted, WEIGHT float[N] sorted,
 *   INDEX int2[N] sorted cell indices, PERM int[N], HEAD int[base_dim^2+1],
 *   CELL_POSITION float2 / CELL_EFFECT float2 / CELL_WEIGHT float cubes
 *   [levels][base_dim][base_dim] with levels = max(num_level, 1); level-l
 *   cell (X,Y) sits at l*base_dim^2 + (Y<<l)*base_dim + (X<<l)
 *   (source/nbfmm/core/solver/m2m.cu:35).  Cube slots that no cell of that
 *   level uses read as zero. */
#define NBFMM_BUF_POSITION 0
#define NBFMM_BUF_EFFECT 1
#define NBFMM_BUF_WEIGHT 2
#define NBFMM_BUF_INDEX 3
#define NBFMM_BUF_PERM 4
#define NBFMM_BUF_HEAD 5
#define NBFMM_BUF_CELL_POSITION 6
#define NBFMM_BUF_CELL_EFFECT 7
#define NBFMM_BUF_CELL_WEIGHT 8

typedef struct nbfmm_solver nbfmm_solver;

/* Version of this ABI (major*10000 + minor*100 + patch). */
int nbfmm_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char* nbfmm_last_error(void);

/* nbfmm::Solver::Solver(num_level, max_num_particle, position_limits)
 * (include/nbfmm/core/solver.hpp:87, source/nbfmm/core/solver.cu:14-46).
 * position_limits = {x_min, y_min, x_max, y_max}.  epsilon is the kernel
 * softening that the reference hard-codes to 0.1
 * (include/nbfmm/core/kernel_function.hpp:35); it must be > 0.  Unlike the
 * reference, max_num_particle is not limited to 65535 (solver.cu:26).
 * Allocates all scratch on the current HIP device. */
int nbfmm_solver_create(int num_level, int max_num_particle, const float position_limits[4], float epsilon,
                        nbfmm_solver** out);

/* nbfmm::Solver::~Solver (source/nbfmm/core/solver.cu:49-61). */
int nbfmm_solver_destroy(nbfmm_solver* solver);

/* Sizes the solver was created with. */
int nbfmm_solver_info(const nbfmm_solver* solver, int* num_level, int* base_dim, int* max_num_particle);

/* nbfmm::Solver::solve (include/nbfmm/core/solver.hpp:102-103,
 * source/nbfmm/core/solver/solve.cu:11-27): effects of all particles in the
 * caller's original order.  d_position_origin float2[N], d_weight_origin
 * float[N], d_effect_origin float2[N] (output). */
int nbfmm_solve(nbfmm_solver* solver, int num_particle, const float* d_position_origin,
                const float* d_weight_origin, float* d_effect_origin, void* stream);

/* No counterpart in the reference: nbfmm_solve for callers that cannot
 * allocate device memory (code under hipGraph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No solver: the sizes come with the call,
 * and all scratch lives in a static 16 MiB device array per GPU (held until
 * the process ends); the call allocates nothing.  Same kernels as
 * nbfmm_solve with a cold sort, so the same bits.  A size whose scratch
 * exceeds 16 MiB (about 40 B per particle, 24 B per cell of the tree and
 * 8 B per leaf: 262144 particles at num_level 7 fit) fails with
 * NBFMM_ERR_CAPACITY.  Calls that overlap on one GPU share the array and
 * race; out-of-box particles are clamped but no flag is reported. */
int nbfmm_solve_noalloc(int num_level, int num_particle, const float position_limits[4], float epsilon,
                        const float* d_position_origin, const float* d_weight_origin, float* d_effect_origin,
                        void* stream);

/* The eight stages of nbfmm::Solver (include/nbfmm/core/solver.hpp:111-207),
 * with the same pre/post conditions on the solver's buffers.  solve() is
 * their composition; the stages exist so each one can be checked in
 * isolation, as the reference's unit tests do (check/core/solver/ sources). */
int nbfmm_predo(nbfmm_solver* solver, int num_particle, const float* d_position_origin,
                const float* d_weight_origin, void* stream);                                 /* predo.cu:105-135 */
int nbfmm_p2p(nbfmm_solver* solver, int num_particle, void* stream);                        /* p2p.cu:166-176 */
int nbfmm_p2m(nbfmm_solver* solver, int num_particle, void* stream);                        /* p2m.cu:79-110 */
int nbfmm_m2m(nbfmm_solver* solver, void* stream);                                          /* m2m.cu:59-74 */
int nbfmm_m2l(nbfmm_solver* solver, void* stream);                                          /* m2l.cu:65-81 */
int nbfmm_l2l(nbfmm_solver* solver, void* stream);                                          /* l2l.cu:41-56 */
int nbfmm_l2p(nbfmm_solver* solver, int num_particle, void* stream);                        /* l2p.cu:40-48 */
int nbfmm_postdo(nbfmm_solver* solver, int num_particle, float* d_effect_origin, void* stream); /* postdo.cu:38-45 */

/* Access to the solver's buffers in the reference layouts (the NBFMM_CHECK
 * public members, include/nbfmm/core/solver.hpp:24-28).  `count` is the
 * number of particles for the per-particle buffers and is ignored for HEAD
 * and the cubes.  nbfmm_buffer_bytes gives the byte size for that count.
 * The copies are device-to-device on `stream`. */
int nbfmm_buffer_bytes(const nbfmm_solver* solver, int buffer_id, int count, size_t* bytes);
int nbfmm_buffer_read(nbfmm_solver* solver, int buffer_id, int count, void* d_dst, void* stream);
int nbfmm_buffer_write(nbfmm_solver* solver, int buffer_id, int count, const void* d_src, void* stream);

/* Synchronises the device and returns the accumulated NBFMM_FLAG_* bits of
 * this solver; clear = 1 resets them. */
int nbfmm_solver_flags(nbfmm_solver* solver, unsigned* flags, int clear);

/* updateDevice (source/nbfmm/display/stars/update.cu:26-42):
 *   x_cur' = 2*x_cur - x_pre + acc*tick*tick,  x_pre' = x_cur
 * for num_star float2 entries.  tick already includes sqrt(G)
 * (source/nbfmm/display/stars.cu:31). */
int nbfmm_update(int num_star, float tick, float* d_position_cur, float* d_position_pre,
                 const float* d_acceleration, void* stream);

/* nbfmm::Stars::update without prune (source/nbfmm/display/stars/update.cu:47-55):
 * solve at d_position_cur into d_acceleration, then the Verlet update.
 * The L2P, un-permute and Verlet stages run fused in one kernel. */
int nbfmm_step(nbfmm_solver* solver, int num_particle, float tick, float* d_position_cur,
               float* d_position_pre, const float* d_weight, float* d_acceleration, void* stream);

/* nbfmm::Stars::prune (source/nbfmm/display/stars/prune.cu:28-62): stars with
 * x < x_min, y < y_min, x > x_max or y > y_max (the solver's position limits;
 * NaN stays) are moved behind the others; both groups keep their order, and
 * d_position_cur, d_position_pre and d_weight are permuted in place.  The
 * number kept (the new num_star) is written to *num_kept.  Synchronises the
 * stream to return it, as the reference's thrust::find does. */
int nbfmm_prune(nbfmm_solver* solver, int num_star, float* d_position_cur, float* d_position_pre,
                float* d_weight, int* num_kept, void* stream);

/* num_steps calls of nbfmm_step (the same bits; with the warm start on, the
 * P2P of every step but the last also makes the first pass of the next
 * step's sort, and leaves d_acceleration alone, which no step reads: the
 * call's last step writes it, so it ends with the last step's acceleration
 * as after num_steps nbfmm_step calls).  use_graph = 1 captures one step into a
 * hipGraph on the first call and replays it.  A graph cannot be captured on
 * the NULL stream: with stream NULL the replays run on a stream the solver
 * makes on first use, ordered after the work already queued on the NULL
 * stream and before the work queued on it afterwards. */
int nbfmm_run(nbfmm_solver* solver, int num_particle, float tick, int num_steps, float* d_position_cur,
              float* d_position_pre, const float* d_weight, float* d_acceleration, int use_graph, void* stream);

/* Stage timing (no reference counterpart): when enabled, solve, step and run
 * (run then steps eagerly, with the launches its graph would replay) record
 * HIP events on `stream` around the four stages of each call: the sort, the
 * upward pass (P2M + M2M), M2L, and P2P (with L2L, L2P, un-permute and the
 * Verlet epilogue).  nbfmm_stage_times synchronises and returns the summed
 * milliseconds of each stage (ms[4]) and the number of timed calls since the
 * last reset; nbfmm_p2p_time returns the P2P stage's sum alone.
 * Pair counts for roofline figures come from nbfmm_pair_count (synchronises;
 * ordered pairs of the P2P stage for the current head, self pairs excluded). */
int nbfmm_profile_enable(nbfmm_solver* solver, int enable);
int nbfmm_p2p_time(nbfmm_solver* solver, double* total_ms, int* launches, int reset);
int nbfmm_stage_times(nbfmm_solver* solver, double* ms, int* calls, int reset);
int nbfmm_pair_count(nbfmm_solver* solver, int num_particle, double* pairs, int* occupied_cells);

/* ---------------------------------------------------------------- warm start
 * No reference counterpart (the reference sorts from scratch every call,
 * source/nbfmm/core/solver/predo.cu:122-127).  solve, step, run and the
 * adjoints start their sort from the solver's previous sort when that was of
 * the same num_particle: particles whose leaf did not change keep their order,
 * and only the others are placed.  The result is the same stable sort, bit
 * for bit.  When a leaf gains more than 512 particles in one call, the cold
 * sort's passes run instead (decided on the device, inside the same
 * launches), so a call never costs much more than a cold sort.  Writing a
 * buffer through nbfmm_buffer_write makes the next sort cold.  Enabled by
 * default; disabling it forces the cold sort every call. */
int nbfmm_set_warm_start(nbfmm_solver* solver, int enable);

/* ---------------------------------------------------------------- small trees
 * No reference counterpart.  For num_level 1..4 and 1..4096 particles,
 * solve, step and run sort the particles and build every multipole level in
 * one kernel of one block (the reference's predo, P2M and M2M,
 * source/nbfmm/core/solver/predo.cu:26-135, p2m.cu:24-110, m2m.cu:22-54),
 * which at these sizes replaces five latency-bound launches; M2L and the
 * P2P epilogue follow as usual.  Its sort is cold, so the warm start does not
 * apply there.  Every array and result is the same, bit for bit.  Enabled
 * by default; disabling it takes the general kernels. */
int nbfmm_set_small_front(nbfmm_solver* solver, int enable);

/* ---------------------------------------------------------------- models
 * Initial conditions of the reference's star models (the .cu files under source/nbfmm/model/,
 * selected by name in main/main.cu:78-222), written into device arrays
 * (position_cur, position_pre: float2[n]; weight: float[n]).  Uniform draws
 * come from a counter-based hash of (seed, particle, draw) in (0, 1] instead
 * of curand_init(0, i, 0) streams, so a model is reproducible across
 * devices; the geometry, the chord-angle position_pre of the rotating models
 * and the orbit offsets of the double disks follow the reference.  eps is
 * the kernel function's softening (the reference's kernelFunction uses 0.1);
 * tick already includes sqrt(G) as Stars passes it. */
int nbfmm_model_disk(int num_particle, float center_x, float center_y, float radius, float weight, float tick,
                     float eps, unsigned seed, float* d_position_cur, float* d_position_pre, float* d_weight,
                     void* stream);  /* model/disk.cu:27-81 */
int nbfmm_model_disk_static(int num_particle, float center_x, float center_y, float radius, float weight,
                            unsigned seed, float* d_position_cur, float* d_position_pre, float* d_weight,
                            void* stream);  /* model/disk_static.cu */
int nbfmm_model_disk_center(int num_particle, float center_x, float center_y, float radius, float weight,
                            float center_weight, float tick, float eps, unsigned seed, float* d_position_cur,
                            float* d_position_pre, float* d_weight, void* stream);  /* model/disk_center.cu */
/* uniform_angles = 0: model/circle_center.cu; 1: model/circle_uniform_center.cu */
int nbfmm_model_circle_center(int num_particle, float center_x, float center_y, float radius, float weight,
                              float center_weight, float tick, float eps, int uniform_angles, unsigned seed,
                              float* d_position_cur, float* d_position_pre, float* d_weight, void* stream);
int nbfmm_model_rectangle(int num_particle, float center_x, float center_y, float width, float height,
                          float max_weight, unsigned seed, float* d_position_cur, float* d_position_pre,
                          float* d_weight, void* stream);  /* model/rectangle.cu */
/* centered = 0: model/double_disk.cu (center weights unused); 1: model/double_disk_center.cu.
 * Disk 1 fills entries [0, n1), disk 2 [n1, n1 + n2). */
int nbfmm_model_double_disk(int num_particle1, int num_particle2, const float center1[2], const float center2[2],
                            float radius1, float radius2, float weight, float center_weight1, float center_weight2,
                            float eccentricity, float tick, float eps, int centered, unsigned seed,
                            float* d_position_cur, float* d_position_pre, float* d_weight, void* stream);

/* ---------------------------------------------------------------- video
 * Frames are planar YUV 4:2:0 in device memory: width*height luma bytes, then
 * width*height/2 chroma bytes (main/main.cu:234, FRAME_SIZE).  width and
 * height must be even. */

/* Stars::display (source/nbfmm/display/stars/display.cu:219-226) without its
 * trailing Stars::prune: clears the frame (luma 0, chroma 128) and lights each
 * star with a luma-255 stamp sized floor(sqrt(weight) / size_scale), clipped
 * to the frame.  display_limits = {x_min, y_min, x_max, y_max} (host); row 0
 * is y_max.  Asynchronous on `stream`; no allocation. */
int nbfmm_display(int num_star, int width, int height, float size_scale, const float display_limits[4],
                  const float* d_position, const float* d_weight, unsigned char* d_frame, void* stream);

/* The YUV4MPEG2 file of main/main.cu:234-250.  open writes the stream header
 * "YUV4MPEG2 W<w> H<h> F<fps>:1 Ip A1:1 C420"; write_frame appends "FRAME\n"
 * and one frame copied from device memory on `stream` (through pinned
 * staging: the file write of a frame happens during the next call or in
 * close, so it overlaps the device work queued meanwhile); close writes what
 * is pending, reports the number of frames written and frees the writer. */
typedef struct nbfmm_y4m nbfmm_y4m;
int nbfmm_y4m_open(const char* path, int width, int height, int fps, nbfmm_y4m** writer);
int nbfmm_y4m_write_frame(nbfmm_y4m* writer, const unsigned char* d_frame, void* stream);
int nbfmm_y4m_close(nbfmm_y4m* writer, long* num_frames);

/* ---------------------------------------------------------------- tuning
 * The P2P kernel comes in variants (LDS staging size, unroll, waves per
 * block) and runs on a grid of blocks_per_cu blocks per compute unit.  Every configuration gives
 * bit-identical results; only the time differs.  No reference counterpart
 * (the reference fixes 64-thread blocks, source/nbfmm/core/solver/p2p.cu:16). */
int nbfmm_p2p_variant_count(void);
int nbfmm_p2p_variant_info(int variant, int* stage, int* unroll, int* waves_per_block);
int nbfmm_set_p2p_config(nbfmm_solver* solver, int variant, int blocks_per_cu);
int nbfmm_get_p2p_config(const nbfmm_solver* solver, int* variant, int* blocks_per_cu);
/* Bound on every wait of a P2P wave for another wave of its block (each poll a
 * short sleep; default 2^20, >= 1).  After it the wave gives up the wait and
 * sums the run alone, segment after segment (NBFMM_FLAG_P2P_WAIT): the same
 * bits, only later.  Exists so that tests can force the give-up path from the
 * host; no reference counterpart. */
int nbfmm_set_p2p_wait_polls(nbfmm_solver* solver, int polls);
/* Times every configuration on this input (repeats launches each), keeps the
 * fastest in the solver and returns it.  Synchronises. */
int nbfmm_autotune(nbfmm_solver* solver, int num_particle, const float* d_position_origin,
                   const float* d_weight_origin, int repeats, int* best_variant, int* best_blocks_per_cu,
                   float* best_ms, void* stream);

/* ---------------------------------------------------------------- adjoint
 * No reference counterpart (the reference only evaluates forces).  The
 * gradients treat each particle's cell as fixed (the cell assignment is
 * piecewise constant in the positions). */

/* Vector-Jacobian product of solve: given d_grad_effect = dloss/d effect
 * (float2[N], original order) at positions/weights (as for nbfmm_solve),
 * writes dloss/d position (float2[N]) and dloss/d weight (float[N]). */
int nbfmm_solve_adjoint(nbfmm_solver* solver, int num_particle, const float* d_position_origin,
                        const float* d_weight_origin, const float* d_grad_effect, float* d_grad_position,
                        float* d_grad_weight, void* stream);

/* Adjoint of one Verlet step x_{k+1} = 2 x_k - x_{k-1} + tick^2 a(x_k, m),
 * at positions d_position = x_k.  In: lambda = dloss/dx_{k+1} and
 * pi = dloss/d(x_pre)_{k+1}; out (in place): lambda = dloss/dx_k,
 * pi = dloss/d(x_pre)_k; d_grad_weight += dloss/dm through this step. */
int nbfmm_step_adjoint(nbfmm_solver* solver, int num_particle, float tick, const float* d_position,
                       const float* d_weight, float* d_lambda, float* d_pi, float* d_grad_weight, void* stream);

/* Checkpointed trajectories: the forward stores (x, x_pre) every
 * checkpoint_every steps and the positions at the observation steps; the
 * backward recomputes each segment from its checkpoint and applies
 * nbfmm_step_adjoint step by step, adding the observation gradients.
 * Sizes of the caller-allocated device buffers: */
int nbfmm_trajectory_workspace(int num_particle, int num_steps, int checkpoint_every, size_t* checkpoint_bytes,
                               size_t* work_bytes);
/* obs_steps: host array of n_obs step indices in [0, num_steps];
 * d_obs_position: n_obs x N float2 (output). */
int nbfmm_trajectory_forward(nbfmm_solver* solver, int num_particle, float tick, int num_steps,
                             int checkpoint_every, float* d_position_cur, float* d_position_pre,
                             const float* d_weight, float* d_acceleration, float* d_checkpoints, int n_obs,
                             const int* obs_steps, float* d_obs_position, void* stream);
/* d_obs_grad: n_obs x N float2 = dloss/d(observed positions).  Writes the
 * gradients wrt the initial positions, initial x_pre and the weights.  The
 * step adjoint at a recomputed step uses the sort and tree that step built
 * (the same bits as sorting again).  The recompute of each segment runs on a
 * stream of the library's own (ordered after the work queued on `stream`
 * before the call) while `stream` runs the adjoint steps of the segment
 * after it; all of the call's work is done once `stream` reaches the end of
 * what the call queued.  The solver keeps two sets of checkpoint_every such
 * states (one when num_steps <= checkpoint_every), allocated by the first
 * call that needs them (a synchronous allocation) and freed with the solver,
 * each 20 B per particle of its capacity plus its leaf grid and tree.
 * Concurrent calls on one solver race. */
int nbfmm_trajectory_backward(nbfmm_solver* solver, int num_particle, float tick, int num_steps,
                              int checkpoint_every, const float* d_weight, const float* d_checkpoints, int n_obs,
                              const int* obs_steps, const float* d_obs_grad, float* d_grad_position0,
                              float* d_grad_position_pre0, float* d_grad_weight, float* d_work, void* stream);

/* ---------------------------------------------------------------- batches
 * Many independent systems with their own particle counts, levels, boxes and
 * softenings, advanced together: each stage of the step is one launch over
 * all systems, and nbfmm_batch_run replays the whole step as one hipGraph.
 * No reference counterpart: the reference solves one system per Solver
 * (include/nbfmm/core/solver.hpp:22); a batch of S systems gives, system by
 * system, the same bits as S separate solvers.
 * Particle arrays are the systems' arrays concatenated in system order;
 * nbfmm_batch_offsets gives the S+1 offsets (in particles). */
typedef struct nbfmm_batch nbfmm_batch;

int nbfmm_batch_create(int num_systems, const int* num_particle, const int* num_level,
                       const float* position_limits /* 4 per system */, const float* epsilon,
                       nbfmm_batch** out);
int nbfmm_batch_destroy(nbfmm_batch* batch);
int nbfmm_batch_info(const nbfmm_batch* batch, int* num_systems, int* total_particles);
int nbfmm_batch_offsets(const nbfmm_batch* batch, int* part_off /* num_systems + 1 */);
/* Verlet step sizes (dt*sqrt(G)) per system, host array; synchronises. */
int nbfmm_batch_set_ticks(nbfmm_batch* batch, const float* ticks);
/* solve() of every system. */
int nbfmm_batch_solve(nbfmm_batch* batch, const float* d_position_origin, const float* d_weight_origin,
                      float* d_effect_origin, void* stream);
/* Stars::update without prune for every system. */
int nbfmm_batch_step(nbfmm_batch* batch, float* d_position_cur, float* d_position_pre, const float* d_weight,
                     float* d_acceleration, void* stream);
int nbfmm_batch_run(nbfmm_batch* batch, int num_steps, float* d_position_cur, float* d_position_pre,
                    const float* d_weight, float* d_acceleration, int use_graph, void* stream);
/* Synchronises; flags[s] = NBFMM_FLAG_* bits of system s. */
int nbfmm_batch_flags(nbfmm_batch* batch, unsigned* flags, int clear);
/* P2P configuration of the batch (see nbfmm_set_p2p_config): the grid of
 * blocks_per_cu blocks per CU is shared among the systems in proportion to
 * their P2P work (particle counts until the batch is balanced). */
int nbfmm_batch_set_p2p_config(nbfmm_batch* batch, int variant, int blocks_per_cu);
int nbfmm_batch_get_p2p_config(const nbfmm_batch* batch, int* variant, int* blocks_per_cu);
/* nbfmm_set_p2p_wait_polls for every system of the batch. */
int nbfmm_batch_set_p2p_wait_polls(nbfmm_batch* batch, int polls);
/* Sorts the given positions and shares the P2P grids (forward and adjoint)
 * among the systems in proportion to their P2P pairs plus 32 per particle.
 * The first solve, step, run or adjoint call of a batch does this by itself
 * after its sort (one synchronisation; skipped while the stream is being
 * captured), so an explicit call is needed only to re-balance.  Results do
 * not depend on the shares, only the time does.  Synchronises.  No
 * reference counterpart. */
int nbfmm_batch_balance(nbfmm_batch* batch, const float* d_position, const float* d_weight, void* stream);
/* Warm-started sort of every system (see nbfmm_set_warm_start); on by default. */
int nbfmm_batch_set_warm_start(nbfmm_batch* batch, int enable);
/* nbfmm_solve_adjoint of every system (concatenated arrays). */
int nbfmm_batch_solve_adjoint(nbfmm_batch* batch, const float* d_position_origin, const float* d_weight_origin,
                              const float* d_grad_effect, float* d_grad_position, float* d_grad_weight,
                              void* stream);
/* nbfmm_step_adjoint of every system, each with its own tick. */
int nbfmm_batch_step_adjoint(nbfmm_batch* batch, const float* d_position, const float* d_weight, float* d_lambda,
                             float* d_pi, float* d_grad_weight, void* stream);
/* For a checkpointed trajectory's recompute (no reference counterpart):
 * keep the batch's current sort and tree (of the positions of its last
 * step or solve) in slot `slot` (0..4096; the first use of a slot allocates
 * it, synchronously, as large as the batch's own sorted arrays and trees),
 * and run nbfmm_batch_step_adjoint on a kept state without sorting again:
 * the same bits as nbfmm_batch_step_adjoint at those positions. */
int nbfmm_batch_keep_sort(nbfmm_batch* batch, int slot, void* stream);
int nbfmm_batch_step_adjoint_kept(nbfmm_batch* batch, int slot, float* d_lambda, float* d_pi, float* d_grad_weight,
                                  void* stream);
/* num_steps steps as nbfmm_batch_run's eager steps, keeping the sort and
 * tree of step k (of the positions it starts from) in slot first_slot + k,
 * then the sort and tree of the final positions in slot first_slot +
 * num_steps (as nbfmm_batch_step_adjoint would build them there): the same
 * bits and kept states as nbfmm_batch_step followed by
 * nbfmm_batch_keep_sort, step by step, with the sorts chained as in a run
 * (first_slot >= 0, first_slot + num_steps <= 4096).  d_acceleration is
 * scratch.  No reference counterpart. */
int nbfmm_batch_run_keep(nbfmm_batch* batch, int num_steps, int first_slot, float* d_position_cur,
                         float* d_position_pre, const float* d_weight, float* d_acceleration, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NBFMM_AMD_H_ */
