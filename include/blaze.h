/*
 * blaze.h — C ABI of the MI355X smoke/fire solver step.
 *
 * Plain C: POD structs, pointers, sizes and int status codes. No HIP or
 * torch types cross this boundary. Every call returns BLZ_OK (0) or a
 * negative BLZ_E_* code and leaves a thread-local message readable with
 * blz_last_error().
 *
 * Reference interface that this ABI replaces (asiroisvigneux/Blaze):
 *   class FluidSolver            src/FluidSolver.h:38-110
 *     FluidSolver(Timer*, SceneSettings*)   src/FluidSolver.cu:671-741  -> blz_create
 *     ~FluidSolver()                        src/FluidSolver.cu:743-768  -> blz_destroy
 *     addSource()                           src/FluidSolver.cu:770-821  -> blz_add_source
 *     step()                                src/FluidSolver.cu:858-955  -> blz_step
 *     getTempGrid()                         src/FluidSolver.h:108       -> blz_get_field(BLZ_FIELD_T)
 *   struct SceneSettings (solver subset)    src/SceneSettings.h:32-71   -> blz_settings
 *   Parser::sceneParse                      src/Parser.h:41-172         -> blz_parse_scene,
 *                                                                          blz_parse_render
 *   RenderEngine::render                    src/RenderEngine.cu:312-365 -> blz_render
 *   RenderEngine::writeToDisk               src/RenderEngine.cu:367-476 -> blz_write_image
 *   Parser::sourceParticleParse             src/Parser.h:174-211        -> blz_read_particles
 *
 * Threading: one handle per host thread. A handle owns all of its device
 * memory and one HIP stream; calls are synchronous from the host's view
 * unless named *_async. Distinct handles may run concurrently.
 */
#ifndef BLAZE_H
#define BLAZE_H

#ifdef __cplusplus
extern "C" {
#endif

#define BLZ_ABI_VERSION 1

/* opaque solver handle (owns device memory and one HIP stream) */
typedef struct blz_solver blz_solver;

/* status codes */
#define BLZ_OK 0
#define BLZ_E_INVALID (-1)   /* bad argument or settings */
#define BLZ_E_HIP (-2)       /* HIP runtime error */
#define BLZ_E_NOMEM (-3)     /* device or host allocation failed */
#define BLZ_E_IO (-4)        /* file missing or unreadable */
#define BLZ_E_PARSE (-5)     /* malformed .blz / missing key */
#define BLZ_E_NUMERIC (-6)   /* non-finite state detected */

/* field selectors (front buffers unless stated) */
#define BLZ_FIELD_T 0        /* temperature, W x H x D */
#define BLZ_FIELD_U 1        /* x face velocity, (W+1) x H x D */
#define BLZ_FIELD_V 2        /* y face velocity, W x (H+1) x D (grid y is flipped) */
#define BLZ_FIELD_W 3        /* z face velocity, W x H x (D+1) */
#define BLZ_FIELD_DENSITY 4  /* smoke density, W x H x D (only when enabled) */
#define BLZ_FIELD_P 5        /* pressure, W x H x D (natural layout on copy; device view is split by colour) */
#define BLZ_FIELD_RHS 6      /* divergence right-hand side, W x H x D */
#define BLZ_NUM_FIELDS 7

/* single stages, for per-kernel parity harnesses (blz_run_stage) */
#define BLZ_STAGE_COOLDOWN 0   /* temperature_cooldown_kernel  src/FluidSolver.cu:118-124 */
#define BLZ_STAGE_DRAG 1       /* apply_drag_kernel            src/FluidSolver.cu:126-140 */
#define BLZ_STAGE_VORTICITY 2  /* vorticity_confinement_kernel src/FluidSolver.cu:203-248 */
#define BLZ_STAGE_BUOYANCY 3   /* add_buoyancy_kernel          src/FluidSolver.cu:142-156 */
#define BLZ_STAGE_WIND 4       /* add_wind_kernel              src/FluidSolver.cu:337-390 */
#define BLZ_STAGE_TURBULENCE 5 /* compute_turbulence + add_curl_noise src/FluidSolver.cu:271-335 */
#define BLZ_STAGE_DIVERGENCE 6 /* compute_divergence_kernel    src/FluidSolver.cu:475-515 */
#define BLZ_STAGE_PRESSURE 7   /* gs_solve_kernel              src/FluidSolver.cu:517-599 */
#define BLZ_STAGE_PROJECT_UPDATE 8 /* pressure_gradient_update src/FluidSolver.cu:601-669 */
#define BLZ_STAGE_ADVECT 9     /* advect_RK3_kernel x4 + swap  src/FluidSolver.cu:431-465,942-953 */
#define BLZ_NUM_STAGES 10

/*
 * Solver subset of SceneSettings (src/SceneSettings.h:32-71) with explicit
 * arrays in place of CUDA vector types, plus optional extensions that
 * default to the reference behaviour (see blz_settings_default).
 */
typedef struct blz_settings {
    int grid_res[3];          /* gridRes: W, H, D (cells) */
    float dx;                 /* dx */
    float dt;                 /* dt (per frame) */
    float domain_min[3];      /* domainBboxMin (world) */
    float domain_max[3];      /* domainBboxMax (world) */
    int closed_bounds[6];     /* closedBounds: xMin,xMax,yMin,yMax,zMin,zMax */
    int max_iter_solve;       /* maxIterSolve (cap when residual_tol > 0) */
    float density;            /* density rho of the projection */
    float gravity;            /* gravity (negative = world down) */
    float buoyancy;
    float cooling_rate;
    float vorticity_conf;
    float drag;
    int wind_dir;             /* 0:X+ 1:X- 2:Z+ 3:Z- */
    float wind_amp;
    float wind_speed;
    float wind_turb_amp;
    float wind_turb_scale;
    float turbulence_amp;
    float turbulence_scale;
    float turb_mask_temp_ramp[2];
    float turb_mask_vel_ramp[2];
    /* source description (src/SceneSettings.h:36-39) */
    int source_animated;
    int source_range[2];
    int source_max_particle_count;
    /* extensions (no reference counterpart; defaults reproduce the reference) */
    int noise_seed;           /* turbulence Perlin seed, reference 123 */
    float residual_tol;       /* >0: RBGS until ||div u||_inf <= tol (cap max_iter_solve) */
    int residual_check_every; /* sweeps between residual checks (default 10) */
    float cfl_number;         /* >0: substeps n = ceil(max|u| dt / (cfl dx)) */
    int max_substeps;         /* cap on n (default 8) */
    float smoke_density;      /* >0: enable smoke density grid, splat value */
    int warm_start;           /* 0 plain RBGS; 1 learned correction before the sweeps; N >= 2 also
                                 after every N sweeps (blz_set_warm_start attaches the network) */
    int residual_check_adaptive; /* 1: residual-driven solves space their checks from the measured
                                    contraction (doubling the interval while the tolerance is far,
                                    down to residual_check_every near it; solves with a warm start keep
                                    every interval) and stop at an FP32 fixed point (the same FP64
                                    residual at three checks in a row); 0: a check every
                                    residual_check_every sweeps until tol or the cap */
} blz_settings;

/* per-step statistics of the last blz_step / blz_run_stage call */
typedef struct blz_stats {
    int substeps;             /* CFL substeps taken in the last step */
    int rbgs_iters;           /* RBGS sweeps in the last projection (sum over substeps) */
    int rbgs_iters_last;      /* RBGS sweeps in the last substep's projection */
    double residual;          /* ||div u||_inf proxy after the last projection (FP64), -1 if not computed */
    double residual_initial;  /* same measure before the first sweep of the last projection */
    int warm_start_used;      /* warm-start corrections kept in the solves of the last step (or stage) */
    int warm_start_fallbacks; /* learned corrections rejected so far (pressure restored) */
    double stage_ms[BLZ_NUM_STAGES]; /* device time per stage from HIP events (0 if timing off) */
    double source_ms;         /* device time of the last blz_add_source */
    double step_ms;           /* device time of the last blz_step */
    double rbgs_kernel_ms;    /* device time of the RBGS sweep batches alone */
    int rbgs_launches;        /* RBGS kernel launches behind rbgs_kernel_ms */
    int residual_checks;      /* FP64 residual evaluations in the last step */
    int rbgs_sweeps_timed;    /* RBGS sweeps behind rbgs_kernel_ms */
    int rbgs_kernel;          /* kernel of the last sweeps: 1 two colour passes, 2 fused sweep, 3 row-pair fused sweep */
    int rbgs_sweeps_executed; /* sweeps run in the last step, including the warm start judge's batches on the
                                 path not kept (rbgs_iters counts each judged batch once) */
} blz_stats;

/* synthetic particle set: uniform in a ball, deterministic for a seed */
typedef struct blz_particle_gen {
    int count;
    unsigned long long seed;
    float center[3];          /* world */
    float radius;             /* world */
    float pscale;             /* per-particle pscale (diameter of the splat) */
    float temp;
    float vel[3];             /* constant velocity */
    float radial_vel;         /* + radial_vel * unit(pos - center) */
} blz_particle_gen;

/* ---- settings, scene files, particles (host only, no GPU needed) ---- */
int blz_abi_version(void);
int blz_settings_default(blz_settings* s);
/* .blz JSON -> settings. scene_dir receives the directory of the file
   (src/Parser.h:43-45); particle_file receives the "particleFile" value. */
int blz_parse_scene(const char* path, blz_settings* out,
                    char* scene_dir, int scene_dir_len,
                    char* particle_file, int particle_file_len,
                    blz_particle_gen* gen_out, int* has_gen);
/* .part: 8 float32 per particle [px,py,pz,vx,vy,vz,pscale,temp], no header */
int blz_read_particles(const char* path, float* part8, int max_count, int* count);
int blz_count_particles(const char* path, int* count);
int blz_generate_particles(const blz_particle_gen* g, float* part8);

/* ---- solver ---- */
int blz_create(const blz_settings* s, blz_solver** out);
int blz_destroy(blz_solver* h);
/* Particle-to-grid source (addSource). part8 != NULL uploads count particles
   first; part8 == NULL re-splats the previously uploaded set. */
int blz_add_source(blz_solver* h, const float* part8, int count);
/* Animated sources without a synchronous upload (no reference counterpart;
   the reference parses and copies each frame inside addSource,
   src/FluidSolver.cu:781-791): blz_stage_source copies the next frame's
   particles into one of two pinned buffers and starts the upload on a copy
   stream, returning at once; it touches no state a step uses, so it may run
   on another host thread while blz_step is in flight. blz_add_source_staged
   then splats the frame staged last (the solver stream waits for its
   upload on the device). Same result as blz_add_source with that array. */
int blz_stage_source(blz_solver* h, const float* part8, int count);
int blz_add_source_staged(blz_solver* h);
/* One frame (step): forces, projection, advection, with CFL substeps when
   enabled. time is the frame's simulation time (Timer::time). */
int blz_step(blz_solver* h, float time);
/* Enqueue the same frame without the final stream synchronisation. */
int blz_step_async(blz_solver* h, float time);
int blz_sync(blz_solver* h);
/* One stage with an explicit dt, for parity harnesses. */
int blz_run_stage(blz_solver* h, int stage, float time, float dt);
int blz_get_field(blz_solver* h, int which, void** dev_ptr, int dims[3]);
int blz_copy_field(blz_solver* h, int which, float* host);
int blz_set_field(blz_solver* h, int which, const float* host);
int blz_get_stats(blz_solver* h, blz_stats* out);
/* Enable per-stage HIP event timing (adds events around each stage). */
int blz_set_timing(blz_solver* h, int on);
/* RBGS kernel variant (no reference counterpart): 2, 4, 8 = two in-place
   colour passes per sweep with that many four-slab groups in flight per
   thread; 102, 103 = fused whole-sweep kernel (2 or 3 groups in flight),
   sweeps in pairs through a spare plane set, with 1/2/4 waves per block for
   x = 1/2/4 in 10x; 1x03 .. 1x06 = row-pair fused whole-sweep kernel (two
   rows per wave, 3 to 6 groups in flight, x = 1/2/4 waves per block);
   0 = default: the autotuner's choice for this grid, made at the first
   solve (every variant gives the same bits). */
int blz_set_rbgs_variant(blz_solver* h, int variant);
/* Autotuner (no reference counterpart): time `sweeps` sweeps of every
   variant above on this grid (pressure restored afterwards) and keep the
   fastest; ms_per_sweep (optional, 17 entries) receives each time in the
   order 2, 4, 8, 102, 103, 202, 203, 402, 403, 1103, 1104, 1105, 1106,
   1203, 1204, 1403, 1404 (fused ones skipped when the spare planes do not
   exist). */
int blz_autotune_rbgs(blz_solver* h, int sweeps, int* variant, float* ms_per_sweep);

/* Change the pressure solve's limits of a live handle (no reference
   counterpart; the reference reads maxIterSolve once, src/FluidSolver.cu:
   723-724): the residual_tol, max_iter_solve and residual_check_every
   fields of blz_settings for the following steps. */
int blz_set_solve_limits(blz_solver* h, float residual_tol, int max_iter_solve, int residual_check_every);

/* Residual history of residual-driven solves (no reference counterpart; the
   reference solve runs a fixed sweep count and evaluates no residual,
   src/FluidSolver.cu:517-599). With max_entries > 0 every FP64 residual
   evaluation of a solve is recorded (at most max_entries per solve, the
   rest dropped); 0 turns recording off. blz_get_residual_trace copies the
   history of solve `solve` (0-based, in enqueue order: one per substep) of
   the last blz_step or blz_run_stage into out (at most max entries) and
   sets *n to the number recorded. */
#define BLZ_TRACE_CHECK 0     /* a residual check after a batch of sweeps */
#define BLZ_TRACE_KEPT 1      /* a warm-start correction kept by the judge (residual after its batch) */
#define BLZ_TRACE_REJECTED 2  /* a correction rejected (residual of the uncorrected path after the batch) */
typedef struct blz_trace_entry {
    int iters;                /* sweeps counted when evaluated */
    int kind;                 /* BLZ_TRACE_* */
    double res;               /* ||div u||_inf proxy (FP64) */
} blz_trace_entry;
int blz_set_residual_trace(blz_solver* h, int max_entries);
int blz_get_residual_trace(blz_solver* h, int solve, blz_trace_entry* out, int max, int* n);

/* ---- renderer: the solver's consumer (RenderEngine::render,
   src/RenderEngine.cu:312-365, render_kernel :149-225; Camera, Light,
   Shader, ColorRamp in src/Camera.h, src/Light.h, src/Shader.h,
   src/ColorRamp.h). Ray-marches the temperature grid as emission +
   single-scattered light from directional lights with shadow rays, plus a
   blurred 4x-downsampled multi-scatter term. Texture filtering is explicit
   FP32 trilinear with clamped edges (the reference's linear-filtered,
   clamp-addressed textures, whose hardware weights are 8-bit fixed point). */
#define BLZ_MAX_LIGHTS 10
#define BLZ_MAX_RAMP 10
typedef struct blz_render_settings {
    int width, height;                              /* renderRes */
    float cam_focal, cam_aperture;                  /* cam.focal, cam.aperture */
    float cam_u[3], cam_v[3], cam_w[3];             /* rows of cam.rotMat */
    float cam_trans[3];                             /* cam.t */
    float p_step, s_step, cutoff;                   /* primarySampleRate, shadowSamplingRate, cutoff */
    int light_count;                                /* lightCount (<= BLZ_MAX_LIGHTS) */
    float light_dir[BLZ_MAX_LIGHTS][3];
    float light_color[BLZ_MAX_LIGHTS][3];
    float light_intensity[BLZ_MAX_LIGHTS];
    float light_exposure[BLZ_MAX_LIGHTS];
    int light_samples[BLZ_MAX_LIGHTS];              /* read, unused (as the reference) */
    float density_scale, shadow_density_scale;      /* densityScale, shadowDensityScale */
    float volume_color[3], scattering[3];           /* volumeColor, scatteringColor */
    float absorption, gain, emission_scale;         /* absorption, gain, emissionScale */
    int ramp_size;                                  /* shaderColorRamp.count (<= BLZ_MAX_RAMP) */
    float ramp_keys[BLZ_MAX_RAMP];
    float ramp_colors[BLZ_MAX_RAMP][3];
    float temp_range[2];                            /* shaderTempRange */
    float ms_scale;                                 /* multiScatterScale */
    float ms_mask[2];                               /* multiScatterDensityMask */
    float ms_color[3];                              /* multiScatterColor */
    int ms_blur_iter;                               /* multiScatterBlurIter */
} blz_render_settings;
int blz_render_settings_default(blz_render_settings* r);
/* The render keys of a .blz (src/Parser.h:109-170); all required. */
int blz_parse_render(const char* path, blz_render_settings* out);
/* Render the solver's current temperature (RenderEngine::render): rgb
   receives width*height*3 floats, pixel (i, j) at 3*(i + width*j). Width
   and height must be multiples of 8 (as the reference's 8x8 launch). */
int blz_render(blz_solver* h, const blz_render_settings* r, float* rgb);
/* blz_render without any allocation: no counterpart in the reference; for
   callers that cannot allocate device memory (code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc). The
   scatter volumes and the device frame buffer live in the library's static
   per-device scratch of 8M floats (32 MiB, held until the process ends):
   2 * (W/4)(H/4)(D/4) + width*height*3 floats must fit, else BLZ_E_INVALID.
   Same image as blz_render. Every call on a device shares that device's
   scratch: calls that overlap on one device race. */
int blz_render_noalloc(blz_solver* h, const blz_render_settings* r, float* rgb);

/* Frame writer (RenderEngine::writeToDisk, src/RenderEngine.cu:367-476): by
   extension .ppm (ASCII P3, int(255.99 v)), .png (8-bit RGBA, (int)(255 v)
   clamped) or .exr (FLOAT B, G, R, uncompressed); rgb as blz_render. */
int blz_write_image(const char* path, const float* rgb, int width, int height);

/* ---- learned warm start (no reference counterpart; the reference solve
   always starts from the previous pressure, src/FluidSolver.cu:517-599) ----
   A 3-level 3D U-Net (16/32/64 channels, BF16 weights, FP32 accumulation)
   on a coarser grid (F = 4, doubled until at most 32 cells per axis) maps
   (residual of the current pressure, solid mask) to a pressure correction,
   followed by an exact Galerkin solve on at most 16^3 blocks (the global
   correction). Network grids are (batch, D, H, W) with D, H,
   W multiples of 4; inputs are two FP32 channels (residual, mask), each
   batch*D*H*W values, x fastest. */
typedef struct blz_net blz_net;
int blz_net_create(unsigned long long seed, blz_net** out); /* He-normal init from seed */
int blz_net_destroy(blz_net* net);
int blz_net_num_params(blz_net* net, long long* n);
int blz_net_get_params(blz_net* net, float* host);       /* FP32 master weights */
int blz_net_set_params(blz_net* net, const float* host);
int blz_net_get_grads(blz_net* net, float* host);        /* gradient of the last training step */
int blz_net_forward(blz_net* net, const float* in2, const int dims[4], float* out);
/* One step of mean-squared-error training: forward, backward, and (update
   != 0) an Adam step with learning rate lr; *loss = loss before the step. */
int blz_net_train_step(blz_net* net, const float* in2, const float* target, const int dims[4], float lr,
                       int update, double* loss);
/* Attach a network (NULL detaches). With warm_start != 0 and residual_tol
   > 0, each solve applies the corrections before its sweeps (and, for
   warm_start = N >= 2, after every N sweeps) and keeps them only if the
   FP64 residual after the next batch of sweeps beats the uncorrected path
   (measured on the first application, extrapolated from the last check on
   re-applications); a rejection continues exactly as without them. The
   network must outlive the attachment. */
int blz_set_warm_start(blz_solver* h, blz_net* net);
/* Which corrections a warm start applies (no reference counterpart):
   bit 0 the network's, bit 1 the global one (an exact Galerkin solve on a
   grid of at most 16^3 blocks, which sets the smoothest error modes),
   bit 2 an exact solve of the network grid's coarse problem in place of
   the network (an ablation: what a perfect prediction would gain; needs
   whole blocks and at most 32 coarse cells per axis), bit 3 the network's
   correction before the global one (default: the global one first, then
   the network's correction of the residual it leaves); default 3. The
   corrections of one application are judged together. */
int blz_set_warm_start_parts(blz_solver* h, int parts);
/* The network input for the current pressure and rhs (for generating
   training data): in2 = 2 x net_dims values; dims reported in (W, H, D). */
int blz_warm_start_input(blz_solver* h, float dt, float* in2, int net_dims[3], int coarse_dims[3]);
/* The global correction alone, applied to the current pressure without a
   judge: the residual the network's correction meets in a warm-started
   solve (blz_warm_start_input afterwards gives that input; for generating
   training data). */
int blz_warm_start_global(blz_solver* h, float dt);
/* No reference counterpart. The network's input as a correction of the
   default order meets it: the global correction's scale chosen and the
   residual it would leave, r - alpha_G A e_G, restricted and normalised,
   with the pressure left unchanged (in2 as for blz_warm_start_input). For
   checking the input the solver computes without adding the global
   correction first. */
int blz_warm_start_input_global(blz_solver* h, float dt, float* in2);
/* The correction step with a caller-given coarse prediction (network grid,
   x fastest) in place of the network's, scaled to the least error energy and
   followed by the global correction when enabled: FP64 residual before and
   right after, kept only if lower (else the pressure is restored). For
   ablations (an exact coarse solve in place of the network). */
int blz_warm_start_apply(blz_solver* h, float dt, const float* pred, double* before, double* after, int* kept);

/* ---- batched scenes (no reference counterpart; the reference steps one
   scene per process, src/main.cu:72-117) ----
   n independent scenes of any grid sizes; one blz_batch_step advances every
   scene by one step as a single replay of a captured HIP graph, with the
   pressure sweeps and residual checks of all scenes in single launches
   (with a sweep cap above 160, a step is a short sequence of graphs: the
   later sweeps in graphs of 160, skipped once every scene's solve has
   ended, then the pressure update and advection).
   Each scene computes what blz_step computes for it alone. Scenes need
   residual_tol > 0, cfl_number 0 and one even residual_check_every that
   divides their max_iter_solve. Sourcing (blz_add_source on the scene
   handles) happens between steps. */
typedef struct blz_batch blz_batch;
#define BLZ_BATCH_TUNE 10 /* batch autotuner candidates */
typedef struct blz_batch_stats {
    int scenes;
    int variant;              /* fused sweep variant in use (blz_set_rbgs_variant numbering) */
    int graph_nodes;          /* nodes of one captured step (all its graphs) */
    int graphs_run;           /* graphs the last blz_batch_step launched */
    long long voxels;         /* cells over all scenes */
    double capture_ms;        /* host time spent capturing and instantiating */
    double step_ms;           /* wall time of the last blz_batch_step replay */
    int tune_variant[BLZ_BATCH_TUNE]; /* autotuner candidates (fused and row-pair sweeps) ... */
    double tune_ms[BLZ_BATCH_TUNE];   /* ... and their time per sweep over the batch (0 if not tuned) */
} blz_batch_stats;
int blz_batch_create(const blz_settings* settings, int n, blz_batch** out);
int blz_batch_destroy(blz_batch* b);
/* Borrowed handle of scene i (sourcing, fields, stats); do not destroy. */
int blz_batch_solver(blz_batch* b, int i, blz_solver** out);
int blz_batch_set_warm_start(blz_batch* b, blz_net* net);
/* Time `sweeps` sweeps of each candidate tile configuration (slab groups in
   flight x waves per block) over the whole batch, keep the fastest; the
   pressure fields are restored afterwards. */
int blz_batch_autotune(blz_batch* b, int sweeps, int* variant);
/* Use one of the autotuner's candidates (blz_batch_stats.tune_variant) for
   the batch's sweeps (no reference counterpart; every candidate gives the
   same bits). */
int blz_batch_set_variant(blz_batch* b, int variant);
int blz_batch_step(blz_batch* b, float time);
int blz_batch_get_stats(blz_batch* b, blz_batch_stats* out);
/* Self-test of the pressure solve's division by a neighbour count d in
   3..6: runs the solver's device division on every 32-bit float and counts
   results that differ from IEEE a / d (expected 0). */
int blz_selftest_division(int divisor, unsigned long long* mismatches);

const char* blz_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* BLAZE_H */
