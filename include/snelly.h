/* snelly.h -- C ABI of libsnelly_hip.so, the MI355X (gfx950) implementation of
 * Snelly's path-tracing hot path.
 *
 * Every entry point replaces one WebGL2 operation that the reference's JS host
 * performs for its shader programs (the "GLU" layer, js/gl.js, driven by
 * js/renderer.js and js/snelly.js).  Handles are opaque; every call returns 0
 * on success or a negative SNL_E_* code, with a message from snl_last_error().
 * All calls on one context must come from one host thread (the reference's JS
 * is single-threaded too).  Nothing in this header uses torch or HIP types.
 */
#ifndef SNELLY_H
#define SNELLY_H

#ifdef __cplusplus
extern "C" {
#endif

#define SNL_OK 0
#define SNL_E_ARG -1      /* bad argument */
#define SNL_E_HIP -2      /* HIP runtime error */
#define SNL_E_COMPILE -3  /* scene code failed to translate or compile */
#define SNL_E_NOMEM -4    /* device allocation failed */
#define SNL_E_UNSUPPORTED -5

/* uniform / value type codes (GLSL types) */
#define SNL_T_FLOAT 1
#define SNL_T_VEC2 2
#define SNL_T_VEC3 3
#define SNL_T_VEC4 4
#define SNL_T_INT 5
#define SNL_T_MAT2 6
#define SNL_T_MAT3 7
#define SNL_T_MAT4 8
#define SNL_T_SAMPLER2D 9
#define SNL_T_SAMPLER3D 10

/* program build flags */
#define SNL_BUILD_PRECISE 1  /* no FMA contraction (-ffp-contract=off) */
#define SNL_BUILD_FAST 2     /* approximate (1-ulp) float divide and sqrt outside the RNG; the scene
                                code divides by multiplying with a 1-ulp reciprocal; subnormal
                                results flush to zero */
/* Renderer.renderMode (js/renderer.js:334-352): the program renders with the
 * ambient-occlusion integrator (shaders/ao-fragment-shader.glsl) or the
 * simplified path tracer (shaders/simplepathtracer-fragment-shader.glsl)
 * instead of the path tracer; both are RGB integrators, DISPERSION_ENABLED
 * does not apply.  At most one of the two. */
#define SNL_BUILD_AO 4
#define SNL_BUILD_PTSIMPLE 8

typedef struct snl_ctx_s* snl_ctx;
typedef struct snl_prog_s* snl_prog;
typedef struct snl_tex_s* snl_tex;
typedef struct snl_state_s* snl_state;
typedef struct snl_batch_s* snl_batch;
typedef struct snl_graph_s* snl_graph;

/* Context on one GPU: its stream and last-error string.
 * Replaces the WebGL2 context set-up of GLU (js/gl.js:16-29). */
int snl_ctx_create(int device, snl_ctx* out);
int snl_ctx_destroy(snl_ctx ctx);
const char* snl_last_error(snl_ctx ctx);   /* never NULL */
const char* snl_version(void);

/* Compile a program from what Renderer.compileShaders hands GLU.Shader
 * (js/renderer.js:217-359 -> js/gl.js:179-229): the scene snippet with its
 * default hooks (__SHADER__), the IOR code of the loaded materials
 * (__IOR_FUNC__), the #define block (__DEFINES__) and the integer tokens
 * __MAX_MARCH_STEPS__, __MAX_BOUNCES__, __MAX_ATMOSPHERE_SCATTERS__,
 * __MAX_SAMPLES_PER_FRAME__.  One program holds the pathtracer, normals, pick
 * and tonemapper kernels.  hipRTC output is cached on disk by source hash.
 * On a compile error the translated source's diagnostics are in log. */
int snl_program_build(snl_ctx ctx, const char* shader_glsl, const char* ior_glsl, const char* defines,
                      int max_march_steps, int max_bounces, int max_atmosphere_scatters,
                      int max_samples_per_frame, unsigned flags, snl_prog* out, char* log, int loglen);
/* Compile and cache only (no GPU needed); same inputs as snl_program_build. */
int snl_program_precompile(const char* shader_glsl, const char* ior_glsl, const char* defines,
                           int max_march_steps, int max_bounces, int max_atmosphere_scatters,
                           int max_samples_per_frame, unsigned flags, char* log, int loglen);
int snl_program_destroy(snl_prog prog);

/* GLU.Shader.uniformIndex (js/gl.js:255-260): -1 when the program has no such
 * uniform, so setters on unknown names are silent no-ops as in WebGL. */
int snl_uniform_index(snl_prog prog, const char* name);
/* GLU.Shader.uniformF/I/2F/3F/1Fv/2Fv/3Fv/4Fv/Matrix4fv (js/gl.js:262-406):
 * `data` holds `count` values of `type`; idx < 0 is ignored (returns 0). */
int snl_uniform_set(snl_prog prog, int idx, int type, const void* data, int count);
/* GLU.Shader.uniformTexture (js/gl.js:262-267) */
int snl_uniform_texture(snl_prog prog, int idx, snl_tex tex);

/* GLU.Texture(width, height, channels, isFloat, isLinear, isClamped, texels)
 * (js/gl.js:485-523); depth > 1 makes a 3-D texture.  texels may be NULL. */
int snl_texture_create(snl_ctx ctx, int width, int height, int depth, int channels, int linear, int clamp,
                       const float* texels, snl_tex* out);
/* GLU.Texture.copy (js/gl.js:503-507): width*height*depth*channels floats.
 * Submitted batch jobs of programs that sample the texture finish first (they
 * read the old texels). */
int snl_texture_upload(snl_tex tex, const float* texels);
/* Deletes the texture (gl.deleteTexture), after the submitted batch jobs of
 * programs that sample it; sampler uniforms still bound to it then read
 * (0, 0, 0, 1), as a texture unit without a texture does in WebGL. */
int snl_texture_destroy(snl_tex tex);
/* The texture's shape, for bindings that check buffer sizes (a WebGL texture
 * object carries its own; no C counterpart in the reference). */
int snl_texture_info(snl_tex tex, int* width, int* height, int* depth, int* channels);

/* PathtracerState pair (js/renderer.js:692-745): per-pixel radiance
 * (X, Y, Z, N) and RNG state, resident in HBM. */
int snl_state_create(snl_ctx ctx, int width, int height, snl_state* out);
int snl_state_destroy(snl_state st);
/* Radiance cleared to (0,0,0,1) as PathtracerState.clear does; RNG seeded per
 * pixel from PCG32(pixel_index ^ seed) (integer states, exactly reproducible). */
int snl_state_seed(snl_state st, unsigned seed);
int snl_state_clear(snl_state st);
/* which: 0 = radiance RGBA32F, 1 = RNG RGBA32F; 4*width*height floats, row 0 = bottom */
int snl_state_read(snl_state st, int which, float* out);
int snl_state_write(snl_state st, int which, const float* in);
/* The state's shape (out buffers above hold 4 * width * height * chains
 * floats), for bindings that check buffer sizes; no counterpart in the
 * reference, whose render targets carry their own size. */
int snl_state_info(snl_state st, int* width, int* height, int* chains);

/* Renderer.render's pathtracer draw (js/renderer.js:462-644) repeated for
 * `frames` frames: one path sample per pixel per frame per sample-per-frame. */
int snl_render(snl_prog prog, snl_state st, int frames);
/* snl_render without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The render's
 * queue counters and its tile-order buffers live in a file-scope device array
 * of the library (one per GPU, about 2.1 MB, held until the process ends)
 * instead of the state's buffers; everything else is the caller's program and
 * state.  Results are identical to snl_render(prog, st, frames).  Limits: the
 * image has at most SNL_NOALLOC_MAX_TILES 8x8 tiles (4096 x 4096; larger ->
 * SNL_E_ARG); programs that render with the wavefront engine or sample texture
 * unit 0 (an unbound sampler2D) -> SNL_E_UNSUPPORTED.  Every call on one GPU
 * shares that GPU's array, so calls that overlap on one GPU (from several
 * contexts or threads) race; calls on one context are ordered. */
#define SNL_NOALLOC_MAX_TILES 262144
int snl_render_noalloc(snl_prog prog, snl_state st, int frames);
/* the 'normals' render mode (normals-fragment-shader.glsl) */
int snl_render_normals(snl_prog prog, snl_state st, int frames);
/* Renderer.pick (js/renderer.js:398-460) for n pixel positions (gl_FragCoord
 * units): out[2k] = hit distance (maxLengthScale if none), out[2k+1] = material */
int snl_pick(snl_prog prog, const float* xy, int n, float* out);
/* tonemapper pass (js/renderer.js:622-642); rgb (3 floats/px) and/or rgba8 may be NULL */
int snl_tonemap(snl_prog prog, snl_state st, float* rgb, unsigned char* rgba8);
int snl_sync(snl_ctx ctx);

/* Timing and counters of the last snl_render call. */
typedef struct snl_stats {
  double ms_total;         /* wall time of the render call on the GPU (hipEvents) */
  double ms_trace;         /* summed time of the trace kernel launches */
  long long trace_steps;   /* SDF march steps executed (all rays) */
  long long paths;         /* camera paths started */
  int trace_launches;
  long long normal_evals;  /* surface normals computed (6 SDF evaluations each) */
} snl_stats;
int snl_get_stats(snl_prog prog, snl_stats* out);
/* launch-shape knobs: trace grid (blocks), jobs fetched per wave, shade grid;
 * 0 keeps the current value.  Also set by the autotuner. */
int snl_set_tuning(snl_prog prog, int trace_blocks, int trace_chunk, int shade_blocks);
/* 1 = time every trace launch with events (wavefront engine) */
int snl_set_profiling(snl_prog prog, int on);

/* Render engines.  WAVEFRONT: per bounce, a camera / trace / material-sorted
 * shade launch sequence over SoA path queues.  PERSISTENT: one launch per
 * pixel tile in which every lane runs whole paths from registers (no
 * per-bounce barrier) and a pixel's frames back to back; tiles run on separate
 * streams.  MEGAKERNEL: the same launch shape, every lane running cameraPath
 * as written, shadow and probe rays traced in place.  All draw the same random
 * numbers per path.  A program starts with the megakernel.  Programs with the
 * homogeneous atmosphere (HAS_ATMOSPHERE) always render with the megakernel;
 * those whose surface has sub-surface scattering use the wavefront engine when
 * the persistent one is chosen.  snl_render_batch runs the megakernel, or the
 * persistent kernel where that is chosen and the program can use it.  0 keeps
 * a knob. */
#define SNL_ENGINE_WAVEFRONT 0
#define SNL_ENGINE_PERSISTENT 1
#define SNL_ENGINE_MEGAKERNEL 2
int snl_set_engine(snl_prog prog, int engine, int tiles, int persist_blocks, int shade_thresh, int chunk);
/* persistent engine: march steps a tracing lane takes between two checks of
 * the shading gate (1..64, default 4); results do not depend on it */
int snl_set_march_burst(snl_prog prog, int steps);
/* The megakernel asks for four waves per SIMD; where that build spills
 * registers the program renders with a build without the bound (fewer waves,
 * no spills).  snl_mega_variants is 2 when both builds exist (else 1, or a
 * negative code); snl_set_mega_variant(prog, 1) renders with the four-wave
 * build that spills (loaded on first use), 0 with the default.  Same code and
 * arithmetic: results do not depend on the variant, only the speed does (the
 * autotuner times both).  No counterpart in the reference (launch tuning). */
int snl_mega_variants(snl_prog prog);
int snl_set_mega_variant(snl_prog prog, int variant);

/* Batched rendering: frames[j] frames of every job j = (progs[j], states[j]),
 * each with its program's current uniforms, submitted together -- the jobs'
 * launches are spread over the context's batch streams
 * (SNL_BATCH_STREAMS, default 16), longest first, so that many small images
 * fill the device at once; they run concurrently up to the process's hardware
 * queue count (GPU_MAX_HW_QUEUES).  The reference renders
 * one scene per canvas per frame (js/renderer.js:462-644); per job the result
 * is identical to snl_render(progs[j], states[j], frames[j]).  All programs
 * belong to one context; every job has its own state and its own program (a
 * program's uniforms are the job's; a repeated one is SNL_E_ARG).  Blocks until the batch
 * is done; `stats` (may be NULL) receives the batch's wall time and summed
 * counters. */
int snl_render_batch(int njobs, const snl_prog* progs, const snl_state* states, const int* frames,
                     snl_stats* stats);

/* The batch without blocking (continuous batching; SURVEY's
 * snl_batch_submit / snl_job_poll -- no counterpart in the reference, which
 * draws one canvas).  snl_batch_submit uploads each program's uniforms,
 * queues the jobs' launches as snl_render_batch does and returns at once with
 * a handle.  Jobs of batches submitted while earlier ones run are queued on
 * the least loaded streams, counting the unfinished jobs of those batches, so
 * their work overlaps on the device.  Until job j completes, progs[j] and
 * states[j] are busy: any call that touches them (render, read, write, seed,
 * error, tonemap, destroy, a later submit) first waits for that job, so results
 * never depend on the timing; uniform setters only change the host copy, which
 * the next render or submit uploads.
 *   snl_job_poll(b, j): 1 when job j has completed, 0 while it runs (j = -1:
 *     the whole batch), < 0 on an error;
 *   snl_batch_wait(b, stats): blocks until every job has completed; stats (may
 *     be NULL) as for snl_render_batch, ms_total from the submit to the last
 *     job's end;
 *   snl_batch_job_stats(b, j, out): waits for job j; its time from the submit
 *     to its end and its counters;
 *   snl_batch_release(b): waits for the batch and frees the handle (a context
 *     releases its batches when it is destroyed).
 * Per job the result is identical to snl_render(progs[j], states[j], frames[j]). */
int snl_batch_submit(int njobs, const snl_prog* progs, const snl_state* states, const int* frames,
                     snl_batch* out);
int snl_job_poll(snl_batch batch, int job);
int snl_batch_wait(snl_batch batch, snl_stats* stats);
int snl_batch_job_stats(snl_batch batch, int job, snl_stats* out);
int snl_batch_release(snl_batch batch);

/* Batch modules: one code object holding the megakernel of several programs,
 * so that a batch of jobs of different scenes, resolutions and frame counts
 * runs as one persistent launch (north-star batching; the reference draws one
 * scene per canvas).  snl_batch_module builds (or fetches from the code-object
 * cache) the module of the given programs' kinds -- a kind is a program's
 * translated scene with its #defines and __MAX_*__ limits; programs that
 * differ only in uniforms share one.  The programs must be path-tracer
 * programs of one build group (dispersion, interactive mode, PRECISE/FAST).
 * While batch fusion is on (snl_ctx_set_batch_fusion, default on), a
 * submitted batch's jobs that one loaded module covers -- megakernel programs
 * that do not sample texture unit 0 -- run in that module's persistent launch,
 * heaviest first: its waves take pixels from one job until its queue is
 * empty, then from the next; the other jobs run as before.  Per job the
 * result is identical to snl_render.  snl_batch_module_precompile compiles the
 * module of programs given as snl_program_build's sources into the cache,
 * without a device. */
int snl_batch_module(snl_ctx ctx, int nprogs, const snl_prog* progs);
int snl_batch_module_precompile(int nprogs, const char* const* shaders, const char* const* iors,
                                const char* const* defines, const int* max_march, const int* max_atm,
                                unsigned flags);
int snl_ctx_set_batch_fusion(snl_ctx ctx, int on);

/* A batch captured as a hipGraph (no counterpart in the reference): the jobs'
 * whole launch sequence -- counters zeroed by kernels (no memset nodes), the
 * tile-order passes, the batch module's persistent launch where fusion is on,
 * the other jobs' launches over the batch streams -- recorded once by
 * snl_batch_graph_create and replayed by snl_batch_graph_launch, which first
 * uploads the programs' current uniforms and blocks until the replay is done.
 * Each replay renders frames[j] more frames of every job, exactly as
 * snl_render_batch would.  Programs that sample texture unit 0 (a snapshot
 * per frame) are refused (SNL_E_UNSUPPORTED).  The programs and states must
 * outlive the graph. */
int snl_batch_graph_create(int njobs, const snl_prog* progs, const snl_state* states, const int* frames,
                           snl_graph* out);
int snl_batch_graph_launch(snl_graph graph, snl_stats* stats);
int snl_batch_graph_destroy(snl_graph graph);

/* Served rendering support (no counterpart in the reference, which renders one
 * interactive canvas).  A state of K chains holds K independent accumulations
 * of one width x height image: pixel p of chain k is state pixel k*w*h + p,
 * rendered like pixel p of the image from its own seed (PCG32 of its state
 * index).  One launch renders all chains, so small images still fill the
 * device, and the chains' spread gives the error of their mean:
 *   out[0] = sum over pixels and X,Y,Z of var_k(u_k) / K,
 *   out[1] = sum of m^2,
 * where u_k are the chains' unbiased means XYZ * N / (N - 1) and m their mean;
 * the relative RMS error of the mean image is about sqrt(out[0] / out[1]).
 * Computed on the device; only the two sums are copied back.  snl_state_read
 * returns all chains, chain after chain. */
int snl_state_create_chains(snl_ctx ctx, int width, int height, int chains, snl_state* out);
int snl_state_error(snl_state st, double* out);
/* free and total device memory of the context's GPU (hipMemGetInfo), for
 * admission control */
int snl_ctx_mem_info(snl_ctx ctx, unsigned long long* free_bytes, unsigned long long* total_bytes);
/* device bytes of a state of `chains` width x height images rendered by the
 * megakernel, the persistent engine or snl_render_batch: radiance and RNG,
 * counters, the tile-order buffers and the error-estimate partials (the
 * wavefront engine adds its path buffers on first use) */
unsigned long long snl_state_bytes(int width, int height, int chains);
/* the same for a state rendered by `prog`: adds the 16 B/px radiance snapshot
 * of programs that sample texture unit 0 (an unbound sampler2D, such as a
 * tabulated metal's kTex, js/materials.js:664-665) */
unsigned long long snl_program_state_bytes(snl_prog prog, int width, int height, int chains);

#ifdef __cplusplus
}
#endif
#endif
