/*
 * radgpu.h - C ABI of the MI355X rtrace / rcontrib / dctimestep hot path.
 *
 * Plain pointers and sizes only.  Every call returns 0 on success and a
 * negative status on failure; rg_last_error() then returns the message.
 * Nothing in the library calls exit(): the reference's error() -> quit()
 * chain (common/rterror.h:30-35) becomes a status code here.
 *
 * A context owns one scene, its device tables and BVH, and one HIP stream.
 * Contexts are independent (the reference allows one simulation manager per
 * process, rt/RtraceSimulManager.cpp:237); calls on one context must not
 * overlap.
 */
#ifndef RADGPU_H
#define RADGPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rg_ctx rg_ctx;

/* Ray-tracing options: the subset of RAYPARAMS (rt/ray.h:157-190) the GPU
 * path honours.  Field meaning and units follow the reference options. */
typedef struct rg_params {
  int ambounce;        /* -ab */
  int ambdiv;          /* -ad */
  int maxdepth;        /* -lr */
  int backvis;         /* -bv */
  double minweight;    /* -lw */
  double dstrsrc;      /* -dj */
  double srcsizerat;   /* -ds (local sources are sampled unpartitioned) */
  double specthresh;   /* -st */
  double specjitter;   /* -ss */
  double ambval[3];    /* -av */
  int directvis;       /* -dv */
  int contrib;         /* rcontrib -V */
  unsigned long long seed;
} rg_params;

/* rcontrib modifier with its bin function (rt/rcontrib.c:92 addmodifier).
 * bin_kind: 0 constant bin `cbin`, 1 Reinhart rbin (util/reinhartb.cal),
 * 2/3/4 Klems full/half/quarter kbin (util/klems_*.cal), 5 rfluxmtx uniform
 * surface receiver, bin 0 when -D.nrm > 0 (util/rfluxmtx.c:465-470),
 * 6 Shirley-Chiu scbin with SCdim in `mf` (util/disk2square.cal),
 * 7 CIE sky-scanner cbin (util/cieskyscan.cal). */
typedef struct rg_modifier {
  const char *name;
  int bin_kind;
  int nbins;
  int cbin;
  double mf;           /* Reinhart MF, Shirley-Chiu SCdim */
  double rhs;          /* RHS: +1 right-handed */
  double nrm[3];       /* rNx,rNy,rNz (Reinhart, SC, CIE, uniform) / Nx,Ny,Nz (Klems) */
  double up[3];        /* Ux,Uy,Uz */
} rg_modifier;

/* Replaces process start-up + ray_init() (rt/raycalls.c:172-200). */
int rg_ctx_create(rg_ctx **out, int device);
void rg_ctx_destroy(rg_ctx *ctx);
/* Message of the last failed call on ctx (or of rg_ctx_create if ctx==NULL). */
const char *rg_last_error(const rg_ctx *ctx);

/* Scene input, replaces readoct() (common/readoct.c:36) and readobj()
 * (common/readobj.c:28): `path` is an octree or a scene file.  Inline
 * "!command" lines run only when allow_exec is non-zero. */
int rg_scene_load(rg_ctx *ctx, const char *path, int allow_exec);
/* Appends scene text, as readobj() on a stream. */
int rg_scene_load_text(rg_ctx *ctx, const char *text, int allow_exec);
/* Compiles the objects, uploads them and builds the BVH on the GPU
 * (replaces marksources(), rt/source.c:47, and the octree). */
int rg_scene_commit(rg_ctx *ctx);
/* info[0..5] = objects, primitives, materials, sources, BVH nodes, warnings */
int rg_scene_info(rg_ctx *ctx, long long *info, int n);

/* Replaces getrenderopt() option state (rt/renderopts.c:124). */
int rg_set_params(rg_ctx *ctx, const rg_params *p);
/* Library defaults of rt/raycalls.c:118-153 (what rtrace uses). */
void rg_default_params(rg_params *p);
/* rcontrib defaults after override_options() (rt/rcontrib.c:24-58, rt/rcmain.c:164-171). */
void rg_default_params_rcontrib(rg_params *p);

/* rtrace: one value per ray (rt/rtrace.c:436 rtcompute).
 * rays: n x 6 doubles (origin, direction).  flags: 8 = -I (RTimmIrrad of
 * rt/RtraceSimulManager.h:86), 16 = -i.  out: n x 3 doubles (RGB value). */
int rg_trace(rg_ctx *ctx, const double *rays, long long n, int flags, double *out);

/* rcontrib: contributions binned per modifier (rt/rcontrib.c:272-338,
 * rt/rc2.c:294).  rays: nrec*accum x 6 doubles; each record averages
 * `accum` consecutive rays.  out: nrec x (3 * sum of nbins) doubles in
 * modifier order.  flags as rg_trace. */
int rg_contrib(rg_ctx *ctx, const double *rays, long long nrec, int accum, int flags,
               const rg_modifier *mods, int nmods, double *out);

/* rg_contrib with the rays and the output in device memory of the context's
 * device (the in-process manager's resident buffers, rt/RcontribSimulManager
 * get_output_array, src/binding/radiance_ext.cpp:346-362): no host copies.
 * The rays are read only after all work queued before the call on the
 * caller's stream (rg_set_stream; the null stream by default) has finished,
 * so a kernel that has just written them on that stream is complete; work
 * on other streams must be synchronized by the caller.  The call returns
 * when the output is written. */
int rg_contrib_device(rg_ctx *ctx, const double *d_rays, long long nrec, int accum, int flags,
                      const rg_modifier *mods, int nmods, double *d_out);
/* The caller's HIP stream (a hipStream_t; NULL: the null stream) that the
 * device-memory entry points order their input reads after (no reference
 * counterpart: the reference has no device buffers). */
int rg_set_stream(rg_ctx *ctx, void *stream);

/* Batched rcontrib: the records of nctx scenes (one committed context each,
 * same device) traced in one wavefront.  rays holds scene 0's nrec[0]*accum
 * rays, then scene 1's, ...; out holds the records in the same order.
 * Parameters and the ray counter are those of ctxs[0].  At most 4096
 * scenes per call. */
int rg_contrib_batch(rg_ctx *const *ctxs, int nctx, const double *rays, const long long *nrec, int accum,
                     int flags, const rg_modifier *mods, int nmods, double *out);

/* Batched rtrace: n[i] rays of scene i, concatenated; out in the same order
 * (at most 4096 scenes per call). */
int rg_trace_batch(rg_ctx *const *ctxs, int nctx, const double *rays, const long long *n, int flags, double *out);

/* Materials of the committed scene (no reference counterpart: the reference
 * changes a material by editing the scene file and rebuilding the octree).
 * rg_material_index: index of the material object `name`.  rg_material_get:
 * its type (0 plastic, 1 metal, 2 trans, ...) and real arguments.
 * rg_material_set: replace the real arguments (same count) on the device. */
int rg_material_index(rg_ctx *ctx, const char *name, int *index);
int rg_material_get(rg_ctx *ctx, int index, int *kind, double *f, int n);
int rg_material_set(rg_ctx *ctx, int index, const double *f, int n);

/* Path-replay adjoint of rg_contrib (no reference counterpart).  For the
 * loss whose derivative with respect to each record's binned irradiance
 * E[rec][c] = sum_col out[rec][col][c] * lcol[col][c] is dlde[rec][c]
 * (nrec x 3), adds into grad[i][0..6] (nmats x 7; zeroed here) the
 * derivative with respect to material mats[i]'s real arguments: red, green,
 * blue reflectance, specularity, roughness, and for trans the
 * transmissivity and transmitted specularity (columns 5 and 6, zero for
 * plastic and metal).  Plastic, metal and trans materials (rt/normal.c
 * m_normal) can be differentiated.  It traces the same rays with the same
 * random numbers as rg_contrib.  lcol: 3 * sum of nbins doubles (radiance
 * of every output column).  Every deposit is differentiated, shadow-ray
 * (direct light) deposits included.  Sampling decisions are held fixed, and
 * a rough material's Gaussian lobe samples add their likelihood-ratio score
 * for the roughness (rt/normal.c:364-495 gaussamp): the roughness column is
 * exact in expectation, not sample by sample.  That score needs untruncated
 * lobe samples: with a rough differentiated material, -ss in [0, 1) is an
 * error. */
int rg_contrib_adjoint(rg_ctx *ctx, const double *rays, long long nrec, int accum, int flags,
                       const rg_modifier *mods, int nmods, const double *dlde, const double *lcol,
                       const int *mats, int nmats, double *grad);

/* First intersection of each ray, as rtrace -oLN prints it
 * (rt/rtrace.c:758-804): t[i] distance (1e10 on a miss), nrm[i] the
 * surface's unflipped normal (zero on a miss), prim[i] the primitive index in
 * scene order (-1 on a miss).  Any output pointer may be NULL. */
int rg_first_hit(rg_ctx *ctx, const double *rays, long long n, double *t, double *nrm, int *prim);

/* Ray segments traced by the last rg_trace / rg_contrib (closest-hit
 * queries, the reference's nrays of rt/raytrace.c:606). */
int rg_last_nrays(rg_ctx *ctx, unsigned long long *nrays);

/* Measurement support (no reference counterpart).  level 0: off, 1: time
 * every generation kernel with HIP events on the context stream, 2: also
 * count BVH nodes visited and primitives tested per ray. */
int rg_set_profiling(rg_ctx *ctx, int level);

/* Spatial index builder used by the next rg_scene_commit (no reference
 * counterpart: the reference indexes scenes with an octree,
 * rt/raytrace.c:595-760).  RG_BVH_AUTO: host binned-SAH tree up to 4M
 * primitives, device LBVH above; RG_BVH_SAH / RG_BVH_LBVH force one.  Both
 * give the same closest hits; only the traversal cost differs. */
/* Launch shape of the renderer (no reference counterpart).  Keys:
 * "trace_kernel" 0 one child ray per thread, 1 persistent wavefronts with
 * dynamic ray fetch (default); "refill" idle lanes (1..64) that make a
 * persistent wavefront fetch new rays (default 16); "window" children per
 * generation launch (default 2M); "blocks_per_cu" persistent grid size;
 * "max_big" (0..32, default 8, applies at the next commit): room-sized
 * primitives are kept out of the BVH and tested before the walk when there
 * are at most this many, else all go into the BVH; "wide" 1 (default) the
 * persistent walk fetches four child boxes per node when the scene has a
 * four-wide tree (host SAH trees shallow enough for the traversal stack),
 * 0 the binary walk; "shadow_cache" 1 (default) shadow rays are first
 * checked against the blocker last recorded for their source and cell, as
 * the reference's shadow cache does (rt/srcobstr.c:28-339, filled by cast
 * rays at the first call after a commit; those rays count once, in the call
 * that first shadow-tests their source, as in the reference's counter), 0
 * every shadow ray is traced; "child_order" 1 (default) hands hemisphere
 * children to wavefronts in 8 x 8 blocks of strata, 0 by stratum row (same
 * samples); "inflight" 1 (default) runs one window at a time, 2 queues the
 * next window before reading back the counts of the one before (the GPU
 * then waits less for the host, but the traversal slows by more on config 2:
 * 53.2 against 52.7 ms per step).
 * Results do not depend on these; only the time does (the walks visit
 * primitives in different orders, which only matters for two surfaces
 * within FTINY of each other along a ray, where the reference's own choice
 * depends on its octree order). */
int rg_set_tuning(rg_ctx *ctx, const char *key, long long value);

#define RG_BVH_AUTO 0
#define RG_BVH_SAH 1
#define RG_BVH_LBVH 2
int rg_set_bvh_builder(rg_ctx *ctx, int kind);
/* out[0..8] = rays, generation launches, threads launched, summed traversal-
 * kernel milliseconds, nodes visited, primitives tested, batches, summed
 * shading-kernel milliseconds, bytes per node visited (64 binary, 128
 * four-wide) */
int rg_last_stats(rg_ctx *ctx, double *out, int n);

/* DC x sky product, replaces cm_multiply() (util/cmatrix.c:421-474):
 * C[M][N][3] = A[M][K][3] . B[K][N][3], each channel independently.
 * Host pointers.  mode 0 = FP32 accumulate, 1 = BF16 hi/lo split MFMA
 * (K <= 160: rows of A held in registers, sky columns streamed through LDS;
 * K > 256: LDS-staged 128 x 128 tiles; otherwise per-wavefront operand
 * loads), 2 / 3 / 4 = the split product forced to per-wavefront loads /
 * LDS-staged tiles / the short-K form (K <= 160).  The split planes live in
 * a buffer kept per device and stream, grown as needed and held until the
 * process ends (2 * 3 * (M + N) * ceil(K / 32) * 32 * 2 bytes at most, with
 * K rounded up to 64 for the staged form). */
int rg_dcsky(const float *A, const float *B, float *C, int M, int K, int N, int mode);
/* Same with device pointers on the current HIP device. */
int rg_dcsky_device(const float *A, const float *B, float *C, int M, int K, int N, int mode);

/* A view after setview() (common/image.c:23-118): type 'v' 'l' 'c' 'a' 'h'
 * 's', unit direction, horizontal and vertical vectors scaled by the view
 * size, their squared half sizes, offsets and clipping distances. */
typedef struct rg_view {
  int type;
  double vp[3], vdir[3], hvec[3], vvec[3];
  double hn2, vn2, horiz, vert, hoff, voff, vfore, vaft;
} rg_view;

/* A picture: one primary ray per pixel generated on the GPU from the view
 * (rpict -ps 1, rt/rpict.c:668-712, pixel jitter pj), traced as rg_trace
 * does; out[yres][xres][3], top scanline first.  Pixels the view has no ray
 * for are 0.  flags as rg_trace. */
int rg_trace_view(rg_ctx *ctx, const rg_view *v, int xres, int yres, double pj, int flags, double *out);

/* gendaymtx (gen/gendaymtx.c): annual Perez sky matrix from WEA or EPW
 * weather text.  The host reads the tape and sets up each step; the GPU
 * evaluates the sky patches.  Options as the reference's flags. */
typedef struct rg_skyopts {
  int mfactor;            /* -m: Reinhart subdivision (1 = 145 sky patches + ground) */
  double sky_color[3];    /* -c; 0 0 0 with -d / -n (sun only) */
  double ground_color[3]; /* -g; 0 0 0 with -d / -n */
  double sun_color[3];    /* 1 1 1; 0 0 0 with -s (sky only) */
  double rotation;        /* -r degrees */
  int nsuns;              /* patches the sun is spread over: 4; 1 with -5 */
  double fixed_sun_sa;    /* -5: solar solid angle (sr), <= 0 otherwise */
  int output;             /* -O0 visible radiance, -O1 solar radiance */
  int sun_hours_only;     /* -u */
  int average;            /* -A */
} rg_skyopts;
typedef struct rg_sky rg_sky;

/* Reads the weather data and prepares every time step (replaces the
 * gendaymtx process up to its matrix).  On failure returns -1 with the
 * message in rg_sky_error(NULL). */
int rg_sky_create(rg_sky **out, const char *weather_text, const rg_skyopts *opts);
void rg_sky_destroy(rg_sky *s);
const char *rg_sky_error(const rg_sky *s);
/* out[5]: sky patches (rows, ground first), stored columns (time steps, 1
 * with -A), stored steps, steps counted, suns placed (for -D). */
int rg_sky_shape(const rg_sky *s, int *out);
/* latitude north, longitude east, degrees (LATLONG header) */
int rg_sky_site(const rg_sky *s, double *latlong);
/* per placed sun: id (solar minute) and good flag; altitude, azimuth
 * (radians) and direct illuminance, as gendaymtx -D writes them */
int rg_sky_suns(const rg_sky *s, int *id_good, double *alt_azi_illum);
/* The matrix [patch][column][rgb] in FP32 into a host buffer, or into device
 * memory of the current device (the B operand of rg_dcsky_device). */
int rg_sky_matrix(rg_sky *s, float *out);
int rg_sky_matrix_device(rg_sky *s, float *dev_out);
/* rg_sky_matrix_device without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * step tables, patch tables and luminance plane the kernels use live in a
 * fixed 96 MiB device array that each GPU holds until the process ends
 * (Reinhart MF:4 over 8,760 steps fits, -A over that does not); a larger
 * plan returns -1 with the message in rg_sky_error(s).  Calls on one GPU
 * share the array and are serialised by the library; the result equals
 * rg_sky_matrix_device's. */
int rg_sky_matrix_device_noalloc(rg_sky *s, float *dev_out);

/* ASCII records on the host cores: nrecords records of per_record values,
 * each value printed with fmt ("%e" or "%g", optional ".N", N <= 17),
 * values separated by sep, then tail and '\n', and one more '\n' after every
 * `group` records (group 0: none).  Replaces the per-value printf of the
 * reference's writers: rcontrib's "%.6e\t" records (rt/rc2.c:306),
 * rtrace's "%e\t" values (rt/rtrace.c:912-916), gendaymtx's
 * "%.3g %.3g %.3g\n" rows with a blank line after each patch
 * (gen/gendaymtx.c:652-663).  Returns the
 * bytes written to out, -1 for bad arguments, or -(2 + bytes needed) when
 * cap is too small.  nthreads <= 0: one per host core. */
long long rg_format_ascii(const double *values, long long nrecords, int per_record, long long group,
                          const char *fmt, const char *sep, const char *tail, char *out, long long cap, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
