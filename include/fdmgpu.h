/* fdmgpu: MI355X (gfx950) fast Poisson solve -- C ABI.
 *
 * Drop-in boundary for the fdm fast Poisson solvers.  Plain pointers, sizes
 * and integer error codes; no C++ or torch types.  The header-only C++ shim
 * include/fdm/lapl_cube.h / include/fdm/lapl_rect.h rebuilds the reference's
 * templated class API (same constructors, same solve(ans, rhs)) on top of it.
 *
 * Array convention (every entry point): a field is dense, x fastest,
 * [nz][ny][nx] for a cube and [ny][nx] for a rect, interior points only --
 * the array the reference's solve(T* ans, T* rhs) receives (reference
 * src/lapl_cube.h:80, src/lapl_rect.h:65-68, src/tensor.h:192-205).
 * Element type: float for FDMGPU_F32 plans, double for FDMGPU_F64 and
 * FDMGPU_F32_REFINE64 plans.
 *
 * Threading: a plan is not reentrant (the reference's solver object is not
 * either: member scratch, src/lapl_cube.h:96-97); different plans may be used
 * concurrently from different threads on different streams.
 */
#ifndef FDMGPU_H
#define FDMGPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fdmgpu_plan_t* fdmgpu_plan;

/* ---- error codes --------------------------------------------------------
 * The reference has no return codes: it calls verify() -> fprintf + abort()
 * (src/verify.h:1-36).  The C++ shim maps every non-zero code back to that
 * behaviour. */
#define FDMGPU_OK 0
#define FDMGPU_ERR_INVALID_SIZE 1  /* transform length not a power of two: verify((1<<n)==N), src/fft.cpp:67,105 */
#define FDMGPU_ERR_INVALID_ARG 2   /* null plan/pointer, bad precision or mask */
#define FDMGPU_ERR_UNSUPPORTED 3   /* valid for the reference, not implemented on the GPU */
#define FDMGPU_ERR_OUT_OF_MEMORY 4 /* device allocation failed */
#define FDMGPU_ERR_DEVICE 5        /* HIP runtime error (no device, launch failure) */
#define FDMGPU_ERR_SINGULAR 6      /* singular LaplRect row system: gtsv info != 0, src/lapl_rect.cpp:91 */
#define FDMGPU_ERR_NOT_CONVERGED 7 /* FP32+refine did not reach the requested residual */
#define FDMGPU_ERR_ADMISSION 8     /* serving: request larger than the HBM budget */
#define FDMGPU_ERR_TOO_LARGE 9     /* fdmgpu_solve_batch_noalloc: scratch beyond the static scratch array */

/* ---- precision -----------------------------------------------------------*/
#define FDMGPU_F32 0          /* float in/out, float arithmetic: LaplCube<float>/LaplRect<float> */
#define FDMGPU_F64 1          /* double in/out, double arithmetic: LaplCube<double>/LaplRect<double> */
#define FDMGPU_F32_REFINE64 2 /* double in/out; FP32 transforms + tridiagonal, FP64 residual/update/norm */

/* ---- periodic axes (tensor_flag::periodic per axis, src/tensor.h:12-32) ---*/
#define FDMGPU_PERIODIC_X 1
#define FDMGPU_PERIODIC_Y 2
#define FDMGPU_PERIODIC_Z 4

/* Replaces LaplCube<T,check,F>::LaplCube(dx,dy,dz,lx,ly,lz,nx,ny,nz)
 * (src/lapl_cube.h:58-100) and its init_lm (src/lapl_cube.cpp:145-172).
 * Transform lengths nx+1, ny+1, nz+1 (Dirichlet) or nx, ny, nz (periodic) must
 * be powers of two.  periodic_mask: 0 (tensor_flags<>) or 7 (periodic^3), the
 * reference's instantiations (src/lapl_cube.cpp:174-182).  Periodic cubes
 * return the zero-mean solution of A u = f - mean(f) (the reference zeroes
 * the (0,0,0) mode, src/lapl_cube.cpp:82-84); with FDMGPU_F32_REFINE64 the
 * refinement certifies ||(f - mean f) - A u|| / ||f - mean f|| (requires
 * l = n h per axis). */
int fdmgpu_plan_cube(fdmgpu_plan* plan, int precision, int periodic_mask, double dx, double dy, double dz,
                     double lx, double ly, double lz, int nx, int ny, int nz);

/* Replaces LaplRect<T,check,F>::LaplRect(dx,dy,lx,ly,nx,ny) (src/lapl_rect.h:40-63).
 * periodic_mask: 0 or FDMGPU_PERIODIC_Y (tensor_flags<periodic>, F::head = y).
 * ny+1 (Dirichlet) / ny (periodic) must be a power of two; nx is arbitrary. */
int fdmgpu_plan_rect(fdmgpu_plan* plan, int precision, int periodic_mask, double dx, double dy, double lx,
                     double ly, int nx, int ny);

/* Replaces LaplRectFFT2<T,check,F>::LaplRectFFT2(dx,dy,lx,ly,nx,ny)
 * (src/lapl_rect.h:76-109; solve src/lapl_rect.cpp:113-206): sine or real
 * DFT transforms along both axes and a division by the eigenvalues.
 * periodic_mask: 0 (tensor_flags<>), FDMGPU_PERIODIC_Y (tensor_flags<periodic>)
 * or FDMGPU_PERIODIC_Y | FDMGPU_PERIODIC_X (periodic^2, whose zero mode gets
 * the reference's coefficient 1, src/lapl_rect.cpp:169-172); an x-periodic,
 * y-Dirichlet plan is accepted too.  ny+1 / ny and nx+1 / nx (Dirichlet /
 * periodic axis) must be powers of two up to 512.  FDMGPU_F32_REFINE64 only
 * for Dirichlet axes with l = (n+1) h.  fdmgpu_rect_set_scales sets
 * lm_y_scale (indexed by the x mode, entries x1..xn as the reference reads
 * them); L_scale and U_scale do not enter this solver. */
int fdmgpu_plan_rect_fft2(fdmgpu_plan* plan, int precision, int periodic_mask, double dx, double dy, double lx,
                          double ly, int nx, int ny);

/* Replaces LaplCyl3FFT2<T,check,zflag,use_cyclic_reduction>::LaplCyl3FFT2(dr, dz,
 * r0, lr, lz, nr, nz, nphi) (src/lapl_cyl.h:173-249; solve
 * src/lapl_cyl.cpp:11-128): the cylindrical Poisson solve on [nphi][nz][nr]
 * (r fastest; z indices 1..nz, or 0..nz-1 when z is periodic) by real DFTs
 * along phi, sine or real DFTs along z and, per (phi, z) mode, the r
 * tridiagonal system solved by cyclic reduction (the use_cyclic_reduction
 * instantiation; the default one factors the same systems with LAPACK
 * gttrf/gttrs).  periodic_mask: 0 (zflag none) or FDMGPU_PERIODIC_Z.  nphi and
 * nz+1 (nz periodic) powers of two up to 512, nr <= 4096.  FDMGPU_F32 or
 * FDMGPU_F64 (no FP32+refine: the residual stencil is not built). */
int fdmgpu_plan_cyl3_fft2(fdmgpu_plan* plan, int precision, int periodic_mask, double dr, double dz, double r0,
                          double lr, double lz, int nr, int nz, int nphi);

/* Replaces the public LaplRect members lm_y_scale, L_scale, U_scale
 * (src/lapl_rect.h:32, 57-59): arrays of nx+1 doubles indexed 1..nx (entry 0
 * unused, as in the reference); NULL keeps the default of all ones.  Row
 * systems that are all diagonally dominant are solved by cyclic reduction;
 * otherwise every row by Gaussian elimination with partial pivoting, as the
 * reference's ?gtsv (src/lapl_rect.cpp:90).  Returns FDMGPU_ERR_SINGULAR when
 * a row system meets a zero pivot (gtsv info != 0, where the reference
 * aborts, src/lapl_rect.cpp:91). */
int fdmgpu_rect_set_scales(fdmgpu_plan plan, const double* lm_y_scale, const double* L_scale,
                           const double* U_scale);

/* FDMGPU_F32_REFINE64 only: stop when ||f - A u||_2 / ||f||_2 <= rel_tol
 * (default 1e-10) or after max_iter FP32 inner solves (default 4).  No
 * reference counterpart (the reference solves in one precision). */
int fdmgpu_set_refine(fdmgpu_plan plan, double rel_tol, int max_iter);

/* Replaces LaplCube::solve / LaplRect::solve(T* ans, T* rhs)
 * (src/lapl_cube.cpp:9-142, src/lapl_rect.cpp:63-110): host pointers,
 * synchronous, ans fully overwritten, rhs not modified, ans == rhs allowed. */
int fdmgpu_solve_host(fdmgpu_plan plan, void* ans, const void* rhs);

/* Same solve on device pointers, stream-ordered on `stream` (a hipStream_t,
 * NULL = default stream); returns after enqueueing.  ans == rhs allowed.
 * No reference counterpart (the reference has no device path). */
int fdmgpu_solve_device(fdmgpu_plan plan, void* d_ans, const void* d_rhs, void* stream);

/* `count` independent problems of the plan's shape laid out back to back
 * (stride = field size) in one launch sequence.  No reference counterpart. */
int fdmgpu_solve_batch_device(fdmgpu_plan plan, int count, void* d_ans, const void* d_rhs, void* stream);

/* The same batch with problem p at d_ans_tab[p] / d_rhs_tab[p]: DEVICE arrays
 * of `count` device pointers (problems in separate buffers).  For
 * FDMGPU_F32_REFINE64 plans every ans pointer must differ from its rhs
 * pointer.  No reference counterpart. */
int fdmgpu_solve_table_device(fdmgpu_plan plan, int count, void* const* d_ans_tab, const void* const* d_rhs_tab,
                              void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same solve as fdmgpu_solve_batch_device (same
 * kernels, bitwise the same results) with no device allocation.  The
 * FP32+refine scratch (and, for an in-place refine solve, the copy of the
 * right-hand side) comes from a static device array of 64 MiB that each GPU
 * the library runs on holds until the process ends; a solve that needs more
 * returns FDMGPU_ERR_TOO_LARGE before enqueueing anything (a 127^3 refine
 * problem needs about 16.5 MB, so three fit; plain F32/F64 solves need no
 * scratch).  Every call on one GPU shares that array: calls that overlap on
 * one GPU race (order them, e.g. on one stream).  fdmgpu_get_stats reports
 * the call until another call on the GPU overwrites the scratch. */
int fdmgpu_solve_batch_noalloc(fdmgpu_plan plan, int count, void* d_ans, const void* d_rhs, void* stream);

/* Kernel variants (tile shapes) of a plan, as the autotuner chooses them:
 * plane = 1 runs square planes that fit LDS through the plane kernel (0: x
 * and y tile kernels; 2: planes of N = 128 through a streaming plane kernel,
 * blocks walking their planes with the next plane's loads in flight -- one
 * block per compute unit for FP64, the default there; two for FP32 and the
 * FP32 transforms of FDMGPU_F32_REFINE64, a variant the autotuner may pick);
 * tile_x / tile_y = pencils per block of the tile
 * kernels (16, 32, 64; 0 = default); zcr_gmul = z cyclic reduction
 * layout: 1, 2 = lanes per system multiplier inside a wavefront, 3, 4 = one
 * wavefront per 32-row chunk of 64 consecutive modes, 64 (3) or 128 (4)
 * modes per block (Dirichlet z, 32 to 512 rows; 4 up to 128), 5 = the same
 * with 16-row chunks (16 to 256 rows); 0 = default
 * (Dirichlet: 5 up to 128 rows, 3 from 256 rows, else 5 where valid;
 * circulant: 2 from 64 rows, else 1); tile_y = 0 takes 64 for FP32 planes at
 * least 128 wide when two such blocks fit a CU, else 32 (16 where 32 does not
 * fit); fuse (FP32+refine plane path) = FP64 residual fused into the next
 * forward transform (1) or a separate residual pass (0, the default; plain
 * plans list 0 only).  Every variant solves the same system; results differ
 * at rounding level.  No reference counterpart. */
typedef struct fdmgpu_variant {
    int plane, tile_x, tile_y, zcr_gmul;
    int fuse; /* FP32+refine on the plane path: 1 = residual fused into the forward transform */
} fdmgpu_variant;
int fdmgpu_plan_variants(fdmgpu_plan plan, int max, fdmgpu_variant* out, int* count);
int fdmgpu_plan_set_variant(fdmgpu_plan plan, const fdmgpu_variant* variant);
int fdmgpu_plan_get_variant(fdmgpu_plan plan, fdmgpu_variant* variant);

/* Statistics of the plan's last solve (synchronises the stream it ran on).
 * iterations: FP32 inner solves (refine plans; 1 otherwise); rel_residual:
 * ||f - A u||/||f|| of the last problem (refine plans; -1 otherwise). */
typedef struct fdmgpu_stats {
    int iterations;
    double rel_residual;
    double max_rel_residual; /* over a batch */
} fdmgpu_stats;
int fdmgpu_get_stats(fdmgpu_plan plan, fdmgpu_stats* stats);

/* Device bytes the plan holds, and the extra bytes one solve of `count`
 * problems needs (scratch); used for HBM-footprint admission. */
int fdmgpu_plan_footprint(fdmgpu_plan plan, int count, unsigned long long* resident,
                          unsigned long long* per_solve);

/* Per-kernel timing with HIP events recorded around every launch on the
 * solve's stream (instrumentation for measurement; no reference counterpart --
 * the reference times its phases with std::chrono, src/ns_cube.cpp:27-62).
 * fdmgpu_set_timing(plan, 1) enables and resets the counters;
 * fdmgpu_get_timing synchronises and returns the summed milliseconds and the
 * launch count of one kernel class since the last reset. */
#define FDMGPU_K_XY_FWD 0   /* plane DST kernel, forward (x and y of whole z-planes) */
#define FDMGPU_K_XY_INV 1   /* plane DST kernel, inverse */
#define FDMGPU_K_X_TILE 2   /* x-axis DST tiles (planes too large for LDS) */
#define FDMGPU_K_Y_TILE 3   /* y-axis DST tiles */
#define FDMGPU_K_ZCR 4      /* z cyclic reduction */
#define FDMGPU_K_ROWCR 5    /* LaplRect row cyclic reduction */
#define FDMGPU_K_RESIDUAL 6 /* FP64 residual + norm + convergence flag */
#define FDMGPU_K_OTHER 7    /* flag fills, FP32->FP64 updates, copies */
#define FDMGPU_K_XY_FWD_RESID 8 /* FP64 residual fused into the forward plane DST (refinement solves >= 2) */
#define FDMGPU_K_XY_INV_ACC 9   /* inverse plane DST adding the correction into the FP64 iterate */
#define FDMGPU_K_RESIDUAL_RHS 10 /* FP64 residual writing the next inner solve's FP32 rhs (+ norm) */
#define FDMGPU_K_XY_FWD_R32 11   /* forward plane DST of that FP32 rhs (refinement solves >= 2) */
#define FDMGPU_NUM_KERNELS 12
int fdmgpu_set_timing(fdmgpu_plan plan, int enable);
int fdmgpu_get_timing(fdmgpu_plan plan, int kernel, double* total_ms, int* launches);
/* The same with events around the launches of the kernel classes in `mask`
 * only (bit k = class k): a measurement of one class's launches that leaves
 * the others uninstrumented.  fdmgpu_set_timing(plan, e) is mask ~0 or 0. */
int fdmgpu_set_timing_kernels(fdmgpu_plan plan, unsigned mask);

/* ---- heterogeneous batches and serving ------------------------------------
 * No reference counterpart: the reference solves one problem per call
 * (LaplCube::solve).  A problem names its solver (the constructor arguments
 * of LaplCube or LaplRect) and its device arrays. */
#define FDMGPU_KIND_CYL 4   /* plans only (not in groups or the server) */
#define FDMGPU_KIND_RECT 2
#define FDMGPU_KIND_CUBE 3
typedef struct fdmgpu_problem {
    int kind;                 /* FDMGPU_KIND_CUBE or FDMGPU_KIND_RECT */
    int precision;            /* FDMGPU_F32, FDMGPU_F64, FDMGPU_F32_REFINE64 */
    int periodic;             /* mask as for fdmgpu_plan_cube / fdmgpu_plan_rect */
    int nx, ny, nz;           /* nz = 1 for a rect */
    double dx, dy, dz;        /* dz, lz unused for a rect */
    double lx, ly, lz;
    void* d_ans;              /* device array of the plan's element type; != d_rhs for refine */
    const void* d_rhs;
} fdmgpu_problem;

/* Per-problem outcome. */
typedef struct fdmgpu_result {
    int state;                /* 0 queued, 1 on the GPU, 2 done */
    int status;               /* FDMGPU_OK, or the error code of the solve */
    int iterations;           /* FP32 inner solves (refine), 1 otherwise */
    double rel_residual;      /* ||f - A u|| / ||f|| (refine), -1 otherwise */
    double queue_ms;          /* submit -> launch */
    double latency_ms;        /* submit -> completion observed */
} fdmgpu_result;

/* A fixed set of `count` problems of any shapes solved together: problems of
 * one solver form one batch (one launch per stage), batches run concurrently
 * on `nstreams` internal streams (0 = 4).  The first fdmgpu_group_solve runs
 * eagerly; with FDMGPU_GROUP_GRAPH later calls replay a captured hipGraph of
 * the whole sequence (one graph launch).  Kernel variants come from the
 * autotuner cache when FDMGPU_TUNE_CACHE names one (see fdmgpu_tune). */
typedef struct fdmgpu_group_t* fdmgpu_group;
#define FDMGPU_GROUP_GRAPH 1
int fdmgpu_group_create(fdmgpu_group* group, int count, const fdmgpu_problem* problems, int nstreams, int flags);
int fdmgpu_group_solve(fdmgpu_group group, void* stream);
/* Refinement tolerance and inner-solve limit of the group's FP32+refine
 * solvers (defaults 1e-10, 4), as fdmgpu_set_refine per solver; synchronises
 * the group and drops a captured graph (the next solve captures again).
 * No reference counterpart. */
int fdmgpu_group_set_refine(fdmgpu_group group, double rel_tol, int max_iter);
/* After the solve completed (synchronises): result of problem i. */
int fdmgpu_group_result(fdmgpu_group group, int i, fdmgpu_result* result);
int fdmgpu_group_destroy(fdmgpu_group group);

/* Serving: requests are admitted in arrival order while their scratch fits
 * the HBM budget (an arena of `hbm_budget` bytes allocated once), batched per
 * solver as they are admitted (continuous batching: requests arriving while
 * others run form new batches on other streams) and completed
 * asynchronously.  Each stream holds at most two batches in flight: under
 * load, arrivals wait on the host and join larger batches.  Problems of at
 * least 2^22 points run on the first stream, smaller ones on the others
 * (nstreams > 1); the first stream then runs on all but 32 compute units
 * (a CU mask; it synchronizes with the legacy null stream), which stay free
 * for the small requests' kernels.  A request that must wait (no stream of its size class
 * free, or its scratch does not fit yet) can be passed by later requests of
 * other solvers; while it waits for scratch, at most 8 batches pass it before
 * admission waits for it.  Solver tables stay resident between requests.  A
 * request whose scratch alone exceeds the budget fails with
 * FDMGPU_ERR_ADMISSION.  Results of a request do not depend on what else is
 * in flight. */
typedef struct fdmgpu_server_t* fdmgpu_server;
int fdmgpu_server_create(fdmgpu_server* server, unsigned long long hbm_budget, int max_batch, int nstreams);
int fdmgpu_server_submit(fdmgpu_server server, const fdmgpu_problem* problem, unsigned long long* ticket);
/* Refinement tolerance and inner-solve limit of the server's FP32+refine
 * solvers (defaults 1e-10, 4); applies to requests launched afterwards. */
int fdmgpu_server_set_refine(fdmgpu_server server, double rel_tol, int max_iter);
int fdmgpu_server_poll(fdmgpu_server server, unsigned long long ticket, fdmgpu_result* result);
int fdmgpu_server_wait(fdmgpu_server server, unsigned long long ticket, fdmgpu_result* result);
/* Blocks until every submitted request completed. */
int fdmgpu_server_drain(fdmgpu_server server);
/* Forgets completed requests' results (they are kept until then). */
int fdmgpu_server_forget(fdmgpu_server server, unsigned long long ticket);
int fdmgpu_server_destroy(fdmgpu_server server);

/* Autotuner: times the kernel variants of every solver in `problems` (each
 * solved `reps` times per variant on device copies of its right-hand side,
 * into scratch answers: the problems' own arrays are only read, and a
 * problem may alias d_ans and d_rhs) and records the fastest
 * in the process-wide table, and in the JSON file `cache_path` when not NULL
 * (read back by later processes through FDMGPU_TUNE_CACHE). */
int fdmgpu_tune(int count, const fdmgpu_problem* problems, int reps, const char* cache_path);
/* Pins kernel variant *v for problems of the shape of *problem in the
 * process-wide table (as a tuner run that found it would): plans and server
 * batches created afterwards for that shape use it.  FDMGPU_ERR_INVALID_ARG
 * if the variant is not valid for the shape (fdmgpu_plan_variants).  No
 * reference counterpart. */
int fdmgpu_tune_set(const fdmgpu_problem* problem, const fdmgpu_variant* v);

/* ---- NSCube: the caller of the solve --------------------------------------
 * Replaces NSCube<double,check>::NSCube(const Config&) (src/ns_cube.h:46-78)
 * with the configuration values as arguments: domain [x1,x2]x[y1,y2]x[z1,z2],
 * lid velocity u0, Re, dt, sizes; the pressure Poisson solve (the member
 * LaplCube lapl_solver, src/ns_cube.h:77) runs at solver_precision
 * (FDMGPU_F64 or FDMGPU_F32_REFINE64).  Fields start at rest.  nz > nx is
 * FDMGPU_ERR_UNSUPPORTED (the reference's lid loop would index u out of
 * range, src/ns_cube.cpp:68). */
typedef struct fdmgpu_ns_t* fdmgpu_ns;
int fdmgpu_ns_create(fdmgpu_ns* ns, int solver_precision, int nx, int ny, int nz, double x1, double x2, double y1,
                     double y2, double z1, double z2, double u0, double Re, double dt);
/* Replaces NSCube::step() (src/ns_cube.cpp:27-62), `count` times, ordered on
 * `stream` (returns after enqueueing); steps from the third on are replayed
 * from a hipGraph of one step. */
int fdmgpu_ns_step(fdmgpu_ns ns, int count, void* stream);
/* Fields with the reference's index ranges (src/ns_cube.h:66-75), x fastest:
 * U [0,nz+1]x[0,ny+1]x[-1,nx+1], V [0,nz+1]x[-1,ny+1]x[0,nx+1],
 * W [-1,nz+1]x[0,ny+1]x[0,nx+1], P [0,n+1]^3, X and RHS [1,n]^3,
 * F [1,nz]x[1,ny]x[0,nx], G [1,nz]x[0,ny]x[1,nx], H [0,nz]x[1,ny]x[1,nx]. */
#define FDMGPU_NS_U 0
#define FDMGPU_NS_V 1
#define FDMGPU_NS_W 2
#define FDMGPU_NS_P 3
#define FDMGPU_NS_X 4
#define FDMGPU_NS_F 5
#define FDMGPU_NS_G 6
#define FDMGPU_NS_H 7
#define FDMGPU_NS_RHS 8
/* Copy a field to / from host memory (synchronises); count = its element count.
 * With the FP32+refine pressure solve, fdmgpu_ns_get copies the field and
 * then returns FDMGPU_ERR_NOT_CONVERGED if any projection so far ended above
 * the solver's tolerance (a flag kept on the device by every step, replayed
 * ones included; it stays set for the rest of the run).  Warm-started
 * projections run inner solves 2 .. max(2, max_iter - 1) of the solver plan's
 * refinement settings as they were when the step was first enqueued. */
int fdmgpu_ns_get(fdmgpu_ns ns, int which, double* host, unsigned long long count);
int fdmgpu_ns_set(fdmgpu_ns ns, int which, const double* host, unsigned long long count);
/* Element count and device address of a field (device-side post-processing). */
int fdmgpu_ns_info(fdmgpu_ns ns, int which, unsigned long long* count, void** device_ptr);
/* One step, enqueued eagerly (the kernels and arguments of a replayed step,
 * so the fields are bitwise those of fdmgpu_ns_step), with GPU events at the
 * bounds of the reference's four phases; waits for it and writes their
 * durations in ms to ms[0..3]: init_bound (lid, ghost layers, pressure
 * boundary), FGH, poisson (right-hand side and the solve), update_uvwp.
 * Replaces the wall-clock timers of NSCube::step() (src/ns_cube.cpp:27-61)
 * that feed the "Step times (ms)" line of ~NSCube (src/ns_cube.cpp:15-24). */
int fdmgpu_ns_step_timed(fdmgpu_ns ns, double* ms, void* stream);
/* The largest |value| of each field array, ghost layers included, in
 * FDMGPU_NS_* order (out[0..8]); NaNs are skipped, as tensor::maxabs
 * (src/tensor.h:269-273) skips them.  Reduced on the device, synchronises.
 * Feeds the verbose per-step line of NSCube::step() (src/ns_cube.cpp:52-56). */
int fdmgpu_ns_maxabs(fdmgpu_ns ns, double* out);
/* The pressure solver plan (statistics, timing, refinement settings). */
int fdmgpu_ns_solver(fdmgpu_ns ns, fdmgpu_plan* plan);
int fdmgpu_ns_destroy(fdmgpu_ns ns);

/* Replaces ~LaplCube / ~LaplRect. */
int fdmgpu_destroy(fdmgpu_plan plan);

const char* fdmgpu_strerror(int code);

/* ABI version (major*10000 + minor*100 + patch). */
int fdmgpu_version(void);

#ifdef __cplusplus
}
#endif

#endif /* FDMGPU_H */
