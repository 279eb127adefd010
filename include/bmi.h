/*
 * bmi.h -- C ABI of the MI355X BEM matrix-assembly library (libbmi.so).
 *
 * Drop-in target: BEAST's assemble(op, test_space, trial_space) as replaced by
 * the reference's assemble_gpu (assamble_gpu.jl:25-137).  A Julia caller
 * reaches these entry points through ccall; tests and benchmarks use ctypes.
 *
 * Conventions
 *   - Arrays are plain host pointers with explicit sizes; no language types
 *     cross the boundary.  Indices are 0-based (a Julia shim subtracts 1).
 *   - Matrices are column-major (row index fastest), as Julia's Matrix and the
 *     reference's store index mapping (store_with_kernel.jl:94).
 *   - Complex values are interleaved (re, im) pairs.
 *   - The library owns every device object; the caller owns host buffers.
 *   - No exception or longjmp crosses the ABI; every call returns a status and
 *     bmi_last_error() gives a thread-local message for the last failure.
 *   - A context is externally synchronised: one thread at a time per context.
 *   - Objects may be freed in any order: a free touches only its own object
 *     (a mesh, space, plan, matrix or batch freed after its context, or a
 *     space after its mesh, releases its device memory all the same).  An
 *     object whose parents are gone accepts no other call.
 */
#ifndef BMI_H
#define BMI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t bmi_status;
#define BMI_OK 0
#define BMI_EINVAL (-1)
#define BMI_ENOMEM (-2)
#define BMI_EDEVICE (-3)
#define BMI_EBUSY (-4)
#define BMI_ECANCELED (-5)
#define BMI_EUNSUPPORTED (-6)
#define BMI_EINTERNAL (-7)  /* unexpected failure inside the library (bmi_last_error says which) */

/* function spaces (BEAST lagrangec0d1 / raviartthomas) */
#define BMI_SPACE_LAGRANGE_C0D1 1
#define BMI_SPACE_RWG 2

/* operators: Helmholtz3D.singlelayer, Maxwell3D.singlelayer, and
 * Maxwell3D.doublelayer tested with rotated RWG functions n x f_m (the n x K
 * block of the MFIE, BEAST's assemble(K, n x X, X)):
 *   Z_mn = alpha < n x f_m, grad_x G x f_n >  (beta unused) */
#define BMI_OP_HH3D_SINGLELAYER 1
#define BMI_OP_MW_SINGLELAYER 2
#define BMI_OP_MW_DOUBLELAYER 3

/* arithmetic of the far-pair double quadrature */
#define BMI_PREC_FP64 0
#define BMI_PREC_MIXED 1

/* matrix element types */
#define BMI_C64 0 /* ComplexF64 */
#define BMI_C32 1 /* ComplexF32 */

/* pair classes, in the order of quadrule_determine.jl:76 (0 far, 1 common
 * vertex, 2 common edge, 3 common face, 4 near); 5 = far pair kept in FP64
 * by the mixed-precision path */
#define BMI_CLASS_FAR 0
#define BMI_CLASS_CV 1
#define BMI_CLASS_CE 2
#define BMI_CLASS_CF 3
#define BMI_CLASS_NEAR 4
#define BMI_CLASS_FAR64 5

typedef struct bmi_context bmi_context;
typedef struct bmi_mesh bmi_mesh;
typedef struct bmi_space bmi_space;
typedef struct bmi_plan bmi_plan;
typedef struct bmi_matrix bmi_matrix;

/* Operator coefficients: Z_mn = alpha <f_m, G f_n> + beta <div f_m, G div f_n>,
 * G = exp(-gamma R)/(4 pi R).  Fields mirror BEAST's operator structs
 * (alpha, gamma; HH3DSingleLayerFDBIO at assamble_chunk_gpu.jl:44-47; beta
 * for MWSingleLayer3D).  beta is ignored by the Helmholtz operator. */
typedef struct {
    int32_t kind;
    double alpha[2];
    double beta[2];
    double gamma[2];
} bmi_operator;

/* Quadrature and precision.  Replaces the quadstrat / qd arguments of
 * assemble_gpu (assamble_chunk_gpu.jl:67-75) and GPUConfiguration fields
 * (utils/configuration.jl:11-26). */
typedef struct {
    int32_t far_test_order;  /* triangle rule for test cells: 1, 2 (3 points) or 3 (4 points) */
    int32_t far_trial_order; /* triangle rule for trial cells */
    int32_t ss_order[3];     /* Gauss-Legendre points per hypercube dimension for CV, CE, CF (1..8) */
    int32_t near_order;      /* collapsed Gauss order for class-4 pairs (1..8) */
    int32_t precision;       /* BMI_PREC_FP64 | BMI_PREC_MIXED */
    double near_factor;      /* mixed: far pairs with centroid distance < near_factor * max edge stay FP64 */
} bmi_options;

/* Caller-supplied quadrature tables, in place of the orders of bmi_options:
 * the far-pair triangle rules of test and trial cells and the Gauss-Legendre
 * rules of the Sauter-Schwab hypercube per touching class.  A point (u, v)
 * of a cell is v3 + u (v1 - v3) + v (v2 - v3), the chart of the reference
 * (load_parameters.jl:25-46); weights are on the reference triangle (they sum
 * to 1/2 for an exact rule; the library applies the Jacobian 2A).  BEAST's
 * quaddata holds the same rules as per-cell Cartesian points and weights
 * times the Jacobian (qd.tpoints / qd.bpoints, read by GPU_scheduler.jl:61-106);
 * its Gauss-Legendre tables qd.gausslegendre feed the Sauter-Schwab rules
 * (nonMainCaseQuadratures.jl:95-97, SauterSchwab.jl:407-414).  Rules equal to
 * the defaults (3 x 4 points, 5-point Gauss-Legendre) run the same kernels as
 * the defaults. */
#define BMI_MAX_RULE_POINTS 4
#define BMI_MAX_GL_POINTS 8
typedef struct {
    int32_t test_points;                          /* 1..BMI_MAX_RULE_POINTS */
    int32_t trial_points;
    double test_uv[BMI_MAX_RULE_POINTS][2];
    double test_weight[BMI_MAX_RULE_POINTS];
    double trial_uv[BMI_MAX_RULE_POINTS][2];
    double trial_weight[BMI_MAX_RULE_POINTS];
    int32_t gl_points[3];                         /* common vertex, edge, face: 1..BMI_MAX_GL_POINTS */
    double gl_node[3][BMI_MAX_GL_POINTS];         /* on [0, 1] */
    double gl_weight[3][BMI_MAX_GL_POINTS];       /* sum to 1 */
} bmi_rules;

/* Per-assembly phase timings (milliseconds, HIP events on the library stream)
 * and pair counts of the last bmi_assemble on a plan. */
typedef struct {
    double classify_ms;
    double special_ms;  /* Sauter-Schwab, near and FP64-kept far pairs */
    double far_ms;      /* far-pair block kernels incl. write-back */
    double total_ms;
    int64_t pairs[6];   /* per BMI_CLASS_* */
    int32_t far_launches;
    int32_t n_colours_test;
    int32_t n_colours_trial;
    int64_t plan_bytes;  /* device memory held by the plan (its spaces' and meshes' not included) */
} bmi_stats;

/* Library version string. */
const char *bmi_version(void);

/* Message of the last failed call on this thread. */
const char *bmi_last_error(void);

/* Fills opts with the defaults: 3 x 4 point far rule, 5-point Gauss-Legendre
 * Sauter-Schwab, near_order 4, FP64, near_factor 2. */
void bmi_options_default(bmi_options *opts);

/* One context per device; owns a HIP stream.
 * Replaces GPUConfiguration.backend (utils/configuration.jl:20). */
bmi_status bmi_context_create(int32_t device, bmi_context **out);
bmi_status bmi_context_destroy(bmi_context *ctx);

/* Triangle mesh: xyz is 3 x nv column-major, tri is 3 x nt (0-based).
 * Replaces the element packing of load_data_elements (load_parameters.jl:25-46). */
bmi_status bmi_mesh_create(bmi_context *ctx, const double *xyz, int64_t nv, const int32_t *tri, int64_t nt,
                           bmi_mesh **out);
bmi_status bmi_mesh_free(bmi_mesh *mesh);

/* Function space: gidx / coeff are 3 x nt (local shape i of cell K -> global
 * function and coefficient; gidx < 0 means no function), the first-slot layout
 * of extract_values_and_indexes (load_parameters.jl:50-55). */
bmi_status bmi_space_create(bmi_mesh *mesh, int32_t kind, const int64_t *gidx, const double *coeff, int64_t nfun,
                            bmi_space **out);
/* bmi_space_create with the far-pair tile shape given: test panels of at most
 * panel_rows functions (1 .. 64 on RWG, 1 .. 128 on Lagrange spaces) and
 * trial patches of at most patch_cols functions (3 .. 40 on RWG, 3 .. 20 on
 * Lagrange spaces); 0 takes the kind's default, the largest of each, which
 * is the fastest shape on every workload measured (bench.py --tile-sweep).
 * Any shape assembles the same matrix up to summation order.
 * No counterpart in the reference; its analogue is the thread and chunk
 * counts that testing/determin_amount_of_threads.jl:44-75 and
 * testing/Amdahl's_law_3.jl:142-154 sweep by hand. */
bmi_status bmi_space_create_shaped(bmi_mesh *mesh, int32_t kind, const int64_t *gidx, const double *coeff,
                                   int64_t nfun, int32_t panel_rows, int32_t patch_cols, bmi_space **out);
bmi_status bmi_space_free(bmi_space *space);

/* Assembly plan for (operator, test space, trial space, options): device
 * geometry, patch colouring and special-pair buffers.  Replaces the set-up half
 * of assemblechunk_gpu! / assemblechunk_body_gpu! (assamble_chunk_gpu.jl:49-144). */
bmi_status bmi_plan_create(bmi_context *ctx, const bmi_operator *op, const bmi_space *test, const bmi_space *trial,
                           const bmi_options *opts, bmi_plan **out);
/* The same with the caller's quadrature tables (rules); opts still gives the
 * precision, near_order and near_factor.  Replaces the quaddata argument of
 * the reference's chunk assembly (assamble_chunk_gpu.jl:75) as consumed by
 * gether_wimps_and_womps / load_data (GPU_scheduler.jl:61-106).
 * Test and trial spaces on different meshes: BMI_EUNSUPPORTED if a vertex of
 * one mesh lies inside a cell of the other without being one of its vertices
 * (one mesh refines or overlaps the other; the reference switches to
 * BEAST's TestRefinesTrialQStrat / TrialRefinesTestQStrat there,
 * assamble_chunk_gpu.jl:67-73).  bmi_plan_create applies the same check. */
bmi_status bmi_plan_create_rules(bmi_context *ctx, const bmi_operator *op, const bmi_space *test,
                                 const bmi_space *trial, const bmi_options *opts, const bmi_rules *rules,
                                 bmi_plan **out);
bmi_status bmi_plan_free(bmi_plan *plan);

/* Device-resident matrix, column-major rows x cols.
 * Replaces the device result cache (CustomDataStructs/GpuWriteBack.jl:11-58). */
bmi_status bmi_matrix_create(bmi_context *ctx, int64_t rows, int64_t cols, int32_t dtype, bmi_matrix **out);
bmi_status bmi_matrix_free(bmi_matrix *m);
bmi_status bmi_matrix_device_ptr(const bmi_matrix *m, void **ptr, int64_t *ld);

/* Assemble the whole matrix into out (blocking).  Classification, singular
 * integrals, far pairs and write-back all run on the device.
 * Replaces assemble_gpu's compute path (assamble_gpu.jl:25-87). */
bmi_status bmi_assemble(bmi_plan *plan, bmi_matrix *out);

/* Timings of the last bmi_assemble on this plan (zero before the first) and
 * the pair counts by class (the first request before any assembly runs the
 * classification once to obtain them). */
bmi_status bmi_plan_stats(const bmi_plan *plan, bmi_stats *stats);

/* Copy a device matrix to caller-owned host memory as C64 or C32.
 * Replaces copy_to_CPU (utils/copy_to_CPU.jl:5-61). */
bmi_status bmi_matrix_to_host(const bmi_matrix *m, void *dst, int32_t dtype);

/* Assemble the whole matrix straight into caller-owned host memory dst
 * (column-major rows x cols, dtype BMI_C64 or BMI_C32, pageable or pinned),
 * for matrices that need not -- or do not -- fit in device memory: row blocks
 * of whole test panels are assembled into two device staging blocks that
 * together hold at most device_budget bytes, and each block is copied out
 * while the next is being assembled (at least 8 blocks when the test space
 * has 8 panels).  The values are bitwise those of bmi_assemble into a
 * matrix of element type dtype followed by bmi_matrix_to_host.  Blocking;
 * BMI_ENOMEM if the budget holds less than two blocks of one test panel.
 * Replaces the storeCPU path of assemble_gpu (assamble_gpu.jl:8-23 with
 * allocatestorage :195-226): the memory-budgeted tiles and producer tasks of
 * schedule_kernel! / producer (GPU_scheduler.jl:134-288) and the host scatter
 * write_to_compact_matrix (CustomDataStructs/GpuWriteBack.jl:60-132). */
bmi_status bmi_assemble_to_host(bmi_plan *plan, void *dst, int32_t dtype, int64_t device_budget);

/* y = Z x for nrhs (1..8) right-hand sides on the device-resident matrix: x is
 * cols x nrhs, y is rows x nrhs, both caller-owned host ComplexF64 arrays in
 * column-major order; products and sums in FP64, deterministic order.
 * Replaces the host-side application of the returned Matrix{ComplexF64}
 * (the solver's Z*x after assemble_gpu, assamble_gpu.jl:135), without the
 * 15 GB device-to-host copy. */
bmi_status bmi_matrix_apply(const bmi_matrix *m, const double *x, double *y, int32_t nrhs);

/* bmi_matrix_apply without any device allocation, for callers that cannot
 * allocate (code under graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The device copies of
 * x and y and the apply kernels' partial sums live in a 128 MiB device array
 * that each GPU the library runs on holds until the process ends; sizes that
 * need more ((cols + rows + chunks x rows) x nrhs x 16 bytes, chunks <= 17 at
 * 30,720 rows) fail with BMI_EINVAL.  Every call on one GPU shares that
 * array: calls that overlap on one GPU race.  Values are bitwise those of
 * bmi_matrix_apply. */
bmi_status bmi_matrix_apply_noalloc(const bmi_matrix *m, const double *x, double *y, int32_t nrhs);

/* out[0] = ||a - ref||_F, out[1] = ||ref||_F, out[2] = max |a_ij - ref_ij|, computed
 * on the device (a and ref of equal shape, any element types).
 * Replaces test(M, M_ref) of test/difference.jl:1-6, which compared one column. */
bmi_status bmi_matrix_compare(const bmi_matrix *a, const bmi_matrix *ref, double out[3]);

/* --------------------------------------------------------------- batching
 * Many assemblies (plans with their output matrices, one context) run as one
 * replayed hipGraph, one branch per operator: one classification launch and
 * one special-pair launch over its problems, then, per bucket of problems that
 * share a far-pair kernel (operator, precision, rule, output type), one of
 * three schedules: one launch per trial colour over every problem's (test
 * panel, trial patch) blocks; one persistent launch whose work-groups pop
 * (problem, test panel) items, largest first, from a device queue and
 * integrate each panel against all trial patches of its problem in colour
 * order (no work-group waits for another); or each problem's own per-colour
 * launches in turn -- whichever the autotuner at creation measured fastest
 * (the persistent launch with 1, 2 or 3 work-groups per CU; results are
 * cached per bucket signature for the process).  Either way the matrices are
 * bitwise those of bmi_assemble on each plan.
 * No counterpart in the reference (its batching precedent is the producer /
 * Channel tiling of GPU_scheduler.jl:160-206 and the hand sweeps of
 * testing/determin_amount_of_threads.jl:44-75). */
typedef struct bmi_batch bmi_batch;

#define BMI_BATCH_MAX_BUCKETS 8
#define BMI_BATCH_CANDIDATES 5

typedef struct {
    double last_ms;               /* device time of the last bmi_batch_assemble (events around the graph) */
    int32_t problems;
    int32_t buckets;
    int32_t singles;              /* problems assembled with their own per-colour launches (no bucket) */
    int32_t bucket_problems[BMI_BATCH_MAX_BUCKETS];
    int32_t bucket_mode[BMI_BATCH_MAX_BUCKETS];        /* 0: per colour, 1: persistent queue, 2: per problem */
    int32_t bucket_wgs_per_cu[BMI_BATCH_MAX_BUCKETS];  /* persistent queue: work-groups per CU */
    double bucket_tuned_ms[BMI_BATCH_MAX_BUCKETS];     /* far phase of the chosen schedule when tuned */
    /* far phase of each schedule the tuner timed for the bucket's signature:
     * [0] per colour, [c] persistent with c = 1 .. 3 work-groups per CU,
     * [4] per problem; 0 if not timed */
    double bucket_candidate_ms[BMI_BATCH_MAX_BUCKETS][BMI_BATCH_CANDIDATES];
} bmi_batch_stats;

/* plans[i] assembles into matrices[i]; every object on ctx.  Runs the
 * prologues once, tunes, captures and instantiates the graph. */
bmi_status bmi_batch_create(bmi_context *ctx, bmi_plan *const *plans, bmi_matrix *const *matrices, int32_t n,
                            bmi_batch **out);
/* Replays the graph; blocking. */
bmi_status bmi_batch_assemble(bmi_batch *batch);
bmi_status bmi_batch_stats_get(const bmi_batch *batch, bmi_batch_stats *stats);
/* Overrides the tuned schedule of one bucket (mode 0: per colour; mode 1:
 * persistent with wgs_per_cu work-groups per CU, 1 .. the kernel's occupancy;
 * mode 2: per problem) and captures the graph again.  BMI_EINVAL for a bucket or schedule that
 * does not exist. */
bmi_status bmi_batch_set_schedule(bmi_batch *batch, int32_t bucket, int32_t mode, int32_t wgs_per_cu);
bmi_status bmi_batch_free(bmi_batch *batch);

/* Free and total device memory of the context's device (bytes).
 * No counterpart in the reference; callers that size work by what fits in
 * HBM (admission control, the full-size tests) use it. */
bmi_status bmi_device_memory(const bmi_context *ctx, int64_t *free_bytes, int64_t *total_bytes);

/* ---------------------------------------------------------------- serving
 * A request loop over one device: assembly, assemble-and-apply and apply
 * requests arrive at any time from any thread; up to `slots` of them run at
 * once, each on its own stream, so the device work of different requests
 * overlaps; a request is admitted when a slot is free and the device memory
 * it needs (its N x N matrix plus plan buffers) fits beside the matrices
 * already resident.  Assembled matrices of assembly requests stay resident
 * in HBM until released; apply requests use them.  Results do not depend on
 * what else runs: every kernel sums in a fixed order.
 * No counterpart in the reference (its only scheduling precedent is the
 * producer / Channel tiling of GPU_scheduler.jl:160-206). */
typedef struct bmi_server bmi_server;

#define BMI_REQ_ASSEMBLE 0       /* assemble and keep the matrix resident */
#define BMI_REQ_ASSEMBLE_APPLY 1 /* assemble, y = Z x, free the matrix */
#define BMI_REQ_APPLY 2          /* y = Z x on a resident matrix */

typedef struct {
    int32_t slots;          /* requests on the device at once, 1..16 */
    int64_t memory_budget;  /* bytes requests may hold together; 0 = free memory at creation less 4 GiB */
} bmi_server_options;

typedef struct {
    int32_t kind;             /* BMI_REQ_* */
    /* geometry and space (assembly kinds); host arrays are copied at submit */
    const double *xyz;        /* 3 x nv */
    int64_t nv;
    const int32_t *tri;       /* 3 x nt, 0-based */
    int64_t nt;
    int32_t space_kind;       /* BMI_SPACE_*; test space = trial space */
    const int64_t *gidx;      /* 3 x nt */
    const double *coeff;      /* 3 x nt */
    int64_t nfun;
    bmi_operator op;
    bmi_options opts;
    int32_t dtype;            /* BMI_C64 | BMI_C32 */
    uint64_t matrix;          /* BMI_REQ_APPLY: id of the assembly request that holds the matrix */
    int32_t nrhs;             /* apply kinds: 1..8 */
    const double *x;          /* apply kinds: nfun x nrhs complex, copied at submit */
} bmi_request;

typedef struct {
    int32_t status;       /* BMI_OK or the failure's status */
    char message[160];    /* failure message */
    double queue_ms;      /* submit -> admitted */
    double prep_ms;       /* host-side preparation: mesh, space, plan */
    double device_ms;     /* assembly and apply */
    double total_ms;      /* submit -> done */
    int64_t entries;      /* matrix entries assembled */
} bmi_result;

bmi_status bmi_server_create(int32_t device, const bmi_server_options *opts, bmi_server **out);
/* Waits for every submitted request, frees resident matrices. */
bmi_status bmi_server_destroy(bmi_server *server);
/* Queue a request; y (nfun x nrhs complex, apply kinds) is written when the
 * request completes and must stay valid until then.  Returns at once. */
bmi_status bmi_server_submit(bmi_server *server, const bmi_request *req, double *y, uint64_t *id);
/* Wait up to timeout_ms (< 0: no limit) for request id; BMI_EBUSY if it has
 * not completed yet.  The result is returned once; then the id is forgotten
 * (an assembly request's id still names its resident matrix). */
bmi_status bmi_server_wait(bmi_server *server, uint64_t id, int32_t timeout_ms, bmi_result *res);
/* Free the resident matrix of assembly request id. */
bmi_status bmi_server_release(bmi_server *server, uint64_t id);
/* Device memory held by requests now and at most so far (bytes). */
bmi_status bmi_server_memory(const bmi_server *server, int64_t *resident, int64_t *peak);
/* Device memory the server admits an assembly request against: its matrix
 * and a bound on its plan's buffers (geometry and slot records, block
 * tables, special-pair entries of touching, near and -- mixed path -- FP64
 * far pairs), computed from the request's mesh.  Apply requests: 0. */
bmi_status bmi_request_bytes(const bmi_request *req, int64_t *matrix_bytes, int64_t *plan_bytes);

#ifdef __cplusplus
}
#endif

#endif /* BMI_H */
