// Internal types shared by the host orchestration (bmi_host.cpp) and the
// gfx950 kernels (bmi_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Cells per trial patch and per test sub-batch: a work-group of 256 threads
// integrates 16 x 16 cell pairs at a time.
#define BMI_PATCH 16
#define BMI_BLOCK_THREADS 256
// at most 3 local shapes per cell -> at most 48 slots per group of cells
#define BMI_MAX_ROWS (3 * BMI_PATCH)
// trial patches are cut to carry at most this many functions (columns)
#define BMI_MAX_COLS 40
// output rows owned by one test panel (consecutive functions); a column of
// a panel is one contiguous 1 KB (ComplexF64) segment of the matrix
#define BMI_PANEL_ROWS 64
// Lagrange spaces: test panels of 128 rows (a vertex function has about 6
// cells, so 64 consecutive vertices are a thin strip whose border cells repeat
// in the neighbouring panels), trial patches of at most 20 columns; the tile
// has the same size
#define BMI_PANEL_ROWS_WIDE 128
#define BMI_MAX_COLS_WIDE 20
static_assert(BMI_PANEL_ROWS_WIDE * BMI_MAX_COLS_WIDE <= BMI_PANEL_ROWS * BMI_MAX_COLS, "wide panels need a larger tile");
// accumulators per thread: 64 rows x 48 columns / 256 threads
#define BMI_PANEL_ACC (BMI_PANEL_ROWS * BMI_MAX_ROWS / BMI_BLOCK_THREADS)
#define BMI_MAX_RULE 4      // far-rule points per cell
#define BMI_MAX_GL 8        // Gauss-Legendre points per hypercube dimension
#define BMI_MAX_NEAR 64     // points of the collapsed near rule per cell
#define BMI_MAX_RHS 8       // right-hand sides per bmi_matrix_apply call
#define BMI_APPLY_CHUNK 512 // columns of x staged in LDS at a time by the apply kernel

enum { OPK_HH3D = 1, OPK_MW = 2, OPK_DL = 3 };
enum { CLS_FAR = 0, CLS_CV = 1, CLS_CE = 2, CLS_CF = 3, CLS_NEAR = 4, CLS_FAR64 = 5 };
// flag on a special-pair entry's class: the pair is integrated in the entry
// of the test cell's home group and copied here
#define SPEC_NOT_OWNER 0x80

// Device view of one function space on one mesh.
struct DevSpace {
    int nt;
    int npatch;
    const double *vtx;   // nt x 9: v1, v2, v3 (xyz each)
    const double *cen;   // nt x 3 centroid
    const double *area;  // nt
    const double *pts;   // nt x BMI_MAX_RULE x 3: far-rule points relative to the centroid
    const int *patch_tri;        // npatch x BMI_PATCH, -1 padded
    const double *patch_sph;     // npatch x 4: bounding sphere (centre, radius) of the cells' vertices
    const double *patch_amax;    // npatch: largest cell area
    const int *home;             // nt: the group a cell calls home (its first group; -1 if in none)
};

struct OpConst {
    double alpha_re, alpha_im;
    double beta_re, beta_im;
    double gamma_re;   // attenuation; 0 for a real wavenumber
    double k2pi;       // Im(gamma) / (2 pi), the wavenumber in turns per unit length
    int lossy;         // gamma_re != 0
    // far kernels with the compile-time rule accumulate entries divided by
    // alpha (alpha != 0) and multiply by alpha when they write the matrix
    double bp_re, bp_im;  // 4 beta / alpha
    double ia_re, ia_im;  // 1 / alpha
};

struct RuleConst {
    int nt, nb;                       // far-rule points per test / trial cell
    double lt[BMI_MAX_RULE][3];       // barycentric coordinates of test points
    double lb[BMI_MAX_RULE][3];
    double w4pi[BMI_MAX_RULE][BMI_MAX_RULE];  // w_I * w_J / (4 pi), reference weights
};

// Packed geometry of a group of BMI_PATCH cells (a test sub-batch or a trial
// patch) for one far rule, BMI_GEO doubles per group, loaded with one
// coalesced read:
#define BMI_GEO_PTS 0                                  // 16 x 12 rule points, relative to the centroid
#define BMI_GEO_CEN (BMI_PATCH * 3 * BMI_MAX_RULE)     // 16 x 3 centroids
#define BMI_GEO_AREA (BMI_GEO_CEN + 3 * BMI_PATCH)     // 16 areas (0 = no cell)
#define BMI_GEO_VREL (BMI_GEO_AREA + BMI_PATCH)        // 16 x 9 vertices, relative to the centroid
#define BMI_GEO (BMI_GEO_VREL + 9 * BMI_PATCH)         // 400
// FP32 copy of the points and vertices in the front of a record's LDS (the
// mixed far kernels): float index of the vertices; it ends before the
// centroids, which stay FP64 in place
#define BMI_GEOF_VREL (BMI_PATCH * 3 * BMI_MAX_RULE)
static_assert((BMI_GEOF_VREL + 9 * BMI_PATCH) * sizeof(float) <= BMI_GEO_CEN * sizeof(double),
              "FP32 geometry overlaps the centroids");

// Far-pair assembly.  Test side: panels of at most BMI_PANEL_ROWS consecutive
// functions (BMI_PANEL_ROWS_WIDE for Lagrange spaces), each with the cells
// that carry them (cells of several panels are repeated), cut into
// sub-batches of BMI_PATCH cells; per sub-batch a slot
// table: for each of the 64 rows the slots (local cell * 3 + shape,
// coefficient) of that row's function.  Trial side: Morton patches of
// BMI_PATCH cells with their columns (functions) and slot tables.  All
// per-group records have fixed sizes, so a work-group needs no dependent
// loads to find them.
struct PanelArgs {
    int npanel;                  // panels of this launch: panel0 .. panel0 + npanel - 1
    int panel0;
    int nrows;                   // functions of the test space
    const int *prow;             // npanel + 1: first row of each panel (at most its height apart)
    const int *sb_off;           // npanel + 1: sub-batches of each panel
    const double *tgeo;          // nsb x BMI_GEO
    const int *trs;              // nsb x (BMI_PANEL_ROWS + 1): slot range of each row (0..48); 64-row panels only
    const unsigned char *tkl;    // nsb x 48
    const double *tcf;           // nsb x 48
    const double *bgeo;          // npb x BMI_GEO
    const int *bnc;              // npb: columns of each trial patch
    const int *bfn;              // npb x 48: function of each column
    const unsigned char *bfirst; // npb x 48: 1 if the column is stored (first trial patch in launch order)
    const int *bcs;              // npb x 49: slot range of each column (0..48)
    const unsigned char *bkl;    // npb x 48
    const double *bcf;           // npb x 48
    // rank-round records (one per cell slot, 48 per group): bits 0-7 row (test)
    // or column (trial) index inside the work-group's tile, 0xff = none;
    // bits 8-15 rank of the slot among the group's slots of that row/column;
    // trial side, bits 16-23: level of the slot, its rank among the patch's
    // slots of that column with the same local shape index
    const int *tslot;            // nsb x 48
    const double *tscf;          // nsb x 48
    const int *bslot;            // npb x 48
    const double *bscf;          // npb x 48
    int rounds_t, rounds_b;      // largest test rank + 1; largest trial rank (level) + 1
    const unsigned char *brounds;  // npb: trial rounds of each patch (by rank, or by level in the Helmholtz kernels)
    int ncm;                     // largest number of columns of a trial patch
    int rr;                      // 1: rank-round kernel, 0: pair buffer + gather
    int fast;                    // 1: default 3 x 4 point far rule and a lossless operator (compile-time rule)
    const int *bq_list;          // trial patches of this launch's colour
    int nbq;
    const int *bq_all;           // every trial patch, colour by colour
    int npb;                     // trial patches in total (block id = sub-batch * npb + q)
    const int *blk_off;          // special-pair entries of each block: first ...
    const int *blk_cnt;          // ... and count
    const unsigned char *spec_loc;
    const double2 *spec_val;
    int capacity;                // special-pair entries allocated
    const double2 *sc;           // 64 x (sin, cos) of 2 pi j / 64, the FP64 Green evaluations' table
    void *Z;
    long long ld;
    OpConst op;
    RuleConst rule;
};

struct ClassifyArgs {
    DevSpace X, Y;
    int npa, npb;
    double k2x16;        // 16 |gamma|^2
    double near_d2;      // (near_factor * max edge)^2, mixed precision only; < 0 disables
    double near_d;       // near_factor * max edge, mixed precision only; < 0 disables
    int *blk_off;        // output: first entry of each block ...
    int *blk_cnt;        // ... and its number of entries
    int *counter;        // entries reserved so far (zeroed before the launch)
    int *spec_K, *spec_L;
    unsigned char *spec_cls, *spec_loc;
    int capacity;
};

struct SpecialArgs {
    DevSpace X, Y;
    const int *spec_K, *spec_L;
    const unsigned char *spec_cls;
    const unsigned char *spec_loc;  // pair's thread in its block (kl * 16 + ll)
    const int *n_special;
    const int *blk_off, *blk_cnt;
    int npb;
    int capacity;        // special-pair entries allocated
    const double2 *sc;   // 64 x (sin, cos) of 2 pi j / 64 (as PanelArgs::sc)
    double2 *spec_val;
    OpConst op;
    RuleConst rule;
    int ss_n[3];
    double glx[3][BMI_MAX_GL], glw[3][BMI_MAX_GL];
    int near_np;
    double nb1[BMI_MAX_NEAR], nb2[BMI_MAX_NEAR], nw[BMI_MAX_NEAR];
    // sym: a symmetric kernel (single layers) on one space as test and trial,
    // so the pair (L, K) has the transposed block of (K, L): common-vertex
    // pairs with K > L are copied from (L, K) transposed, and the common-face
    // regions that are each other's x <-> y swap are evaluated once
    int sym;
    // cf_zero: common-face blocks are zero (the double layer on a flat cell:
    // n.(x - y) and n.(y - c) vanish) and are not integrated
    int cf_zero;
};

// kernel launchers (bmi_kernels.hip)
// classification, one problem or a batch: work-group g takes sub-batch
// g - gbase[i] of problem i (gbase: nprob + 1 prefix of the sub-batch counts)
hipError_t launch_rule_points(const double *vtx, const double *cen, long long nt, int n, const double l[][3],
                              double *pts, hipStream_t s);
hipError_t launch_geo_records(const double *pts, const double *vtx, const double *cen, const double *area,
                              const int *cells, int ngroups, double *g, hipStream_t s);
// the 64-entry sine/cosine table of the FP64 Green evaluations, written to tab
hipError_t launch_sincos_table(double2 *tab, hipStream_t s);
hipError_t launch_classify(const ClassifyArgs *args, const int *gbase, int nprob, int total_groups, hipStream_t s);
hipError_t launch_zero_each(int *const *ptrs, int n, hipStream_t s);
// special pairs, one problem or a batch of one operator: cbase / ebase are
// prefixes of the problems' chunk counts (chunks of cs entries) and capacities
// ng5: every problem uses 5 Gauss-Legendre points per dimension for CV, CE, CF;
// fast: every problem has the default 3 x 4 far rule and a lossless operator
hipError_t launch_special(const SpecialArgs *args, const int *cbase, const int *ebase, int nprob, int total_chunks,
                          int total_entries, int cs, int mixed, int op_kind, int ng5, int fast, hipStream_t s);
int special_chunk(int capacity);
hipError_t launch_panels(const PanelArgs &a, int op_kind, int mixed, int c32, hipStream_t s);
// batched far-pair launches over items (problem, panel, patch, global panel)
hipError_t launch_panels_items(const PanelArgs *args, const int4 *items, int nitems, int ncm, int op_kind, int mixed,
                               int fast, int c32, hipStream_t s);
// persistent, panel-major: items (problem, test panel) popped through *head
// (zeroed before the launch), each against every trial patch in colour order
hipError_t launch_panels_persist(const PanelArgs *args, const int2 *items, int nitems, int ncm, int *head, int grid,
                                 int op_kind, int mixed, int fast, int c32, hipStream_t s);
hipError_t launch_fill_zero(void *p, size_t bytes, hipStream_t s);
hipError_t launch_zero_ints(int *p, int n, hipStream_t s);
// column chunks of the apply kernels (and the chunk width); the partial
// sums take nchunk x rows x nrhs complex values
int apply_chunks(int rows, int cols, int *chunk);
hipError_t launch_apply(const void *Z, int c32, long long ld, int rows, int cols, const double2 *x, int nrhs,
                        double2 *partial, double2 *y, hipStream_t s);
hipError_t launch_compare(const void *a, int a32, const void *b, int b32, long long n, int groups, double *out,
                          hipStream_t s);

// host helpers shared by the C ABI files (bmi_host.cpp)
#include <string>
struct bmi_matrix;
struct bmi_context;
hipStream_t context_stream(const bmi_context *ctx);
// Allocations and uploads of an entry point go to its context's stream
// while one of these is alive on the thread.
class WorkStream {
public:
    explicit WorkStream(hipStream_t s);
    ~WorkStream();
    WorkStream(const WorkStream &) = delete;
    WorkStream &operator=(const WorkStream &) = delete;
private:
    hipStream_t prev_;
};
hipError_t dev_alloc(void **p, size_t bytes);
void dev_free(void *p);
hipError_t dev_upload(void *dst, const void *src, size_t bytes);
int32_t set_error(int32_t status, const std::string &msg);
// status for an exception caught at the C boundary (call inside a catch block)
int32_t fail_exception() noexcept;
int32_t matrix_apply_on(bmi_context *run, const bmi_matrix *m, const double *x, double *y, int32_t nrhs);
// the fields of a matrix the allocation-free apply needs
struct MatrixView {
    const void *d;
    int64_t rows, cols;
    int32_t dtype;
    int device;
    hipStream_t stream;  // its context's
};
MatrixView matrix_view(const bmi_matrix *m);
// address lookup of the allocation-free apply's scratch on the current
// device dev, once per device (bmi_noalloc.hip)
int32_t scratch_setup(int dev);
int32_t matrix_wrap(bmi_context *ctx, void *mem, int64_t rows, int64_t cols, int32_t dtype, bmi_matrix **out);
