// Internal types shared by the host orchestration (bmi_host.cpp) and the
// gfx950 kernels (bmi_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Triangles per patch.  A work-group of 256 threads owns one
// (test patch) x (trial patch) block of 16 x 16 cell pairs.
#define BMI_PATCH 16
#define BMI_BLOCK_THREADS 256
// at most 3 local shapes per cell -> at most 48 rows (or columns) per patch
#define BMI_MAX_ROWS (3 * BMI_PATCH)
#define BMI_MAX_RULE 4      // far-rule points per cell
#define BMI_MAX_GL 8        // Gauss-Legendre points per hypercube dimension
#define BMI_MAX_NEAR 64     // points of the collapsed near rule per cell

enum { OPK_HH3D = 1, OPK_MW = 2 };
enum { CLS_FAR = 0, CLS_CV = 1, CLS_CE = 2, CLS_CF = 3, CLS_NEAR = 4, CLS_FAR64 = 5 };

// Device view of one function space on one mesh.
struct DevSpace {
    int nt;
    int npatch;
    const double *vtx;   // nt x 9: v1, v2, v3 (xyz each)
    const double *cen;   // nt x 3 centroid
    const double *area;  // nt
    const double *pts;   // nt x BMI_MAX_RULE x 3: far-rule points relative to the centroid
    const int *patch_tri;        // npatch x BMI_PATCH, -1 padded
    const int *row_ptr;          // npatch + 1: rows of each patch
    const int *row_fn;           // global function of each row
    const unsigned char *row_first;  // 1 if this patch is the first (in launch order) to write the row
    const int *row_slot_ptr;     // rows + 1
    const unsigned char *slot_kl;    // local cell * 3 + local shape
    const double *slot_cf;       // coefficient of the slot
};

struct OpConst {
    double alpha_re, alpha_im;
    double beta_re, beta_im;
    double gamma_re;   // attenuation; 0 for a real wavenumber
    double k2pi;       // Im(gamma) / (2 pi), the wavenumber in turns per unit length
    int lossy;         // gamma_re != 0
};

struct RuleConst {
    int nt, nb;                       // far-rule points per test / trial cell
    double lt[BMI_MAX_RULE][3];       // barycentric coordinates of test points
    double lb[BMI_MAX_RULE][3];
    double w4pi[BMI_MAX_RULE][BMI_MAX_RULE];  // w_I * w_J / (4 pi), reference weights
};

struct BlockArgs {
    DevSpace X, Y;
    const int *tp_list;  // test patches of this launch's colour
    const int *bq_list;  // trial patches of this launch's colour
    int nbq;
    int npb;             // trial patches in total (block id = p * npb + q)
    const int *blk_off;  // special-pair offsets per block
    const unsigned char *spec_loc;
    const double2 *spec_val;
    int capacity;        // special-pair entries allocated
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
    int *blk_count;      // pass 0 output
    const int *blk_off;  // pass 1 input
    int *spec_K, *spec_L;
    unsigned char *spec_cls, *spec_loc;
    unsigned long long *class_count;  // 6 counters
    int capacity;
};

struct SpecialArgs {
    DevSpace X, Y;
    const int *spec_K, *spec_L;
    const unsigned char *spec_cls;
    const int *n_special;
    int capacity;        // special-pair entries allocated
    double2 *spec_val;
    OpConst op;
    RuleConst rule;
    int ss_n[3];
    double glx[3][BMI_MAX_GL], glw[3][BMI_MAX_GL];
    int near_np;
    double nb1[BMI_MAX_NEAR], nb2[BMI_MAX_NEAR], nw[BMI_MAX_NEAR];
};

// kernel launchers (bmi_kernels.hip)
hipError_t launch_classify(const ClassifyArgs &a, int pass, hipStream_t s);
hipError_t launch_scan(const int *count, int *off, int n, int *partials, int *total, hipStream_t s);
hipError_t launch_special(const SpecialArgs &a, int capacity, int mixed, int op_kind, hipStream_t s);
hipError_t launch_blocks(const BlockArgs &a, int ntp, int op_kind, int mixed, int c32, hipStream_t s);
int scan_partials_needed(int n);
