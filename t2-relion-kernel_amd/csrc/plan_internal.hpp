// Internal view of a batched plan, shared by the C ABI (capi.cpp) and the
// serving engine (server.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "rc_internal.hpp"

struct rc_plan {
    int P, C, O, npix, full_npix, form, kpad, nkc;
    int variant;
    rc::VariantInfo info;
    std::vector<int> trans_offset;          // host copy
    std::vector<int> proj_dims;             // per class: 7 ints for shape checks
    int n_groups, n_mtiles, ntiles_per_class, parts_per_group;
    int n_rows;          // sum of the translation counts
    bool packed;         // wide kernels: rows packed without per-particle padding
    unsigned char *mem = nullptr;
    size_t bytes = 0;
    int2 *groups;
    int *group_begin;
    int *row_part;
    rc::ProjView *vols;
    float *bscale;
    float2 *xw, *coords;
    double *xnorm;
    float *half_corr, *ascale;
    unsigned char *proj;
    float *pabs2, *pnorm;
    double *part_min;
    long long *part_idx;
    unsigned char *a_enc;                   // schedule-8 variants: precomputed A
    int a_chunks;                           // 16-pixel chunks per row tile of a_enc
    std::vector<hipEvent_t> ev;             // timing slots: 2 per launch
    long long launches = 0;
    bool capturing = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int persist_blocks = 0;
};

namespace rc {

// Host-side planning: shapes, row groups, variant, tables to upload.
struct PlanTables {
    std::vector<int2> groups;
    std::vector<int> group_begin;
    std::vector<int> row_part;
    std::vector<ProjView> views;
    std::vector<float> scales;
};
// forced_variant >= 0 picks that cross-term variant; -1 follows
// rc_set_variant_override, then the autotuner's table, then the size rule.
int plan_setup(rc_plan *p, const rc_batch_desc *d, PlanTables &tab, int forced_variant = -1);
// Lay the plan's device buffers out from `mem` (NULL: only size them);
// returns the bytes needed.
size_t plan_bind(rc_plan *p, unsigned char *mem);
// Asynchronous upload of the tables from pinned host memory `staging` of
// at least plan_table_bytes(tab) bytes.
size_t plan_table_bytes(const PlanTables &tab);
int plan_upload_async(rc_plan *p, const PlanTables &tab, unsigned char *staging, hipStream_t s);
// Validate desc against the plan and launch the kernels on s.
int plan_run(rc_plan *p, const rc_batch_desc *d, hipStream_t s);
int validate_batch(const rc_batch_desc *d, bool need_ptrs);
// rc_plan_create with an explicit variant (-1: the automatic choice).
int plan_create_internal(const rc_batch_desc *d, int forced_variant, rc_plan **out);

}  // namespace rc
