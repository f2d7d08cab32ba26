// sphg_kernels.h -- host-side launchers of the kernels in sphg_kernels.hip.
// Each takes the scene table and chunk table of one batch launch.
#pragma once

#include "sphg_internal.h"

namespace sphg {

// begin_active: also the start of a WCSPH or PCISPH step, counting the
// scenes whose pressure loop runs in begin_active[0] (null: keys only)
void launch_keys(hipStream_t st, const DevScene* T, const Chunk* C, int nchunks, int* begin_active = nullptr);
// maxChunks: the largest cell-chunk count of one scene; up to
// kScanFusedChunks each final-scan block sums the totals before its chunk
// itself (one launch fewer), above it a per-scene scan of the totals runs
constexpr int kScanFusedChunks = 1024;
void launch_scan(hipStream_t st, const DevScene* T, const Chunk* CC, int ncellchunks, int nscenes, int maxChunks);
void launch_scatter(hipStream_t st, const DevScene* T, const Chunk* C, int nchunks);
void launch_rank_reorder(hipStream_t st, const DevScene* T, const Chunk* C, int nchunks);
void launch_bvolume(hipStream_t st, const DevScene* T, int scene, int nb, float* tmp);
void launch_build(hipStream_t st, int method, const DevScene* T, const Chunk* C, int nchunks);
void launch_normals(hipStream_t st, const DevScene* T, const Chunk* C, int nchunks);
void launch_forces(hipStream_t st, int method, const DevScene* T, const Chunk* C, int nchunks);
void launch_dt(hipStream_t st, const DevScene* T, const Chunk* C, int nchunks);
void launch_df_predict(hipStream_t st, const DevScene* T, const Chunk* C, int nchunks);
void launch_df_vel(hipStream_t st, int mode, const DevScene* T, const Chunk* C, int nchunks);
void launch_df_err(hipStream_t st, int mode, const DevScene* T, const Chunk* C, int nchunks);
// kind: 0 dt, 1 density (before loop), 2 density, 3 divergence (before loop), 4 divergence, 5 PCISPH
void launch_finalize(hipStream_t st, int kind, const DevScene* T, int nscenes, int* active);
// also clears active[0..1] for the divergence solve
void launch_df_positions(hipStream_t st, const DevScene* T, const Chunk* C, int nchunks, int* active);
void launch_pci_density(hipStream_t st, const DevScene* T, const Chunk* C, int nchunks);
void launch_pci_pacc(hipStream_t st, const DevScene* T, const Chunk* C, int nchunks, int iter);
void launch_pci_integrate(hipStream_t st, const DevScene* T, const Chunk* C, int nchunks);
void launch_cube(hipStream_t st, const DevScene* T, const Chunk* C, int n);
void launch_emit(hipStream_t st, DevScene* Tc, DevScene* To, int nscenes, int cells);
// PCISPH boundary method
void launch_pb_bself(hipStream_t st, const DevScene* T, int scene, int nb);
void launch_pb_bdensity(hipStream_t st, int pred, const DevScene* T, int scene, int nb, int iter);
void launch_gather_f3(hipStream_t st, float4* dst, const float* src, const uint32_t* id, int n);
void launch_pb_correct(hipStream_t st, int mode, const DevScene* T, const Chunk* C, int n);
void launch_pb_predcorrect(hipStream_t st, const DevScene* T, const Chunk* C, int n);
void launch_time_advance(hipStream_t st, const DevScene* T, int nscenes);
void launch_clear_overflow(hipStream_t st, const DevScene* T, int nscenes);
void launch_fill_int(hipStream_t st, int* p, int v, long n);
void launch_unsort_f4(hipStream_t st, const float4* src, const uint32_t* id, float* dst, int n, int comps);
void launch_unsort_f1(hipStream_t st, const float* src, const uint32_t* id, float* dst, int n, double offset = 0.0);
void launch_unsort_i1(hipStream_t st, const int* src, const uint32_t* id, int* dst, int n);
void launch_keys3(hipStream_t st, const DevScene* T, int scene, int n, int* dst);
void launch_load_state(hipStream_t st, const DevScene* T, int scene, int n, const float* xyz, const float* vel,
                       const float* k, const float* kv);

}  // namespace sphg
