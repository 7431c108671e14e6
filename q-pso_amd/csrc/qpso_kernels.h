// qpso_kernels.h -- host-side launchers of the gfx950 kernels.
#ifndef QPSO_KERNELS_H_
#define QPSO_KERNELS_H_

#include <hip/hip_runtime_api.h>

#include "qpso_internal.h"

namespace qpso_launch {

// Render+fitness kernel variants: (threads per workgroup, int32 tile capacity).
int rf_num_configs();
int rf_config_threads(int cfg);
int rf_config_tile(int cfg);

// Every launcher takes optional events that the launch stamps with the
// kernel's own start and end (hipExtLaunchKernel), for per-kernel timing.

// InitQPSO for every solve in d_solves[0..n).
hipError_t init(const QpsoSolveDev* d_solves, int n_solves, hipStream_t st, hipEvent_t ev0 = nullptr,
                hipEvent_t ev1 = nullptr);
// Whether a mesh of v_eff vertices needs the window-by-window instantiation.
bool rf_needs_windows(int cfg, int v_eff);
// One fused render+fitness pass; d_items[i] = (solve index, particle index),
// or d_items = nullptr for workgroup i = particle i of d_solves[0].
// mode 0: meshes that fit the vertex cache; 1: any mesh (window by window
// beyond the cache); 2: any mesh with the frame dump of the stage call.
hipError_t render_fitness(int cfg, int mode, const QpsoSolveDev* d_solves, const QpsoItem* d_items, int n_items,
                          hipStream_t st, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
// The reference-faithful off-frame mode's render+fitness pass (the solves'
// full per-particle buffers in device memory, linear-index wrap).
hipError_t render_fitness_wrap(const QpsoSolveDev* d_solves, const QpsoItem* d_items, int n_items, hipStream_t st,
                               hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
// Sets n_words depth words to the empty value 60000 (a kernel, not a memset).
hipError_t fill_frames(int32_t* d_frames, size_t n_words, hipStream_t st);
// argmin + pbest/gbest + particle update, one workgroup per solve.
hipError_t pso_step(const QpsoSolveDev* d_solves, int n_solves, hipStream_t st, hipEvent_t ev0 = nullptr,
                    hipEvent_t ev1 = nullptr);

// The address of `device`'s copy of the allocation-free entry point's
// scratch (QPSO_NOALLOC_SCRATCH_BYTES); `device` must be current at the first
// call for it, which takes the address once.
hipError_t scratch_address(int device, void** ptr);
// Per-row observed-pixel prefix counts of a device crop (rw x rh uint16 ->
// rh x (rw + 1) uint16).
hipError_t obs_rows(const uint16_t* d_crop, uint16_t* d_rows, int rw, int rh, hipStream_t st);

}  // namespace qpso_launch

#endif
