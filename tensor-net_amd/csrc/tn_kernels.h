// Host-side launchers for the kernels in tn_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "tn_device.h"

namespace tnmps {

void launch_theta(const UpdDesc* descs, int nt, int gx, hipStream_t s);
void launch_winit(const UpdDesc* descs, int nt, int gx, hipStream_t s);
// Jacobi step with block width bw (8 or 16): one workgroup per block pair.
void launch_jacobi_step(const UpdDesc* descs, int nt, int npairs, int bw, int sweep, int step,
                        double tol_rot, double conv_tol, hipStream_t s);
// Row-split step: S workgroups per block pair (gram, then solve + apply).
void launch_jacobi_gram(const UpdDesc* descs, int nt, int npairs, int bw, int S, int sweep,
                        int step, double conv_tol, hipStream_t s);
void launch_jacobi_apply(const UpdDesc* descs, int nt, int npairs, int bw, int S, int sweep,
                         int step, double tol_rot, double conv_tol, hipStream_t s);
// Fused row-split step with block width 8 (cooperative launch; pair barrier).
hipError_t launch_jacobi_fused(const UpdDesc* descs, int nt, int npairs, int S, int sweep,
                               int step, double tol_rot, double conv_tol, hipStream_t s);
// Largest grid of k_jacobi_fused that is co-resident (0: no cooperative launch).
int jacobi_fused_max_blocks();
hipError_t launch_jacobi_coop(const UpdDesc* descs, int nt, int npairs, int steps,
                              int max_sweeps, double tol_rot, double conv_tol, int* sweeps_out,
                              int grid, hipStream_t s);
// Largest co-resident grid of k_jacobi_coop on the current device (0 when
// cooperative launch is unsupported).
int jacobi_coop_max_blocks();
void launch_finish(const UpdDesc* descs, int nt, const int* sweeps_dev, int sweeps_run,
                   hipStream_t s);
// Factored updates: T = kept singular vectors x isometric factor; grid bounds
// maxX >= max(m, n) and maxRank >= rank over the tasks.
void launch_fmul(const UpdDesc* descs, int nt, int maxX, int maxRank, hipStream_t s);
void launch_split(const UpdDesc* descs, int nt, hipStream_t s);
void launch_renorm(const UpdDesc* descs, int nt, hipStream_t s);
void launch_measure(const MeasDesc* descs, int nm, hipStream_t s);
void launch_entropy(const double* lam, const int* dimp, double a, double* out, hipStream_t s);
void launch_u1(double2* g, const int* dims, const double* op8, int gx, hipStream_t s);

}  // namespace tnmps
