// kernels.hpp -- host-side launchers of the HIP kernels (internal to libnekrs_hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace nrs {

// ax.hip
template <typename T>
void launch_ax(int N, int E, const int *elist, const T *ggeo, const T *D, const T *l0, const T *l1,
               bool poisson, bool lfield, const T *q, T *Aq, const uint32_t *maskBits,
               hipStream_t s);
// the same with an explicit elements-per-block multiplier (1, 2 or 4)
template <typename T>
void launch_ax_variant(int N, int mult, int E, const int *elist, const T *ggeo, const T *D,
                       const T *l0, const T *l1, bool poisson, bool lfield, const T *q, T *Aq,
                       const uint32_t *maskBits, hipStream_t s);
void launch_block_ax(int N, int E, int64_t offset, int64_t loffset, const double *ggeo,
                     const double *D, const double *l0, const double *l1, const double *q,
                     double *Aq, hipStream_t s,
                     bool lfield = true);
template <typename T>
void launch_build_diag(int N, int E, const T *ggeo, const T *D, const T *l0, const T *l1,
                       bool poisson, bool lfield, T *diag, hipStream_t s);
template <typename T>
void launch_mask(int64_t Nm, const int *ids, T *v, hipStream_t s);
// Ax whose consecutive elements e, e+1 with fx[e] = 1 sum their shared face i = N /
// i = 0 interior nodes in the kernel (N = 7); blocks are 4 elements from q's start
template <typename T>
void launch_ax_fused(int N, int E, const T *ggeo, const T *D, const T *l0, const T *l1, bool poisson,
                     const T *q, T *Aq, const uint32_t *maskBits, const uint8_t *fx, hipStream_t s, int nb = 4);

// Ax of trilinear elements with the geometric factors formed in the kernel
// from the element corners [E][3][8] and the GLL nodes and weights zw [2][N+1]
// (the tuned elements-per-block multiplier of launch_ax)
template <typename T>
void launch_ax_trilin(int N, int E, const T *corners, const T *zw, const T *D, const T *l0, const T *l1,
                      bool poisson, const T *q, T *Aq, const uint32_t *maskBits, hipStream_t s);

// mesh.hip
void launch_geofacs(int N, int E, const double *D, const double *w, const double *x,
                    const double *y, const double *z, double *ggeo, hipStream_t s);

// transfer.hip
template <typename T>
// w: w .* qf contracted; w and sub: w .* (sub - qf) (qf left unchanged in both)
void launch_coarsen(int Nf, int Nc, int E, const T *R, const T *qf, T *qc, hipStream_t s,
                    const T *w = nullptr, const T *sub = nullptr, const uint32_t *cmask = nullptr);
template <typename T>
void launch_prolongate(int Nf, int Nc, int E, const T *R, const T *qc, T *qf, hipStream_t s);

}  // namespace nrs
