/* kernels.h - host-side launchers of the kernels in kernels.hip */
#pragma once
#include "common.h"

hipError_t ilm_k_ctl_init(const KParams& P, hipStream_t s);
hipError_t ilm_k_raygen(const KParams& P, int grid, hipStream_t s);
hipError_t ilm_k_extend(const KParams& P, int depth, int* ovf, int grid, hipStream_t s);
hipError_t ilm_k_path_clear(const KParams& P, int grid, hipStream_t s);
hipError_t ilm_k_shade(const KParams& P, int depth, int grid, hipStream_t s);
hipError_t ilm_k_accumulate(const KParams& P, int grid, hipStream_t s);
hipError_t ilm_k_chunk_end(const KParams& P, unsigned long long* seg_total, hipStream_t s);
hipError_t ilm_k_tonemap(const float4* accum, ilm_pixel* out, int n, int spp, int op, hipStream_t s);
hipError_t ilm_k_fill_zero4(float4* p, size_t n, hipStream_t s);
hipError_t ilm_k_sums_to_rgb(const float4* sums, float* rgb, int n, hipStream_t s);
int ilm_k_extend_occupancy(int tri_test, int bvh_layout, int extend_mode);
int ilm_k_shade_occupancy(int sort);
int ilm_k_ovf_stack();
int ilm_k_lds_stack();
int ilm_k_ext_block();
