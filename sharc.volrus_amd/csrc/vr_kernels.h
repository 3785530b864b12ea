// Launch wrappers for the gfx950 kernels (host side of vr_kernels.hip).
#pragma once

#include "vr_device.h"

hipError_t vr_upload_mc_tables(const uint16_t* edge_mask, const uint8_t* ntri, const int8_t* tri,
                               const int8_t* corner, const int8_t* edge, const int8_t* cand);

hipError_t vr_fill_int(int* p, int v, size_t n, hipStream_t s);
hipError_t vr_build_child(const NanoLeaf* leaves, const int* count, int cap, const int4* roots, int root_count,
                          int* child, int4* origins, hipStream_t s);
hipError_t vr_build_nbr(const GridView& g, int cap, int* nbr, hipStream_t s);

size_t vr_scan_scratch_ints(size_t n);
// n_dev (optional, single-block sizes only): scan min(n, *n_dev * n_mult)
// elements, a device-side count below the host bound n.
hipError_t vr_exclusive_scan(const int* in, int* out, size_t n, int* scratch, int* total, hipStream_t s,
                             const int* n_dev = nullptr, int n_mult = 1);
// Two arrays of the same length in one pass when they fit one block.
hipError_t vr_exclusive_scan2(const int* in0, int* out0, int* total0, const int* in1, int* out1, int* total1,
                              size_t n, int* scratch, hipStream_t s, const int* n_dev = nullptr);

// nw = waves per leaf (1 or 4), see vr_advect_waves.
hipError_t vr_launch_advect(const GridView& in, const VelView& vel, NanoLeaf* out, int cap, double dt,
                            const double tf[3], int scheme, int nw, hipStream_t s);
// Waves per leaf for an advect over `leaves` leaves on `cus` compute units:
// one (fewest instructions per leaf) once there are enough leaves to give
// every CU 16 waves, else several (a leaf finishes sooner; the launch picks
// how many per velocity kind).
inline int vr_advect_waves(long long leaves, int cus) { return leaves >= 16LL * (cus > 0 ? cus : 256) ? 1 : 8; }
hipError_t vr_launch_crossings(const GridView& g, int cap, uint64_t* cross, int* any, int gen, hipStream_t s);
hipError_t vr_launch_rebuild(const GridView& g, int cap, const uint64_t* cross, const int* any, int gen, NanoLeaf* out,
                             int* keep, int keep_len, int R, float vs, float band, unsigned long long* active_out,
                             hipStream_t s);
hipError_t vr_launch_trim(const GridView& g, int cap, NanoLeaf* out, int* keep, float threshold, hipStream_t s);
hipError_t vr_launch_nonempty(const NanoLeaf* leaves, const int* count, int cap, int* keep, hipStream_t s);
hipError_t vr_launch_compact(const NanoLeaf* src, const int* count, int cap, const int* keep, const int* pos,
                             NanoLeaf* dst, hipStream_t s, int* child_fill = nullptr, size_t nchild = 0);
hipError_t vr_copy_int(const int* src, int* dst, hipStream_t s, int* dst2 = nullptr);
hipError_t vr_launch_compact_topo(const NanoLeaf* src, const int* count_old, int cap, const int* keep, const int* pos,
                                  const int* total, NanoLeaf* dst, const int4* roots, int root_count, int* child,
                                  int4* origins, const int* nbr_old, int* nbr_new, int* count_out, int* count_out2,
                                  hipStream_t s);
hipError_t vr_zero_tail(int* p, const int* count, int n, hipStream_t s);
hipError_t vr_launch_popcount_words(const uint64_t* w, const int* count, int cap, int* out, hipStream_t s);
hipError_t vr_launch_crossing_coords(const int4* origins, const int* count, int cap, const uint64_t* cross,
                                     const int* pos, int* xyz, hipStream_t s);

hipError_t vr_launch_leaf_popcount(const NanoLeaf* leaves, const int* count, int cap, int* out, hipStream_t s);
