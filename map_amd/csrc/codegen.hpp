// Code generation: one fused group -> HIP source instantiating the templates
// of device/mapx_dev.hiph, compiled by hiprtc for gfx950.
//
// Replaces the reference's OpenCL skeleton generators
// (runtime/skeleton/{Skeleton,LocalSkeleton,FocalSkeleton,ZonalSkeleton}.cpp):
// node statements are still emitted one per node in node-id order with the
// node's promoted type, but the loops, tiling, loads and reductions are the
// hand-written templates.
#pragma once
#include "fuser.hpp"
#include <string>
#include <vector>

namespace mapx {

enum class Form : int { Local = 0, Focal = 1 };

// One member of the kernel's argument struct (P::Args), in declaration order.
struct Param {
  // Mask: the weights of a convolution with a floating mask, an array in
  // P::Args (so kernels do not depend on the weights); integer masks are
  // compiled in as literals
  enum Kind : int { InD2, InD0, OutD2, Mask } kind;
  Node* node = nullptr;
  int dt = 0;
  size_t off = 0, size = 0;  // byte offset and size inside P::Args
};

// A zonal output of the group.
struct ZonalOut {
  Node* node;      // ZonalReduc or ZonalStats
  int red;         // reduction (ZonalReduc) or 0 (ZonalStats)
  int dt;          // data type of the reduced values
  int acc_dt;      // accumulator type (F64 for floating SUM/PROD)
  int nzones;      // ZonalStats only
  size_t result_off;  // byte offset of the final value(s) in the result buffer
};

struct TileShape {
  int TX = 64, TY = 16, VX = 1;
  int PF = 0;  // focal groups: load the next tile's rows while computing the current one
  int ST = 0;  // focal groups that stage only inputs: rows in registers (stream_tile) instead of LDS
  int W = 0;   // wavefronts per SIMD the kernel's registers must allow (0: as the compiler allocates)
  bool operator==(const TileShape& o) const {
    return TX == o.TX && TY == o.TY && VX == o.VX && PF == o.PF && ST == o.ST && W == o.W;
  }
};

struct KernelSpec {
  std::string name, fin_name;  // main kernel; zonal finalisation kernel ("" if none)
  std::string source;
  Form form = Form::Local;
  bool stream = false;  // focal group run by stream_tile (register row window) rather than the LDS tile
  unsigned lane_r = 0;  // byte-lane form for staged values in [0, lane_r] (0: none)
  TileShape tile;
  int H0 = 0, H1 = 0;
  std::vector<Param> params;
  size_t args_size = 0;           // sizeof(P::Args)
  size_t task_size = 0;           // sizeof(mx::TaskRec<P::Args>)
  size_t task_launch_off = 0;     // offset of TaskRec::L
  std::vector<Node*> d2_outs;     // materialised D2 outputs, in parameter order
  std::vector<ZonalOut> zouts;
  size_t part_bytes = 0;          // bytes of one chunk's partial record (= one task's result record)
  std::string signature;          // cache key (group signature + tile shape)
};

// Kernel parameters (both kernels are extern "C", launched with these):
//   main: (P::Args a0, mx::Launch L0, const mx::TaskRec<P::Args>* tasks,
//          const int* starts, int ntasks, int nchunks, mx::u32* queue, char* zpart)
//   fold: (const char* src, char* dst, const mx::TaskRec<P::Args>* tasks,
//          const int* starts, int n_single, int stride, int F)
// Builds the kernel of group `g`. `d2_outs` are the D2 nodes the group must
// write to HBM (group outputs consumed later, written to files or requested).
// `libm`: FP32 sqrt, atan2 and sin/cos pairs call the device library instead of
// the fast paths of mapx_dev.hiph (mx::sqrt_f32, mx::atan2_f32, mx::sincos_f32).
KernelSpec generate_kernel(const Group& g, const std::vector<Node*>& d2_outs, TileShape tile, bool libm = false);

// The group's nodes in the kernel's local numbering: inputs (D2, then D0,
// each by id), then the group's nodes by id.
std::vector<Node*> kernel_order(const Group& g);
// Key of the kernel generate_kernel builds for `g`: groups with equal keys
// (node signatures and links by local number, materialised outputs, tile,
// libm) get the same source and parameter layout, so one compiled kernel
// serves them all with its node references rebound by local number.
std::string kernel_shape_key(const Group& g, const std::vector<Node*>& d2_outs, TileShape tile, bool libm);

// Default tile shape for a group: VX cells per thread along x so that a
// wavefront row covers >= 256 contiguous bytes.
TileShape default_tile(const Group& g);

// Reference quantisation of floating mask weights and D2 constant literals:
// printed with std::to_string (%f, 6 decimals) into the kernel source
// (runtime/skeleton/util.cpp:136-165, util/VariantType.cpp:91) and read back.
Variant quantize_literal(const Variant& v);

}  // namespace mapx
