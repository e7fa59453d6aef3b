// Device storage of a set of solver lanes, shared by the batch and the streaming kernels.
//
// Per lane (L lanes, B = 2^tt_log2 buckets):
//   tags   B * 64 B    tag lines; lane l's TT is one contiguous range
//   ents   B * 256 B   entries
//   stack  52 * 32 B   packed frames; a lane's neighbouring depths share 128-byte lines, so a
//                      push and the matching pop hit the same line
//   gen    4 B         TT generation, kept across launches so that tags written by an earlier
//                      call never match (no per-call clear of the tag lines)
#pragma once
#include <hip/hip_runtime.h>

#include "ddx.h"
#include "ddx/machine.hpp"
#include "ddx_internal.h"

namespace ddx {

constexpr int kStackDepth = 52;  // > the 48 expanded frames a full deal can need

// Frame stacks are lane-major: a lane's frames are contiguous, so a push and the matching pop
// hit the same line. DDX_FRAMES_DEPTH_MAJOR=1 builds the alternative layout -- frame i of every
// lane contiguous, lanes at the same depth side by side -- for measuring one against the other.
#ifndef DDX_FRAMES_DEPTH_MAJOR
#define DDX_FRAMES_DEPTH_MAJOR 0
#endif

struct GpuStack {
  PackedFrame* base;
#if DDX_FRAMES_DEPTH_MAJOR
  long long stride;  // searches
  __device__ __forceinline__ void store(int i, const Frame& f) { base[i * stride] = pack_frame(f); }
  __device__ __forceinline__ Frame load(int i) const { return unpack_frame(base[i * stride]); }
#else
  __device__ __forceinline__ void store(int i, const Frame& f) { base[i] = pack_frame(f); }
  __device__ __forceinline__ Frame load(int i) const { return unpack_frame(base[i]); }
#endif
};

// Largest TT_LOG2: per search, or, in the shared-TT build, of the one TT (2^28 buckets = 155 GB).
constexpr int kMaxTTLog2 = DDX_SHARED_TT ? 28 : 20;

struct LaneStore {
  u32* tags = nullptr;
  Entry* ents = nullptr;
  PackedFrame* stacks = nullptr;
  u32* gen = nullptr;
  long long lanes = 0;
  int tt_log2 = 0;
  uint64_t bytes = 0;
};

// Device view handed to kernels.
struct LaneView {
  u32* tags;
  Entry* ents;
  PackedFrame* stacks;
  u32* gen;
  int tt_log2;
  long long searches;
};

__device__ __forceinline__ void lane_attach(const LaneView& v, long long lane, Lane& L, GpuStack& st) {
  const size_t buckets = (size_t)1 << v.tt_log2;
  TTView tt;
#if DDX_SHARED_TT
  tt.tags = v.tags;  // one TT of 2^tt_log2 buckets serves every search
  tt.ents = v.ents;
#else
  tt.tags = v.tags + (size_t)lane * buckets * kWays;
  tt.ents = (Entry*)((char*)v.ents + (size_t)lane * buckets * kBucketEntryBytes);
#endif
  tt.mask = (u32)(buckets - 1);
#if DDX_FRAMES_DEPTH_MAJOR
  st.base = v.stacks + lane;
  st.stride = v.searches;
#else
  st.base = v.stacks + (size_t)lane * kStackDepth;
#endif
  lane_init(L, tt);
  L.c.gen = v.gen[lane];
}

// Fills n 16-byte words with a 32-bit pattern (a kernel rather than hipMemset, so that it can
// sit in a captured graph).
__global__ void fill_kernel(Q4* p, size_t n, u32 v);
__global__ void fill_bytes_kernel(uint8_t* p, size_t n, uint8_t v);
int launch_fill(void* p, size_t bytes, u32 v, hipStream_t s);

int lanes_alloc(LaneStore& s, long long lanes, int tt_log2);
void lanes_free(LaneStore& s);
inline LaneView lanes_view(const LaneStore& s) {
  return LaneView{s.tags, s.ents, s.stacks, s.gen, s.tt_log2, s.lanes};
}

}  // namespace ddx

#define HIPCHECK(x)                                                                 \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      ddx_set_error("%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
      return DDX_E_HIP;                                                             \
    }                                                                               \
  } while (0)
