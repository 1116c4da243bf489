// Grouped launches: one kernel launch that runs the blocks of many independent invocations of
// the same kernel, one per system of a batched ensemble.
//
// Every pipeline kernel that the ensemble uses takes its arguments as one struct A and reads
// its block index and block count from a Blk, never from blockIdx / gridDim directly. The
// plain launch passes Blk{blockIdx.x, gridDim.x}. The grouped launch gets a device table of
// GroupEntry<A>: entry e owns blocks [b0, b0 + nb) of the launch (b0 ascending, each a
// multiple of 8 so that block % 8 is still the XCD the hardware dispatched it to), and its
// blocks run the body with Blk{blockIdx.x - b0, nb} and the entry's own arguments. A body
// synchronises only within its block, so the entries run independently.
//
// On the host, submit() launches the plain kernel, or, while a Recorder is attached to the
// context, records the invocation (tag, kernel, geometry, argument bytes) instead. The batch
// runtime records every system's pipeline, merges the invocations with the same tag and
// kernel into one grouped launch, and captures the merged sequence.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstring>
#include <vector>

#include "ffcm_internal.hpp"

namespace ffcm {

struct Blk {
  int id, n;
};

__device__ __forceinline__ Blk plain_blk() { return Blk{(int)blockIdx.x, (int)gridDim.x}; }

template <typename A>
struct GroupEntry {
  int b0, nb;
  A a;
};

// Entry owning this block: the last e with g[e].b0 <= blockIdx.x (binary search on a
// wave-uniform index, so every load is a scalar load).
template <typename A>
__device__ __forceinline__ int group_find(const GroupEntry<A>* __restrict__ g, int n) {
  const int b = (int)blockIdx.x;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (g[mid].b0 <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// A kernel K is written as a body K_body(const Blk&, const A&) and two entry points:
//   __global__ void K(A a)                                    { K_body(plain_blk(), a); }
//   __global__ void K_group(const GroupEntry<A>* g, int n)    { group_run(g, n, K_body<...>); }
// Blocks past an entry's nb (padding up to the next entry's multiple-of-8 start) exit at once.
template <typename A, typename Body>
__device__ __forceinline__ void group_run(const GroupEntry<A>* __restrict__ g, int n, Body body) {
  const GroupEntry<A>& ge = g[group_find(g, n)];
  const int id = (int)blockIdx.x - ge.b0;
  if (id >= ge.nb) return;
  body(Blk{id, ge.nb}, ge.a);
}

// Pipeline step of a recorded launch: invocations of different systems with the same tag and
// kernel are merged; tags increase along each system's pipeline.
enum StepTag {
  kTagWrapAssign = 100,
  kTagSortPass = 200,  // + 10 * pass + {0 histogram, 1..4 scan, 5 scatter}
  kTagGatherBin = 300,
  kTagGatherForces = 310,
  kTagBinScan = 320,   // + 1..4 scan
  kTagBinScatter = 330,
  kTagSpread = 1000,    // + first right-hand side of the launch (fewer than 4096)
  kTagFft = 9000,       // + pass
  kTagInterp = 10000,   // + right-hand side
  kTagCorrect = 20000,  // + first right-hand side
};

// Scan steps inside a tag block: offsets added to the scan's base tag.
enum { kScanSingle = 1, kScanReduce = 2, kScanPartials = 3, kScanApply = 4 };

struct LaunchRecord {
  int tag;
  const void* kernel;  // the plain kernel: identifies the instance
  const void* group;   // its grouped form
  unsigned grid;
  unsigned threads;
  size_t lds;
  int64_t work;       // for a persistent kernel, the tiles behind its grid (> grid); else -1
  size_t args_offset;  // offsetof(GroupEntry<A>, a)
  size_t entry_size;   // sizeof(GroupEntry<A>)
  std::vector<unsigned char> args;
};

struct Recorder {
  std::vector<LaunchRecord> launches;
};

template <typename A>
hipError_t submit(ffcm_ctx* c, int tag, void (*plain)(A), void (*group)(const GroupEntry<A>*, int), unsigned grid,
                  unsigned threads, size_t lds, const A& a, int64_t work = -1) {
  if (c->rec) {
    LaunchRecord r;
    r.tag = tag;
    r.kernel = (const void*)plain;
    r.group = (const void*)group;
    r.grid = grid;
    r.threads = threads;
    r.lds = lds;
    r.work = work;
    r.args_offset = offsetof(GroupEntry<A>, a);
    r.entry_size = sizeof(GroupEntry<A>);
    r.args.resize(sizeof(A));
    std::memcpy(r.args.data(), &a, sizeof(A));
    c->rec->launches.push_back(std::move(r));
    return hipSuccess;
  }
  if (grid == 0) return hipSuccess;
  plain<<<grid, threads, lds, c->stream>>>(a);
  return hipGetLastError();
}

}  // namespace ffcm
