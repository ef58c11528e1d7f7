// zframe_dev.hpp -- host interface of the device zstd-format frame builder
// (zframe.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace xnyz {

struct ZBlk {  // one zstd block: bytes [src, src + n) of the streams buffer
  u64 src;
  u32 n;
  u32 frame;
};

struct ZFrame {  // one frame: content size, its blocks [first, first + nblk)
  u64 len;
  u32 first;
  u32 nblk;
};

// Device scratch (bytes) for nblk blocks.
size_t zframe_scratch_bytes(u32 nblk);

// Frames of every ZFrame into `out`: frame f occupies
// [pos[frames[f].first] + 13 f, pos[frames[f].first + frames[f].nblk] + 13 (f + 1)),
// pos holds nblk + 1 entries (pos[nblk] = total block bytes).  blks, frames,
// scratch, out and pos are device memory; src offsets 4-byte aligned.  With
// h_pos (pinned host memory, nblk + 1 entries) the positions are also
// copied there as soon as they are known -- before the blocks are gathered
// into `out` -- and `sizes_ready` is recorded after that copy.
void zframe_encode(const u8* pre, const ZBlk* blks, u32 nblk, const ZFrame* frames,
                   void* scratch, u8* out, u64* pos, Arena& ar, hipStream_t s,
                   u64* h_pos = nullptr, hipEvent_t sizes_ready = nullptr);

}  // namespace xnyz
