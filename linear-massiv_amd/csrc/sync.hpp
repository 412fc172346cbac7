// Data exchange between the workgroups of one running kernel.
//
// An agent-scope release or acquire fence on gfx950 compiles to a write-back
// (buffer_wbl2 sc1) or an invalidate (buffer_inv sc1) of the XCD's whole L2:
// microseconds each, and the invalidate evicts everything the XCD had cached.
// The kernels here therefore use no such fences for their hand-offs.  Every
// word that another workgroup reads while the kernel runs is stored and
// loaded as a relaxed device-scope atomic, which the compiler emits as a
// plain vector access with the sc1 (device-coherent) cache policy.  Such an
// access is performed at the device coherence point, not in the XCD-private
// L2.  The protocol:
//   writer: xstore() the data; xdrain() (s_waitcnt vmcnt(0): this wave's
//           stores have completed); then xstore() the flag;
//   reader: spin on xload() of the flag; then xload() the data.  Its data
//           loads are issued only after the flag load has returned (a
//           workgroup barrier or xdrain() separates them).
// This is the ordering the compiler itself relies on for an agent-scope
// release store (wait for completion, then store) once the write-back is
// unnecessary because every earlier store was already device-coherent.
#pragma once

#include <hip/hip_runtime.h>

namespace lmx {

template <class T>
__device__ inline void xstore(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ inline T xload(const T *p) {
    return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Waits until every memory access this wave issued has completed; also a
// compiler barrier for memory operations.
__device__ inline void xdrain() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

}  // namespace lmx
