// common.hpp -- shared device helpers and host plumbing of the gfx950 library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>

#include "../../include/vren_hip/capi.h"
#include "../../include/vren_hip/vmath.h"

#define VREN_WAVE 64

// ---------------------------------------------------------------------------------------------------------------
// device: wave64 primitives
// ---------------------------------------------------------------------------------------------------------------

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Global-address-space views for pointers held in records (frame tables): loads through them compile to global_load
// instead of flat_load, whose completion also counts against the LDS wait counter. Native vector types, since HIP's
// float4/uint4 copy through a generic reference.
typedef float vf4 __attribute__((ext_vector_type(4)));
typedef uint32_t vu4 __attribute__((ext_vector_type(4)));
typedef uint32_t vu2 __attribute__((ext_vector_type(2)));
typedef const __attribute__((address_space(1))) float gfloat;
typedef const __attribute__((address_space(1))) uint32_t guint;
typedef const __attribute__((address_space(1))) vf4 gvf4;
typedef const __attribute__((address_space(1))) vu4 gvu4;
typedef const __attribute__((address_space(1))) vu2 gvu2;

// index of the calling wave in its workgroup, as a wave-uniform value: the compiler then keeps what is derived
// from it (tile and frame indices, frame-record fields) in scalar registers and reads such records with scalar loads
__device__ __forceinline__ uint32_t wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x / VREN_WAVE); }

// popcount of mask bits below this lane
__device__ __forceinline__ uint32_t mbcnt(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

template <typename T>
__device__ __forceinline__ T shfl(T v, int src) { return __shfl(v, src, VREN_WAVE); }

template <typename T>
__device__ __forceinline__ T shfl_up(T v, unsigned d) { return __shfl_up(v, d, VREN_WAVE); }

template <typename T>
__device__ __forceinline__ T shfl_xor(T v, int m) { return __shfl_xor(v, m, VREN_WAVE); }

// one DPP move of v (source lanes outside the row, or rows that row_mask leaves out, give 0)
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dpp0(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xF, true);
}

// inclusive prefix sum over the wave on the DPP network, with no lane-address registers (a shuffle-based scan keeps
// six per-lane permute addresses live, which the register-capped persistent kernels spilled to scratch):
// row_shr 1, 2, 4, 8 scan each row of 16 lanes; row_bcast:15 adds lane 15's sum to row 1 and lane 47's to row 3,
// row_bcast:31 adds lane 31's to rows 2 and 3. Every lane of the wave must be active.
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v)
{
    v += dpp0<0x111, 0xF>(v); // row_shr:1
    v += dpp0<0x112, 0xF>(v); // row_shr:2
    v += dpp0<0x114, 0xF>(v); // row_shr:4
    v += dpp0<0x118, 0xF>(v); // row_shr:8
    v += dpp0<0x142, 0xA>(v); // row_bcast:15
    v += dpp0<0x143, 0xC>(v); // row_bcast:31
    return v;
}

// one DPP move of v; lanes of rows that ROW_MASK leaves out get `fill` (every source lane of these patterns is
// inside its row)
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dpp_fill(uint32_t v, uint32_t fill)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)fill, (int)v, CTRL, ROW_MASK, 0xF, false);
}

// Wave reduction on the DPP network, returned to every lane as a wave-uniform value: quad_perm [1,0,3,2] and
// [2,3,0,1], row_half_mirror and row_mirror leave every lane with its row's result; row_bcast:15 folds row 0 into
// row 1 and row 2 into row 3, row_bcast:31 folds lane 31 (rows 0-1) into row 3; lane 63 then holds the wave's.
// `id` is the operation's identity (what the rows left out by a row mask contribute). Every lane of the wave must be
// active.
template <typename Op>
__device__ __forceinline__ uint32_t wave_reduce_dpp(uint32_t v, uint32_t id, Op op)
{
    v = op(v, dpp_fill<0xB1, 0xF>(v, id));  // quad_perm [1,0,3,2]
    v = op(v, dpp_fill<0x4E, 0xF>(v, id));  // quad_perm [2,3,0,1]
    v = op(v, dpp_fill<0x141, 0xF>(v, id)); // row_half_mirror
    v = op(v, dpp_fill<0x140, 0xF>(v, id)); // row_mirror
    v = op(v, dpp_fill<0x142, 0xA>(v, id)); // row_bcast:15
    v = op(v, dpp_fill<0x143, 0xC>(v, id)); // row_bcast:31
    return __builtin_amdgcn_readlane(v, VREN_WAVE - 1);
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
    return wave_reduce_dpp(v, 0u, [](uint32_t a, uint32_t b) { return a + b; });
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v)
{
    return wave_reduce_dpp(v, 0xFFFFFFFFu, [](uint32_t a, uint32_t b) { return a < b ? a : b; });
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// 128-bit streaming (non-temporal) load / store of 4 uint32 at a 16-byte aligned address
__device__ __forceinline__ uint4 ntload4(const void* p)
{
    u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void ntstore4(void* p, uint4 q)
{
    u32x4 v = {q.x, q.y, q.z, q.w};
    __builtin_nontemporal_store(v, reinterpret_cast<u32x4*>(p));
}

// ---------------------------------------------------------------------------------------------------------------
// device: agent-scope atomics for cross-workgroup (cross-XCD) publication.
// Look-back status words carry their state and their value in ONE word, so they need no ordering with other
// memory: relaxed atomics at agent scope (coherent across the XCDs' L2s) are enough, and they avoid the L2
// write-back / invalidate that release / acquire at agent scope cost on this part.
// ---------------------------------------------------------------------------------------------------------------

__device__ __forceinline__ uint32_t ld_agent_u32(const uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_agent_u64(const uint64_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent_u32(uint32_t* p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent_u64(uint64_t* p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_add_agent(uint32_t* p, uint32_t v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_or_agent(uint32_t* p, uint32_t v)
{
    __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// f16 -> f32 on the conversion unit (exact); Inf/NaN halves take the shared software path, whose NaN payload
// handling is the oracle's
__device__ __forceinline__ float vren_h2f(uint16_t h)
{
    if ((h & 0x7C00u) == 0x7C00u) return vren_half_to_float(h);
    _Float16 x;
    __builtin_memcpy(&x, &h, 2);
    return (float)x;
}

// min / max in the total order -0 < +0 (order-independent, so any reduction tree gives the same bits)
__device__ __host__ __forceinline__ float tmin(float a, float b)
{
    if (a < b) return a;
    if (b < a) return b;
    return (vren_f32_bits(a) >> 31) ? a : b;
}
__device__ __host__ __forceinline__ float tmax(float a, float b)
{
    if (a > b) return a;
    if (b > a) return b;
    return (vren_f32_bits(a) >> 31) ? b : a;
}

// look-back status words (64 bit): [63:62] state, [31:0] value
#define VREN_LB_EMPTY 0ull
#define VREN_LB_AGG (1ull << 62)
#define VREN_LB_INCL (2ull << 62)
#define VREN_LB_STATE(s) ((s) & (3ull << 62))

// device capacity flags (vren_ctx::d_flags)
#define VREN_FLAG_KEY_CAPACITY 1u
#define VREN_FLAG_INDEX_CAPACITY 2u

// ---------------------------------------------------------------------------------------------------------------
// host plumbing
// ---------------------------------------------------------------------------------------------------------------

struct FrameDev; // clustered.hpp

struct vren_ctx {
    int device = 0;
    int num_cus = 256;
    char last_error[512] = {0};        // fixed buffer: reporting an error allocates nothing
    uint32_t* d_flags = nullptr;       // device capacity flags
    // grow-only device scratch for primitives that do not take a caller scratch buffer
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    // ring of device frame-descriptor slots for the single-frame stage entry points
    FrameDev* frame_slots = nullptr;
    uint32_t frame_slot_next = 0;
    // this device's copy of the fixed workspace of vren_exclusive_scan_u32_noalloc (scan.hip)
    void* noalloc_scan_work = nullptr;
    // radix-sort autotuner: forced configuration (-1 = none) and tuned choices keyed by problem bucket
    int radix_forced = -1;
    std::map<uint64_t, int> radix_tuned;
    // chain autotuner: tuned batch shapes keyed by size class (batch.hip)
    std::map<uint64_t, vren_chain_shape> chain_tuned;
};

#define VREN_FRAME_SLOTS 256

int vren_fail(vren_ctx* ctx, int status, const char* fmt, ...);
int vren_hip_fail(vren_ctx* ctx, hipError_t err, const char* what);
int vren_ensure_scratch(vren_ctx* ctx, size_t bytes, hipStream_t s);
int vren_noalloc_scan_setup(vren_ctx* ctx);

#define VREN_HIP(ctx, expr)                                                                                            \
    do {                                                                                                               \
        hipError_t _e = (expr);                                                                                        \
        if (_e != hipSuccess) return vren_hip_fail((ctx), _e, #expr);                                                  \
    } while (0)

#define VREN_LAUNCH_CHECK(ctx) VREN_HIP(ctx, hipGetLastError())

// Result of a teardown call (free, destroy, the synchronisation before them) made where nothing can report it: a
// destroy function returns nothing, and a sticky device error resurfaces at the next checked call.
static inline void vren_teardown(hipError_t) {}

static inline uint64_t ceil_div_u64(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
static inline uint32_t next_pow2_u32(uint32_t v)
{
    v--;
    v |= v >> 1;
    v |= v >> 2;
    v |= v >> 4;
    v |= v >> 8;
    v |= v >> 16;
    v++;
    return v;
}
static inline uint32_t ilog2_u64(uint64_t v)
{
    uint32_t r = 0;
    while (v > 1) {
        v >>= 1;
        r++;
    }
    return r;
}
