// hip_lanes.h -- runs the library's device code (pm_device.h, pm_invert.h)
// on the host, one std::thread per GPU lane, so that a kernel's logic can be
// checked under AddressSanitizer / UBSan on the CPU.  TEST INFRASTRUCTURE:
// only tests/lanes/lanes_run.cpp includes it (through hip/hip_runtime.h of
// this directory, which shadows the HIP header).
//
// A pixel group of W lanes is a set of W threads with a barrier of its own.
// Every cross-lane operation the device code uses (DPP row moves, the
// ds_swizzle xor-16, readlane, __shfl / __shfl_xor, __ballot) is an exchange
// through the group's slots between two barriers; the partner lane must lie
// in the same group, which is asserted.  wave_sync's wave barrier is the
// group barrier.  One workgroup runs at a time (its __shared__ array is a
// function-local static).  Rounding differs from the GPU where the device
// code uses hardware approximations (rcp): the harness checks memory safety
// and logic, not bits against the GPU.
#pragma once

#include <barrier>
#include <cassert>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <math.h>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline __attribute__((always_inline))
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) __attribute__((aligned(n)))

struct lanes_dim3 {
    unsigned x = 0, y = 0, z = 0;
};
inline thread_local lanes_dim3 threadIdx, blockIdx;
inline lanes_dim3 blockDim, gridDim;

// the group of the calling thread
struct LanesGroup {
    int W = 0;
    std::barrier<> *bar = nullptr;
    uint64_t slot[128];
};
inline thread_local LanesGroup *lanes_grp = nullptr;

inline int __lane_id() { return (int)(threadIdx.x % 64); }

// the lane's index in its group (a group of 128 lanes spans two waves)
inline int lanes_in_group() { return (int)(threadIdx.x % lanes_grp->W); }

inline void lanes_barrier() { lanes_grp->bar->arrive_and_wait(); }

// the workgroup barrier; the harness runs 128-lane groups as one workgroup
// each, the only use the device code makes of it
inline void __syncthreads() { lanes_barrier(); }

// value of lane src_wave_lane (a lane of the same 64-lane wave) after all
// lanes of the group have published theirs
template <typename V>
inline V lanes_xchg(V v, int src_wave_lane)
{
    static_assert(sizeof(V) <= 8, "exchange width");
    LanesGroup *g = lanes_grp;
    const int base = __lane_id() - lanes_in_group();
    const int src = src_wave_lane - base;
    if (src < 0 || src >= g->W) {
        fprintf(stderr, "lanes: lane %d reads lane %d outside its group of %d\n", __lane_id(),
                src_wave_lane, g->W);
        abort();
    }
    uint64_t u = 0;
    memcpy(&u, &v, sizeof(V));
    g->slot[lanes_in_group()] = u;
    lanes_barrier();
    const uint64_t r = g->slot[src];
    lanes_barrier();
    V out;
    memcpy(&out, &r, sizeof(V));
    return out;
}

inline int __builtin_amdgcn_mov_dpp(int v, int ctrl, int row_mask, int bank_mask, bool bound)
{
    (void)row_mask; (void)bank_mask; (void)bound;
    const int l = __lane_id();
    int src;
    if (ctrl <= 0xFF) {                                  // quad_perm
        const int sel = (ctrl >> (2 * (l & 3))) & 3;
        src = (l & ~3) | sel;
    } else if (ctrl == 0x141) {                          // row_half_mirror
        src = (l & ~7) | (7 - (l & 7));
    } else if (ctrl == 0x140) {                          // row_mirror
        src = (l & ~15) | (15 - (l & 15));
    } else {
        fprintf(stderr, "lanes: DPP control 0x%x not emulated\n", ctrl);
        abort();
    }
    return lanes_xchg(v, src);
}

inline int __builtin_amdgcn_ds_swizzle(int v, int pattern)
{
    if (pattern != 0x401F) {
        fprintf(stderr, "lanes: swizzle 0x%x not emulated\n", pattern);
        abort();
    }
    return lanes_xchg(v, __lane_id() ^ 16);
}

inline int __builtin_amdgcn_readlane(int v, int lane) { return lanes_xchg(v, lane); }

template <typename V>
inline V __shfl(V v, int src, int width = 64)
{
    return lanes_xchg(v, (__lane_id() & ~(width - 1)) + (src & (width - 1)));
}

template <typename V>
inline V __shfl_xor(V v, int mask, int width = 64)
{
    return lanes_xchg(v, (__lane_id() & ~(width - 1)) + ((__lane_id() ^ mask) & (width - 1)));
}

inline unsigned long long __ballot(int pred)
{
    LanesGroup *g = lanes_grp;
    const int base = __lane_id() - lanes_in_group();
    g->slot[lanes_in_group()] = pred ? 1 : 0;
    lanes_barrier();
    unsigned long long m = 0;
    for (int i = 0; i < g->W; i++)
        if (g->slot[i]) m |= 1ull << (base + i);
    lanes_barrier();
    return m;
}

inline void __builtin_amdgcn_fence(int order, const char *scope)
{
    (void)order; (void)scope;
}
inline void __builtin_amdgcn_wave_barrier() { lanes_barrier(); }

inline double __builtin_amdgcn_rcp(double x) { return 1.0 / x; }
inline float __builtin_amdgcn_rcpf(float x) { return 1.0f / x; }

inline int atomicAdd(int *p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline unsigned long long atomicAdd(unsigned long long *p, unsigned long long v)
{
    return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}

inline int __double2loint(double d)
{
    uint64_t u;
    memcpy(&u, &d, 8);
    return (int)(uint32_t)u;
}
inline int __double2hiint(double d)
{
    uint64_t u;
    memcpy(&u, &d, 8);
    return (int)(uint32_t)(u >> 32);
}
inline double __hiloint2double(int hi, int lo)
{
    const uint64_t u = ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
    double d;
    memcpy(&d, &u, 8);
    return d;
}
inline int __float_as_int(float f)
{
    int i;
    memcpy(&i, &f, 4);
    return i;
}
inline float __int_as_float(int i)
{
    float f;
    memcpy(&f, &i, 4);
    return f;
}
inline float __fadd_rn(float a, float b) { return a + b; }
inline float __fsub_rn(float a, float b) { return a - b; }
inline float __fmul_rn(float a, float b) { return a * b; }

inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
