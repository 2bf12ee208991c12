// sr_raster_common.hpp — device helpers shared by the raster kernels
// (sr_raster.hip, sr_raster_fmt.hip): the reference's screen tile grid,
// uniform loads through the scalar cache, and the triangle record as the
// per-triangle loop reads it.
#pragma once

#include "sr_device.hpp"

namespace sr {

// Screen tile grid of the reference (fragment.rs:188-216): column j spans
// [j*tw, min((j+1)*tw, W-1)], rows likewise; both ends inclusive.
__device__ __forceinline__ void sem_range(uint32_t lo_px, uint32_t hi_px, uint32_t t, uint32_t n,
                                          uint32_t& jlo, uint32_t& jhi) {
    // tiles whose inclusive span meets [lo_px, hi_px]
    const uint32_t a = (lo_px + t - 1) / t;  // ceil(lo/t)
    jlo = a > 0 ? a - 1 : 0;
    jhi = hi_px / t;
    if (jhi > n - 1) jhi = n - 1;
}

// Loads through the constant address space: read-only for the kernel's
// lifetime, so a wave-uniform address becomes an s_load into SGPRs.
__device__ __forceinline__ uint32_t load_item(const uint32_t* base, uint32_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
    return ((const __attribute__((address_space(4))) uint32_t*)base)[i];
#else
    return base[i];
#endif
}

// 64 consecutive item ids, lane k holding the id at position c0 + k, the
// position clamped into [beg, end) (beg < end) so every lane loads.
__device__ __forceinline__ uint32_t load_ids(const uint32_t* items, uint32_t c0, uint32_t beg,
                                             uint32_t end, int lane) {
    const uint32_t c = min(max(c0 + (uint32_t)lane, beg), end - 1);
#if defined(__HIP_DEVICE_COMPILE__)
    return ((const __attribute__((address_space(1))) uint32_t*)items)[c];
#else
    return items[c];
#endif
}

// The fields of a triangle record the raster loop reads.
struct RecLane {
    double R;
    uint32_t zs;  // bit 0: strip flag; the rest: depth bound key (TriRec::zs)
    float x3, y3, e23y, e23x, e31y, e31x, z1, z2, z3;
    int X0, X1, Y0, Y1;
};

__device__ __forceinline__ uint32_t lane_u32(uint32_t v, int k) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, k);
}
__device__ __forceinline__ float lane_f32(uint32_t v, int k) {
    return __builtin_bit_cast(float, lane_u32(v, k));
}

// Record `id` through the scalar cache: `id` is wave-uniform, so these
// become scalar loads into SGPRs.
__device__ __forceinline__ RecLane rec_s(const TriRec* tris, uint32_t id) {
#if defined(__HIP_DEVICE_COMPILE__)
    typedef const __attribute__((address_space(4))) TriRec CRec;
    const CRec* p = (const CRec*)tris + id;
    RecLane r;
    r.R = p->R;
    r.zs = p->zs;
    r.x3 = p->x3; r.y3 = p->y3;
    r.e23y = p->e23y; r.e23x = p->e23x; r.e31y = p->e31y; r.e31x = p->e31x;
    r.z1 = p->z1; r.z2 = p->z2; r.z3 = p->z3;
    r.X0 = p->X0; r.X1 = p->X1; r.Y0 = p->Y0; r.Y1 = p->Y1;
    return r;
#else
    const TriRec& t = tris[id];
    return RecLane{t.R, t.zs, t.x3, t.y3, t.e23y, t.e23x, t.e31y, t.e31x, t.z1, t.z2, t.z3, t.X0, t.X1, t.Y0, t.Y1};
#endif
}

// The box of record `id` into rc (scalar loads).
__device__ __forceinline__ void rec_box(const TriRec* tris, uint32_t id, RecLane& rc) {
#if defined(__HIP_DEVICE_COMPILE__)
    typedef const __attribute__((address_space(4))) TriRec CRec;
    const CRec* p = (const CRec*)tris + id;
    rc.X0 = p->X0; rc.X1 = p->X1; rc.Y0 = p->Y0; rc.Y1 = p->Y1;
#else
    const TriRec& t = tris[id];
    rc.X0 = t.X0; rc.X1 = t.X1; rc.Y0 = t.Y0; rc.Y1 = t.Y1;
#endif
}

// Lane k copies bytes 0-47 of record ids[k] into stage chunks 0-2 (stage[c
// * 64 + k]) by three 16-byte LDS DMA loads (no registers in between).
__device__ __forceinline__ void stage_records(const TriRec* tris, uint32_t id, uint4* stage) {
#if defined(__HIP_DEVICE_COMPILE__)
    typedef __attribute__((address_space(3))) void LVoid;
    const char* src = (const char*)(tris + id);
#pragma unroll
    for (int c = 0; c < 3; c++)
        __builtin_amdgcn_global_load_lds((const void*)(src + 16 * c), (LVoid*)(stage + 64 * c), 16, 0, 0);
#else
    for (int c = 0; c < 3; c++) stage[64 * c] = ((const uint4*)(tris + id))[c];
#endif
}

// s_waitcnt vmcnt(0) (expcnt and lgkmcnt left alone): every vector memory
// operation of the wave, LDS DMA loads included, has completed.
__device__ __forceinline__ void wait_vm_all() {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt [3:0] = 0, expcnt [6:4] = 7, lgkmcnt [11:8] = 15, vmcnt [15:14] = 0
#endif
}

// Record k of the stage (the same for every lane: broadcast reads); the
// box is left unset.  zs goes to an SGPR: the raster branches on it.
__device__ __forceinline__ RecLane rec_stage(const uint4* stage, int k) {
    const uint4 a = stage[k], b = stage[64 + k], c = stage[128 + k];
    RecLane r;
    r.R = __builtin_bit_cast(double, ((uint64_t)a.y << 32) | a.x);
    r.zs = (uint32_t)__builtin_amdgcn_readfirstlane((int)a.z);
    r.x3 = __uint_as_float(a.w);
    r.y3 = __uint_as_float(b.x); r.e23y = __uint_as_float(b.y); r.e23x = __uint_as_float(b.z);
    r.e31y = __uint_as_float(b.w);
    r.e31x = __uint_as_float(c.x); r.z1 = __uint_as_float(c.y); r.z2 = __uint_as_float(c.z);
    r.z3 = __uint_as_float(c.w);
    r.X0 = r.X1 = r.Y0 = r.Y1 = 0;
    return r;
}

}  // namespace sr
