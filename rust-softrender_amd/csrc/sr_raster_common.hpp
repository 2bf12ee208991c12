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

// The fields of a triangle record the raster loop reads, in SGPRs.
struct RecLane {
    double R;
    float x3, y3, e23y, e23x, e31y, e31x, z1, z2, z3;
    int X0, X1, Y0, Y1;
    uint32_t zs;  // bit 0: strip flag; the rest: depth bound key (TriRec::zs)
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
    r.x3 = p->x3; r.y3 = p->y3;
    r.e23y = p->e23y; r.e23x = p->e23x; r.e31y = p->e31y; r.e31x = p->e31x;
    r.z1 = p->z1; r.z2 = p->z2; r.z3 = p->z3;
    r.X0 = p->X0; r.X1 = p->X1; r.Y0 = p->Y0; r.Y1 = p->Y1;
    r.zs = p->zs;
    return r;
#else
    const TriRec& t = tris[id];
    return RecLane{t.R, t.x3, t.y3, t.e23y, t.e23x, t.e31y, t.e31x, t.z1, t.z2, t.z3, t.X0, t.X1, t.Y0, t.Y1,
                   t.zs};
#endif
}

// Orders the scalar load of the next record after the arrival of the
// current one: `next` passes through an empty asm that also reads a field
// of `cur`, so its address is not known before that field is.
__device__ __forceinline__ void after_record(uint32_t& next, const RecLane& cur) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+s"(next) : "s"(cur.X0));
#endif
}

// One vector load per lane from record `id` (a record is one 64-byte
// line): brings it into L2 ahead of the scalar load that reads it.
__device__ __forceinline__ uint32_t touch_rec(const TriRec* tris, uint32_t id) {
#if defined(__HIP_DEVICE_COMPILE__)
    return *(const __attribute__((address_space(1))) uint32_t*)(tris + id);
#else
    return *(const uint32_t*)(tris + id);
#endif
}

// Keeps a loaded value alive (its load is then neither removed nor left
// unwaited): an empty asm that reads it.
__device__ __forceinline__ void keep_live(uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" ::"v"(v));
#endif
}


}  // namespace sr
