// Context, errors, device memory plumbing, host math helpers and the resident
// scene (AoS -> SoA upload) of libhv.so.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <vector>

#include "hv_internal.h"

namespace hv {

static thread_local char g_err[512] = "no error";

hv_status set_error(hv_status code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

hv_status hip_error(hipError_t e, const char* what, const char* file, int line) {
    const char* base = std::strrchr(file, '/');
    set_error(e == hipErrorOutOfMemory ? HV_ENOMEM : HV_EDEVICE, "%s failed: %s (%s:%d)", what,
              hipGetErrorString(e), base ? base + 1 : file, line);
    return e == hipErrorOutOfMemory ? HV_ENOMEM : HV_EDEVICE;
}

hv_status DevBuf::reserve(uint64_t need, bool keep_contents, hipStream_t stream) {
    if (need <= bytes) return HV_OK;
    uint64_t cap = bytes ? bytes : 256;
    while (cap < need) cap *= 2;
    void* p = nullptr;
    HV_HIP(hipMalloc(&p, cap));
    if (keep_contents && ptr && bytes) {
        hipError_t e = hipMemcpyAsync(p, ptr, bytes, hipMemcpyDeviceToDevice, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) {
            (void)hipFree(p);
            return hip_error(e, "grow copy", __FILE__, __LINE__);
        }
    }
    if (ptr) (void)hipFree(ptr);
    ptr = p;
    bytes = cap;
    return HV_OK;
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

}  // namespace hv

using namespace hv;

static hipStream_t pick(hv_ctx* ctx, void* stream) {
    return stream ? static_cast<hipStream_t>(stream) : ctx->stream;
}

// ---------------------------------------------------------------------------
// AoS -> SoA conversion of GpuInstanceData on the device (upload path).
// One thread per instance reads the 16 bytes of bounds and the 12 bytes of
// mesh_id/material_id/flags it needs out of the 144-byte record.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_scene_from_aos(const hv_instance_data* __restrict__ src, uint32_t count,
                                                        float4* __restrict__ bounds, uint32_t* __restrict__ meta) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= count) return;
    const float4 b = *reinterpret_cast<const float4*>(src[i].bounds);
    bounds[i] = b;
    meta[i] = (src[i].mesh_id & 0x7fffffffu) | ((src[i].flags & 1u) << 31);
}

// k_scene_from_aos for a device-resident source, also reducing the largest
// mesh id of the range (one atomic per workgroup) so the host can validate it.
__global__ __launch_bounds__(256) void k_scene_from_aos_max(const hv_instance_data* __restrict__ src, uint32_t count,
                                                            float4* __restrict__ bounds, uint32_t* __restrict__ meta,
                                                            uint32_t* __restrict__ max_id) {
    __shared__ uint32_t s_max[4];
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    uint32_t id = 0;
    if (i < count) {
        const float4 b = *reinterpret_cast<const float4*>(src[i].bounds);
        id = src[i].mesh_id;
        bounds[i] = b;
        meta[i] = (id & 0x7fffffffu) | ((src[i].flags & 1u) << 31);
    }
#pragma unroll
    for (uint32_t off = 32; off; off >>= 1) id = max(id, (uint32_t)__shfl_xor(id, off));
    if ((threadIdx.x & 63u) == 0) s_max[threadIdx.x >> 6] = id;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(max_id, max(max(s_max[0], s_max[1]), max(s_max[2], s_max[3])));
}

extern "C" {

const char* hv_last_error(void) { return g_err; }

hv_status hv_ctx_create(int32_t device, hv_ctx** out) {
    if (!out) return set_error(HV_EINVAL, "hv_ctx_create: out is NULL");
    int n = 0;
    HV_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return set_error(HV_EINVAL, "hv_ctx_create: device %d of %d", device, n);
    HV_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    HV_HIP(hipGetDeviceProperties(&prop, device));
    hv_ctx* c = new hv_ctx();
    c->device = device;
    c->cu_count = (uint32_t)prop.multiProcessorCount;
    c->hbm_bytes = (uint64_t)prop.totalGlobalMem;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    for (int i = 0; e == hipSuccess && i < 64; ++i) e = hipEventCreate(&c->timers[i]);
    if (e != hipSuccess) {
        hv_ctx_destroy(c);
        return hip_error(e, "hv_ctx_create", __FILE__, __LINE__);
    }
    // Once per device: the address of the no-allocation cull's scratch array
    // (a failure is reported by that entry point, which retries the lookup).
    (void)cull_scratch_init(device);
    *out = c;
    return HV_OK;
}

void hv_ctx_destroy(hv_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    for (auto& ev : ctx->timers)
        if (ev) (void)hipEventDestroy(ev);
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

hv_status hv_ctx_stream(hv_ctx* ctx, void** stream) {
    if (!ctx || !stream) return set_error(HV_EINVAL, "hv_ctx_stream: NULL argument");
    *stream = ctx->stream;
    return HV_OK;
}

hv_status hv_stream_create(hv_ctx* ctx, void** stream) {
    if (!ctx || !stream) return set_error(HV_EINVAL, "hv_stream_create: NULL argument");
    HV_HIP(hipSetDevice(ctx->device));
    hipStream_t s = nullptr;
    HV_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return HV_OK;
}

hv_status hv_stream_destroy(hv_ctx* ctx, void* stream) {
    if (!ctx || !stream) return set_error(HV_EINVAL, "hv_stream_destroy: NULL argument");
    HV_HIP(hipSetDevice(ctx->device));
    HV_HIP(hipStreamSynchronize((hipStream_t)stream));
    HV_HIP(hipStreamDestroy((hipStream_t)stream));
    return HV_OK;
}

hv_status hv_stream_synchronize(void* stream) {
    if (!stream) return set_error(HV_EINVAL, "hv_stream_synchronize: NULL stream");
    HV_HIP(hipStreamSynchronize((hipStream_t)stream));
    return HV_OK;
}

hv_status hv_ctx_synchronize(hv_ctx* ctx) {
    if (!ctx) return set_error(HV_EINVAL, "hv_ctx_synchronize: NULL ctx");
    HV_HIP(hipStreamSynchronize(ctx->stream));
    return HV_OK;
}

hv_status hv_ctx_device_info(hv_ctx* ctx, uint32_t* cu_count, uint64_t* hbm_bytes) {
    if (!ctx) return set_error(HV_EINVAL, "hv_ctx_device_info: NULL ctx");
    if (cu_count) *cu_count = ctx->cu_count;
    if (hbm_bytes) *hbm_bytes = ctx->hbm_bytes;
    return HV_OK;
}

hv_status hv_device_alloc(hv_ctx* ctx, uint64_t bytes, void** dptr) {
    if (!ctx || !dptr) return set_error(HV_EINVAL, "hv_device_alloc: NULL argument");
    HV_HIP(hipSetDevice(ctx->device));
    HV_HIP(hipMalloc(dptr, bytes ? bytes : 4));
    return HV_OK;
}

hv_status hv_device_free(hv_ctx* ctx, void* dptr) {
    if (!ctx) return set_error(HV_EINVAL, "hv_device_free: NULL ctx");
    if (dptr) HV_HIP(hipFree(dptr));
    return HV_OK;
}

hv_status hv_memcpy_h2d(hv_ctx* ctx, void* dst, const void* src, uint64_t bytes, void* stream) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return set_error(HV_EINVAL, "hv_memcpy_h2d: NULL argument");
    hipStream_t s = pick(ctx, stream);
    HV_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    HV_HIP(hipStreamSynchronize(s));
    return HV_OK;
}

hv_status hv_memcpy_d2h(hv_ctx* ctx, void* dst, const void* src, uint64_t bytes, void* stream) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return set_error(HV_EINVAL, "hv_memcpy_d2h: NULL argument");
    hipStream_t s = pick(ctx, stream);
    HV_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    HV_HIP(hipStreamSynchronize(s));
    return HV_OK;
}

hv_status hv_memcpy_d2d(hv_ctx* ctx, void* dst, const void* src, uint64_t bytes, void* stream) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return set_error(HV_EINVAL, "hv_memcpy_d2d: NULL argument");
    hipStream_t s = pick(ctx, stream);
    HV_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
    HV_HIP(hipStreamSynchronize(s));
    return HV_OK;
}

hv_status hv_memset_d32(hv_ctx* ctx, void* dst, uint32_t value, uint64_t count, void* stream) {
    if (!ctx || (!dst && count)) return set_error(HV_EINVAL, "hv_memset_d32: NULL argument");
    HV_HIP(hipMemsetD32Async(static_cast<hipDeviceptr_t>(dst), (int)value, count, pick(ctx, stream)));
    return HV_OK;
}

hv_status hv_timer_record(hv_ctx* ctx, uint32_t slot, void* stream) {
    if (!ctx || slot >= 64) return set_error(HV_EINVAL, "hv_timer_record: bad slot %u", slot);
    HV_HIP(hipEventRecord(ctx->timers[slot], pick(ctx, stream)));
    return HV_OK;
}

hv_status hv_timer_elapsed_ms(hv_ctx* ctx, uint32_t a, uint32_t b, float* ms) {
    if (!ctx || a >= 64 || b >= 64 || !ms) return set_error(HV_EINVAL, "hv_timer_elapsed_ms: bad argument");
    HV_HIP(hipEventSynchronize(ctx->timers[b]));
    HV_HIP(hipEventElapsedTime(ms, ctx->timers[a], ctx->timers[b]));
    return HV_OK;
}

// extract_frustum_planes: crates/helio-pass-indirect-dispatch/src/lib.rs:271-300
// (Rust f32 arithmetic; this file is built with -ffp-contract=off).
hv_status hv_extract_frustum_planes(const float vp[16], float planes[24]) {
    if (!vp || !planes) return set_error(HV_EINVAL, "hv_extract_frustum_planes: NULL argument");
    auto row = [&](int r, float out[4]) {
        for (int c = 0; c < 4; ++c) out[c] = vp[c * 4 + r];
    };
    float r0[4], r1[4], r2[4], r3[4];
    row(0, r0);
    row(1, r1);
    row(2, r2);
    row(3, r3);
    float p[6][4];
    for (int c = 0; c < 4; ++c) {
        p[0][c] = r3[c] + r0[c];  // left
        p[1][c] = r3[c] - r0[c];  // right
        p[2][c] = r3[c] + r1[c];  // bottom
        p[3][c] = r3[c] - r1[c];  // top
        p[4][c] = r2[c];          // near (z in [0, 1])
        p[5][c] = r3[c] - r2[c];  // far
    }
    for (int i = 0; i < 6; ++i) {
        const float len = std::sqrt(p[i][0] * p[i][0] + p[i][1] * p[i][1] + p[i][2] * p[i][2]);
        for (int c = 0; c < 4; ++c) planes[4 * i + c] = len > 1e-10f ? p[i][c] / len : p[i][c];
    }
    return HV_OK;
}

// mip_levels: crates/helio-pass-hiz/src/lib.rs:376-379.
uint32_t hv_mip_levels(uint32_t width, uint32_t height) {
    const uint32_t m = width > height ? width : height;
    return m ? 32u - (uint32_t)__builtin_clz(m) : 1u;
}

// InstanceCullData::from_instance: crates/helio-pass-virtual-geometry/src/lib.rs:131-168.
hv_status hv_instance_cull_data_from_instances(const hv_instance_data* instances, uint32_t count,
                                               hv_instance_cull_data* out) {
    if ((!instances || !out) && count) return set_error(HV_EINVAL, "hv_instance_cull_data_from_instances: NULL");
    for (uint32_t i = 0; i < count; ++i) {
        const float* m = instances[i].model;
        const float sx = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        const float sy = std::sqrt(m[4] * m[4] + m[5] * m[5] + m[6] * m[6]);
        const float sz = std::sqrt(m[8] * m[8] + m[9] * m[9] + m[10] * m[10]);
        const float mx = std::fmax(std::fmax(sx, sy), sz);  // f32::max ignores NaN
        const float mn = std::fmin(std::fmin(sx, sy), sz);
        const float dxy = m[0] * m[4] + m[1] * m[5] + m[2] * m[6];
        const float dxz = m[0] * m[8] + m[1] * m[9] + m[2] * m[10];
        const float dyz = m[4] * m[8] + m[5] * m[9] + m[6] * m[10];
        const float det = m[0] * (m[5] * m[10] - m[6] * m[9]) - m[4] * (m[1] * m[10] - m[2] * m[9]) +
                          m[8] * (m[1] * m[6] - m[2] * m[5]);
        const bool affine = std::fabs(m[3]) <= 1.0e-6f && std::fabs(m[7]) <= 1.0e-6f &&
                            std::fabs(m[11]) <= 1.0e-6f && std::fabs(m[15] - 1.0f) <= 1.0e-4f;
        bool finite = true;
        for (int k = 0; k < 16; ++k) finite = finite && std::isfinite(m[k]);
        const bool valid = finite && std::isfinite(mx) && mn > 1.0e-8f && affine;
        const float uniform_tol = mx * 1.0e-4f;
        const float ortho_tol = mx * mx * 1.0e-4f;
        const bool angle_preserving = mx - mn <= uniform_tol && std::fabs(dxy) <= ortho_tol &&
                                      std::fabs(dxz) <= ortho_tol && std::fabs(dyz) <= ortho_tol;
        out[i].max_scale = valid ? mx : 0.0f;
        out[i].min_scale = valid ? mn : 0.0f;
        out[i].cone_cull_enabled = (valid && angle_preserving && det > 0.0f) ? 1u : 0u;
        out[i].valid_transform = valid ? 1u : 0u;
    }
    return HV_OK;
}

// ---------------------------------------------------------------------------
// Resident scene.
// ---------------------------------------------------------------------------
extern "C++" hv_status scene_grow(hv_scene* s, uint32_t need, hipStream_t stream) {
    if (need <= s->capacity) return HV_OK;
    uint64_t cap = s->capacity ? s->capacity : 1024;
    while (cap < need) cap *= 2;
    if (cap > 0xffffffffull) cap = 0xffffffffull;
    HV_TRY(s->bounds.reserve(cap * 16, true, stream));
    HV_TRY(s->meta.reserve(cap * 4, true, stream));
    if (s->explicit_aabb) HV_TRY(s->aabb.reserve(cap * 32, true, stream));
    s->capacity = (uint32_t)cap;
    ++s->storage_gen;
    return HV_OK;
}

hv_status hv_scene_create(hv_ctx* ctx, uint32_t capacity, hv_scene** out) {
    if (!ctx || !out) return set_error(HV_EINVAL, "hv_scene_create: NULL argument");
    HV_HIP(hipSetDevice(ctx->device));
    hv_scene* s = new hv_scene();
    s->ctx = ctx;
    hv_status st = scene_grow(s, capacity ? capacity : 1024, ctx->stream);
    if (st != HV_OK) {
        hv_scene_destroy(s);
        return st;
    }
    *out = s;
    return HV_OK;
}

void hv_scene_destroy(hv_scene* s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    s->bounds.release();
    s->meta.release();
    s->aabb.release();
    s->staging.release();
    s->mesh_lods.release();
    delete s;
}

hv_status hv_scene_upload(hv_scene* s, const hv_instance_data* instances, const hv_instance_aabb* aabbs,
                          uint32_t first, uint32_t count, void* stream) {
    if (!s || (!instances && count)) return set_error(HV_EINVAL, "hv_scene_upload: NULL argument");
    if ((uint64_t)first + count > 0xffffffffull) return set_error(HV_ELIMIT, "hv_scene_upload: range overflows u32");
    uint32_t max_id = 0;
    for (uint32_t i = 0; i < count; ++i) {
        if (instances[i].mesh_id >= 0x80000000u)
            return set_error(HV_EINVAL, "hv_scene_upload: mesh_id %u of instance %u is >= 2^31", instances[i].mesh_id,
                             first + i);
        max_id = std::max(max_id, instances[i].mesh_id);
    }
    hipStream_t st = pick(s->ctx, stream);
    HV_HIP(hipSetDevice(s->ctx->device));
    // An AABB array that differs from sphere_to_aabb(bounds) anywhere switches
    // the scene to explicit AABBs (the reference reads aabbs[] separately).
    bool need_explicit = false;
    if (aabbs) {
        for (uint32_t i = 0; i < count && !need_explicit; ++i) {
            const float* b = instances[i].bounds;
            const hv_instance_aabb& a = aabbs[i];
            for (int k = 0; k < 3; ++k)
                if (a.min[k] != b[k] - b[3] || a.max[k] != b[k] + b[3]) need_explicit = true;
        }
    }
    HV_TRY(scene_grow(s, first + count, st));
    if (need_explicit && !s->explicit_aabb) {
        // Materialise sphere_to_aabb for everything already resident.
        std::vector<hv_instance_aabb> host((size_t)s->capacity);
        std::vector<float> b4((size_t)s->count * 4);
        if (s->count) HV_HIP(hipMemcpy(b4.data(), s->bounds.ptr, (size_t)s->count * 16, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < s->count; ++i) {
            const float* b = &b4[4 * (size_t)i];
            for (int k = 0; k < 3; ++k) {
                host[i].min[k] = b[k] - b[3];
                host[i].max[k] = b[k] + b[3];
            }
            host[i]._pad0 = host[i]._pad1 = 0.0f;
        }
        HV_TRY(s->aabb.reserve((uint64_t)s->capacity * 32, false, st));
        HV_HIP(hipMemcpy(s->aabb.ptr, host.data(), (size_t)s->capacity * 32, hipMemcpyHostToDevice));
        s->explicit_aabb = true;
        ++s->storage_gen;
    }
    if (count) {
        HV_TRY(s->staging.reserve((uint64_t)count * sizeof(hv_instance_data), false, st));
        HV_HIP(hipMemcpyAsync(s->staging.ptr, instances, (size_t)count * sizeof(hv_instance_data),
                              hipMemcpyHostToDevice, st));
        k_scene_from_aos<<<ceil_div(count, 256), 256, 0, st>>>(s->staging.as<hv_instance_data>(), count,
                                                             s->bounds.as<float4>() + first,
                                                             s->meta.as<uint32_t>() + first);
        HV_HIP(hipGetLastError());
        if (s->explicit_aabb) {
            if (aabbs) {
                HV_HIP(hipMemcpyAsync(s->aabb.as<hv_instance_aabb>() + first, aabbs, (size_t)count * 32,
                                      hipMemcpyHostToDevice, st));
            } else {
                std::vector<hv_instance_aabb> host(count);
                for (uint32_t i = 0; i < count; ++i) {
                    const float* b = instances[i].bounds;
                    for (int k = 0; k < 3; ++k) {
                        host[i].min[k] = b[k] - b[3];
                        host[i].max[k] = b[k] + b[3];
                    }
                    host[i]._pad0 = host[i]._pad1 = 0.0f;
                }
                HV_HIP(hipMemcpyAsync(s->aabb.as<hv_instance_aabb>() + first, host.data(), (size_t)count * 32,
                                      hipMemcpyHostToDevice, st));
                HV_HIP(hipStreamSynchronize(st));
            }
        }
        HV_HIP(hipStreamSynchronize(st));
    }
    if (first + count > s->count) s->count = first + count;
    s->max_mesh_id = std::max(s->max_mesh_id, max_id);
    return HV_OK;
}

hv_status hv_scene_upload_soa(hv_scene* s, const float* bounds4, const uint32_t* mesh_ids, const uint32_t* flags,
                              uint32_t first, uint32_t count, void* stream) {
    if (!s || ((!bounds4 || !mesh_ids) && count)) return set_error(HV_EINVAL, "hv_scene_upload_soa: NULL argument");
    if ((uint64_t)first + count > 0xffffffffull)
        return set_error(HV_ELIMIT, "hv_scene_upload_soa: range overflows u32");
    if (s->explicit_aabb)
        return set_error(HV_ESTATE, "hv_scene_upload_soa: scene holds explicit AABBs; use hv_scene_upload");
    hipStream_t st = pick(s->ctx, stream);
    HV_HIP(hipSetDevice(s->ctx->device));
    HV_TRY(scene_grow(s, first + count, st));
    std::vector<uint32_t> meta(count);
    for (uint32_t i = 0; i < count; ++i) {
        if (mesh_ids[i] >= 0x80000000u)
            return set_error(HV_EINVAL, "hv_scene_upload_soa: mesh_id %u of instance %u is >= 2^31", mesh_ids[i],
                             first + i);
        meta[i] = mesh_ids[i] | ((flags ? (flags[i] & 1u) : 0u) << 31);
        s->max_mesh_id = std::max(s->max_mesh_id, mesh_ids[i]);
    }
    if (count) {
        HV_HIP(hipMemcpyAsync(s->bounds.as<float4>() + first, bounds4, (size_t)count * 16, hipMemcpyHostToDevice,
                              st));
        HV_HIP(hipMemcpyAsync(s->meta.as<uint32_t>() + first, meta.data(), (size_t)count * 4, hipMemcpyHostToDevice,
                              st));
        HV_HIP(hipStreamSynchronize(st));
    }
    if (first + count > s->count) s->count = first + count;
    return HV_OK;
}

hv_status hv_scene_upload_device(hv_scene* s, const hv_instance_data* instances, uint32_t first, uint32_t count,
                                 void* stream) {
    if (!s || (!instances && count)) return set_error(HV_EINVAL, "hv_scene_upload_device: NULL argument");
    if ((uint64_t)first + count > 0xffffffffull)
        return set_error(HV_ELIMIT, "hv_scene_upload_device: range overflows u32");
    if (s->explicit_aabb)
        return set_error(HV_ESTATE, "hv_scene_upload_device: scene holds explicit AABBs; use hv_scene_upload");
    if (!count) return HV_OK;
    hipStream_t st = pick(s->ctx, stream);
    HV_HIP(hipSetDevice(s->ctx->device));
    HV_TRY(scene_grow(s, first + count, st));
    HV_TRY(s->staging.reserve(sizeof(uint32_t), false, st));
    uint32_t* d_max = s->staging.as<uint32_t>();
    HV_HIP(hipMemsetAsync(d_max, 0, 4, st));
    k_scene_from_aos_max<<<ceil_div(count, 256), 256, 0, st>>>(instances, count, s->bounds.as<float4>() + first,
                                                             s->meta.as<uint32_t>() + first, d_max);
    HV_HIP(hipGetLastError());
    uint32_t max_id = 0;
    HV_HIP(hipMemcpyAsync(&max_id, d_max, 4, hipMemcpyDeviceToHost, st));
    HV_HIP(hipStreamSynchronize(st));
    if (max_id >= 0x80000000u)
        return set_error(HV_EINVAL, "hv_scene_upload_device: a mesh_id of [%u, %u) is >= 2^31 (range written "
                                    "with the id masked; the instance count is unchanged)", first, first + count);
    if (first + count > s->count) s->count = first + count;
    s->max_mesh_id = std::max(s->max_mesh_id, max_id);
    return HV_OK;
}

hv_status hv_scene_swap_remove(hv_scene* s, uint32_t index, void* stream) {
    if (!s) return set_error(HV_EINVAL, "hv_scene_swap_remove: NULL scene");
    if (index >= s->count) return set_error(HV_EINVAL, "hv_scene_swap_remove: index %u >= %u instances", index, s->count);
    const uint32_t last = s->count - 1u;
    if (index < last) {
        hipStream_t st = pick(s->ctx, stream);
        HV_HIP(hipSetDevice(s->ctx->device));
        HV_HIP(hipMemcpyAsync(s->bounds.as<float4>() + index, s->bounds.as<float4>() + last, 16,
                              hipMemcpyDeviceToDevice, st));
        HV_HIP(hipMemcpyAsync(s->meta.as<uint32_t>() + index, s->meta.as<uint32_t>() + last, 4,
                              hipMemcpyDeviceToDevice, st));
        if (s->explicit_aabb)
            HV_HIP(hipMemcpyAsync(s->aabb.as<hv_instance_aabb>() + index, s->aabb.as<hv_instance_aabb>() + last, 32,
                                  hipMemcpyDeviceToDevice, st));
    }
    s->count = last;
    return HV_OK;
}

hv_status hv_scene_truncate(hv_scene* s, uint32_t new_len) {
    if (!s) return set_error(HV_EINVAL, "hv_scene_truncate: NULL scene");
    if (new_len < s->count) s->count = new_len;
    return HV_OK;
}

hv_status hv_scene_read_soa(hv_scene* s, float* bounds4, uint32_t* meta, uint32_t first, uint32_t count) {
    if (!s) return set_error(HV_EINVAL, "hv_scene_read_soa: NULL scene");
    if ((uint64_t)first + count > s->count)
        return set_error(HV_EINVAL, "hv_scene_read_soa: [%u, %u) exceeds %u instances", first, first + count, s->count);
    HV_HIP(hipSetDevice(s->ctx->device));
    HV_HIP(hipDeviceSynchronize());  // uploads may have gone to any caller stream
    if (bounds4 && count)
        HV_HIP(hipMemcpy(bounds4, s->bounds.as<float4>() + first, (size_t)count * 16, hipMemcpyDeviceToHost));
    if (meta && count) HV_HIP(hipMemcpy(meta, s->meta.as<uint32_t>() + first, (size_t)count * 4, hipMemcpyDeviceToHost));
    return HV_OK;
}

hv_status hv_scene_set_mesh_lods(hv_scene* s, const hv_mesh_lod* table, uint32_t mesh_count, uint32_t lod_count) {
    if (!s || !table || !mesh_count) return set_error(HV_EINVAL, "hv_scene_set_mesh_lods: empty table");
    if (lod_count < 1 || lod_count > HV_MAX_LOD_LEVELS)
        return set_error(HV_EINVAL, "hv_scene_set_mesh_lods: lod_count %u not in 1..8", lod_count);
    const uint64_t n = (uint64_t)mesh_count * lod_count;
    if (n > 14336) return set_error(HV_ELIMIT, "hv_scene_set_mesh_lods: %llu buckets > 14336", (unsigned long long)n);
    HV_HIP(hipSetDevice(s->ctx->device));
    HV_TRY(s->mesh_lods.reserve(n * sizeof(hv_mesh_lod), false, s->ctx->stream));
    HV_HIP(hipMemcpy(s->mesh_lods.ptr, table, n * sizeof(hv_mesh_lod), hipMemcpyHostToDevice));
    ++s->storage_gen;
    s->mesh_count = mesh_count;
    s->lod_count = lod_count;
    return HV_OK;
}

hv_status hv_scene_info(const hv_scene* s, uint32_t* instance_count, uint32_t* mesh_count, uint32_t* lod_count,
                        uint64_t* device_bytes) {
    if (!s) return set_error(HV_EINVAL, "hv_scene_info: NULL scene");
    if (instance_count) *instance_count = s->count;
    if (mesh_count) *mesh_count = s->mesh_count;
    if (lod_count) *lod_count = s->lod_count;
    if (device_bytes)
        *device_bytes = s->bounds.bytes + s->meta.bytes + s->aabb.bytes + s->staging.bytes + s->mesh_lods.bytes;
    return HV_OK;
}

}  // extern "C"
