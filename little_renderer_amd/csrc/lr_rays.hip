// lr_rays.hip - ray batches for the BVH-stress workload: pinhole primaries (camera_ray,
// engine/scene/camera.h:38-42, at each pixel's centre), then per primary hit one
// cosine-weighted diffuse bounce (rng.h:53-70: Duff's basis, Malley's method) and one shadow
// ray toward a point light, both appended to queues compacted per wave (one atomic per wave:
// ballot + popcount of the lanes below), with the source pixel kept beside each entry.
#include "lr_device.h"
#include "lr_internal.h"

namespace {

constexpr int RBLOCK = 256;

__device__ __forceinline__ uint32_t hash32(uint32_t x)   // lowbias32 (util.wgsl:6-10)
{
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__global__ __launch_bounds__(RBLOCK) void k_rays_primary(float ox, float oy, float oz, float ulx, float uly, float ulz,
                                                          float hx, float hy, float hz, float vx, float vy, float vz, int w, int h,
                                                          float tmin, float tmax, float4 *rays)
{
    const int i = blockIdx.x * RBLOCK + threadIdx.x;
    if (i >= w * h) return;
    const int x = i % w, y = i / w;
    const float s = ((float)x + 0.5f) / (float)w, t = ((float)y + 0.5f) / (float)h;
    // target = upper_left + horizontal * s - vertical * t; d = normalize(target - origin)
    const float tx = (ulx + hx * s) - vx * t, ty = (uly + hy * s) - vy * t, tz = (ulz + hz * s) - vz * t;
    const float dx = tx - ox, dy = ty - oy, dz = tz - oz;
    const float l = sqrtf(dx * dx + dy * dy + dz * dz);
    const float il = l > 0.0f ? 1.0f / l : 1.0f;
    rays[2 * (size_t)i] = make_float4(ox, oy, oz, tmin);
    rays[2 * (size_t)i + 1] = make_float4(dx * il, dy * il, dz * il, tmax);
}

__global__ void k_clear(uint32_t *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ __launch_bounds__(RBLOCK) void k_rays_secondary(const float4 *__restrict__ rays, const float4 *__restrict__ hits, int n,
                                                            const float4 *__restrict__ tris, float lx, float ly, float lz, float eps,
                                                            uint32_t seed, float4 *bounce, float4 *shadow, uint32_t *pix,
                                                            uint32_t *count)
{
    const int i = blockIdx.x * RBLOCK + threadIdx.x;
    bool hit = false;
    float4 hv = make_float4(0, 0, 0, 0);
    if (i < n) {
        hv = hits[i];
        hit = __float_as_uint(hv.y) != LR_NONE;
    }
    const uint64_t mask = __ballot(hit);
    if (mask == 0) return;
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if ((int)lane_id() == leader) base = atomicAdd(count, (uint32_t)__popcll(mask));
    base = __shfl(base, leader, 64);
    if (!hit) return;
    const uint32_t slot = base + (uint32_t)__popcll(mask & lanemask_lt());
    const float4 r0 = rays[2 * (size_t)i], r1 = rays[2 * (size_t)i + 1];
    const float t = hv.x;
    const float px = r0.x + r1.x * t, py = r0.y + r1.y * t, pz = r0.z + r1.z * t;
    const uint32_t id = __float_as_uint(hv.y);
    const float4 tu = tris[3 * (size_t)id + 1], tv = tris[3 * (size_t)id + 2];
    float nx = tu.y * tv.z - tu.z * tv.y, ny = tu.z * tv.x - tu.x * tv.z, nz = tu.x * tv.y - tu.y * tv.x;
    float nl = sqrtf(nx * nx + ny * ny + nz * nz);
    nl = nl > 0.0f ? 1.0f / nl : 0.0f;
    nx *= nl; ny *= nl; nz *= nl;
    if (nx * r1.x + ny * r1.y + nz * r1.z > 0.0f) { nx = -nx; ny = -ny; nz = -nz; }
    // two numbers from the pixel and the seed
    const uint32_t h1 = hash32((uint32_t)i ^ hash32(seed)), h2 = hash32(h1 ^ 0x9e3779b9u);
    const float u1 = (float)(h1 >> 8) * (1.0f / 16777216.0f), u2 = (float)(h2 >> 8) * (1.0f / 16777216.0f);
    // sample_cosine_hemisphere_u (rng.h:63-70) with onb_from_normal (rng.h:53-59)
    float sp, cp;
    sincosf(2.0f * 3.14159265358979f * u1, &sp, &cp);
    const float rad = sqrtf(u2), bx = rad * cp, by = rad * sp, bz = sqrtf(fmaxf(0.0f, 1.0f - u2));
    const float sg = copysignf(1.0f, nz), a = -1.0f / (sg + nz), c = nx * ny * a;
    const float t1x = 1.0f + sg * nx * nx * a, t1y = sg * c, t1z = -sg * nx;
    const float b1x = c, b1y = sg + ny * ny * a, b1z = -ny;
    const float dx = t1x * bx + b1x * by + nx * bz, dy = t1y * bx + b1y * by + ny * bz, dz = t1z * bx + b1z * by + nz * bz;
    bounce[2 * (size_t)slot] = make_float4(px, py, pz, eps);
    bounce[2 * (size_t)slot + 1] = make_float4(dx, dy, dz, 1e30f);
    const float sx = lx - px, sy = ly - py, sz = lz - pz;
    const float dist = sqrtf(sx * sx + sy * sy + sz * sz);
    const float id_ = dist > 0.0f ? 1.0f / dist : 0.0f;
    shadow[2 * (size_t)slot] = make_float4(px, py, pz, eps);
    shadow[2 * (size_t)slot + 1] = make_float4(sx * id_, sy * id_, sz * id_, dist - eps);
    pix[slot] = (uint32_t)i;
}

}  // namespace

int lr_launch_rays_primary(LrhContext *c, const float *cam, int w, int h, float tmin, float tmax, float *rays)
{
    const int n = w * h;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_rays_primary, dim3((n + RBLOCK - 1) / RBLOCK), dim3(RBLOCK), 0, c->stream, cam[0], cam[1], cam[2],
                       cam[3], cam[4], cam[5], cam[6], cam[7], cam[8], cam[9], cam[10], cam[11], w, h, tmin, tmax,
                       reinterpret_cast<float4 *>(rays));
    LR_HIP(hipGetLastError());
    return 0;
}

int lr_launch_clear(LrhContext *c, uint32_t *p, int n)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_clear, dim3((n + 255) / 256), dim3(256), 0, c->stream, p, n);
    LR_HIP(hipGetLastError());
    return 0;
}

int lr_launch_rays_secondary(LrhContext *c, const float *rays, const float *hits, int n, const float *tris,
                             const float *light, float eps, uint32_t seed, float *bounce, float *shadow, uint32_t *pix,
                             uint32_t *count)
{
    int rc = lr_launch_clear(c, count, 1);
    if (rc || n <= 0) return rc;
    hipLaunchKernelGGL(k_rays_secondary, dim3((n + RBLOCK - 1) / RBLOCK), dim3(RBLOCK), 0, c->stream,
                       reinterpret_cast<const float4 *>(rays), reinterpret_cast<const float4 *>(hits), n,
                       reinterpret_cast<const float4 *>(tris), light[0], light[1], light[2], eps, seed,
                       reinterpret_cast<float4 *>(bounce), reinterpret_cast<float4 *>(shadow), pix, count);
    LR_HIP(hipGetLastError());
    return 0;
}
