// lr_bsdf.hip - the device BSDF on a batch of cases (include/lr_hip.h: lrh_bsdf_cases), the
// kernel-level check of the shading code the engine runs: bsdf_make / bsdf_eval / bsdf_pick /
// bsdf_sample_lobe of lr_pt_dev.h, exactly as k_shade calls them. The reference checks its WGSL
// BSDF against its C one the same way (gpu_bsdf_run, tests/test_gpu_brdf.c:21-90).
#include "lr_internal.h"
#include "lr_pt_dev.h"

namespace {

// one case a thread: in 17 floats (base rgb, metallic, roughness, ior, transmission, outside,
// v xyz, l xyz, ul, u1, u2; the normal is +z), out 16 floats (lro_bsdf_case's layout: f rgb,
// pdf | ok, lobe, l xyz, w rgb, pdf, delta, through, 0)
__global__ __launch_bounds__(256) void k_bsdf_cases(const float *__restrict__ in, int n, const float *__restrict__ tab,
                                                    float *__restrict__ out)
{
    const int i = (int)(blockIdx.x * 256 + threadIdx.x);
    if (i >= n) return;
    const float *c = in + 17 * (size_t)i;
    // material_pbr (scene.h:70-74) in the engine's 3-float4 record
    float4 mat[3];
    mat[0] = make_float4(c[0], c[1], c[2], (float)MAT_PBR);
    mat[1] = make_float4(0.0f, 0.0f, 0.0f, c[5]);
    mat[2] = make_float4(c[3], c[4], c[6], 0.0f);
    int type;
    const V3 nrm = v3(0.0f, 0.0f, 1.0f), v = v3(c[8], c[9], c[10]), l = v3(c[11], c[12], c[13]);
    Bsdf b = bsdf_make(mat, 0, c[7] > 0.5f, tab, type);
    float pdf;
    const V3 f = bsdf_eval(b, tab, nrm, v, l, pdf);
    float *o = out + 16 * (size_t)i;
    o[0] = f.x;
    o[1] = f.y;
    o[2] = f.z;
    o[3] = pdf;
    bsdf_view(b, tab, vdot(v, nrm));
    const int lobe = bsdf_pick(b, c[14]);
    BSample s;
    s.l = v3(0.0f, 0.0f, 0.0f);
    s.w = v3(0.0f, 0.0f, 0.0f);
    s.pdf = 0.0f;
    s.delta = false;
    const bool ok = bsdf_sample_lobe(b, tab, lobe, nrm, v, c[15], c[16], s);
    o[4] = ok ? 1.0f : 0.0f;
    o[5] = (float)lobe;
    o[6] = s.l.x;
    o[7] = s.l.y;
    o[8] = s.l.z;
    o[9] = s.w.x;
    o[10] = s.w.y;
    o[11] = s.w.z;
    o[12] = s.pdf;
    o[13] = s.delta ? 1.0f : 0.0f;
    o[14] = ok && s.l.z < 0.0f ? 1.0f : 0.0f;   // through: below the surface (l . n < 0)
    o[15] = 0.0f;
}

}  // namespace

extern "C" void lrs_brdf_tables(float *out);

extern "C" int lrh_bsdf_cases(LrhContext *c, const float *d_in, int n, float *d_out)
{
    if (!c || !d_in || !d_out || n < 0) return lr_fail(LRH_EINVAL, "lrh_bsdf_cases: bad arguments");
    if (n == 0) return 0;
    LR_HIP(hipSetDevice(c->device));
    int rc = lr_scratch_reserve(c, LR_SCRATCH_PT, 2112 * sizeof(float));
    if (rc) return rc;
    float tab[2112];
    lrs_brdf_tables(tab);
    float *d_tab = (float *)c->scratch[LR_SCRATCH_PT].base;
    LR_HIP(hipMemcpyAsync(d_tab, tab, sizeof(tab), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_bsdf_cases, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, d_in, n,
                       (const float *)d_tab, d_out);
    LR_HIP(hipGetLastError());
    // the table is host memory on this stack frame: wait for its copy (and the cases) to end
    LR_HIP(hipStreamSynchronize(c->stream));
    return 0;
}
