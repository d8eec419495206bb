// FP32 VALU issue-rate probe for MI355X (gfx950), non-packed arithmetic: how many
// wave64 FP32 FMA instructions a SIMD completes per nanosecond when nothing but
// independent FMAs are in flight, in the operand forms the force kernels use.
// The result is the practical ceiling of any kernel built without packed FP32
// (v_pk_fma_f32 & co, disabled on this platform); the published 157.3 TFLOP/s
// assumes packed math.
//
//   hipcc -O3 --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops \
//       ahotsph-2hot-apps_amd/roofline/valu_issue.hip -o /tmp/valu_issue && /tmp/valu_issue
//
// Modes (8 independent chains per lane, 32 single-wave workgroups per CU):
//   0  v_fmac: VGPR x VGPR + accumulator            (the body-pair loop's form)
//   1  v_fmac: SGPR x VGPR + accumulator            (the cell loop's: record coefficient in an SGPR)
//   2  v_fma + v_mul, VGPR only, three sources      (VOP3 form)
#include <hip/hip_runtime.h>

#include <cstdio>

#define CH 8
#define IT 4096

template <int MODE>
__global__ __launch_bounds__(64) void probe(float *out, float c) {
    float a[CH], v[CH];
#pragma unroll
    for (int i = 0; i < CH; i++) {
        a[i] = 0.999f + threadIdx.x * 1e-7f + i * 1e-5f;
        v[i] = threadIdx.x * 1e-3f + i;
    }
    const float cv = c + threadIdx.x * 1e-7f;
    for (int it = 0; it < IT; it++) {
#pragma unroll
        for (int i = 0; i < CH; i++) {
            if (MODE == 0)
                v[i] = fmaf(a[i], cv, v[i]);
            else if (MODE == 1)
                v[i] = fmaf(a[i], c, v[i]);
            else
                v[i] = fmaf(a[i], a[(i + 1) % CH], v[i]) * cv;
        }
    }
    float s = 0;
#pragma unroll
    for (int i = 0; i < CH; i++) s += v[i];
    out[blockIdx.x * 64 + threadIdx.x] = s;
}

template <int MODE>
static double run(int blocks, float *out, int cus) {
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    double best = 0;
    for (int rep = 0; rep < 3; rep++) {
        (void)hipEventRecord(e0);
        hipLaunchKernelGGL(probe<MODE>, dim3(blocks), dim3(64), 0, 0, out, 1e-6f);
        (void)hipEventRecord(e1);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        const double instr = (double)blocks * IT * CH * (MODE == 2 ? 2 : 1);
        const double rate = instr / (cus * 4.0) / (ms * 1e6);  // wave-instructions per ns per SIMD
        if (rate > best)
            best = rate;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

int main() {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, 0) != hipSuccess)
        return 1;
    const int cus = p.multiProcessorCount, blocks = cus * 32;
    float *out = nullptr;
    if (hipMalloc(&out, sizeof(float) * blocks * 64) != hipSuccess)
        return 1;
    const double r[3] = {run<0>(blocks, out, cus), run<1>(blocks, out, cus), run<2>(blocks, out, cus)};
    const char *name[3] = {"v_fmac VGPR x VGPR", "v_fmac SGPR x VGPR", "v_fma + v_mul VOP3"};
    for (int m = 0; m < 3; m++)
        printf("mode %d %-20s %.3f wave-instr/ns/SIMD = %.1f TFLOP/s if every instruction were an FMA\n",
               m, name[m], r[m], r[m] * 1e9 * 64 * 2 * cus * 4 / 1e12);
    (void)hipFree(out);
    return 0;
}
