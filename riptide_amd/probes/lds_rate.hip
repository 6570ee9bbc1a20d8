// LDS read-rate probe for gfx950 (a standalone program, not part of the
// library): bytes per clock per CU that 16 wavefronts reach with 4-, 8- and
// 16-byte LDS reads, (a) contiguous across the lanes of a wavefront and (b)
// one row per lane at a given row pitch (the S/N phase's access pattern).
// Per-lane base computed once, a uniform offset per iteration, 8 loads in
// flight at immediate offsets; the empty asm statements only mark the loaded
// values as used.  Every index stays inside s[16384] (checked on the host
// for each case before its launch).
//
//   hipcc --offload-arch=gfx950 -O3 -o lds_rate riptide_amd/probes/lds_rate.hip && ./lds_rate
#include <hip/hip_runtime.h>
#include <cstdio>
template <int W, int MODE, int P>
__global__ __launch_bounds__(1024) void lds_read(float* out, long long* cycles, int iters)
{
    __shared__ float s[16384];
    for (int i = threadIdx.x; i < 16384; i += 1024) s[i] = (float)(i & 255);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const float* pb = s + (MODE == 0 ? lane * W : lane * P);
    float acc = 0.0f;
    long long t0 = clock64();
    for (int it = 0; it < iters; ++it) {
        const int k = MODE == 0 ? (it & 7) * 512 * W : (it & 15) * 8 * W;
        const float* p = pb + k;
        constexpr int S = MODE == 0 ? 64 * W : W;  // floats between the 8 loads
        if constexpr (W == 1) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = p[u * S];
            asm volatile("" :: "v"(v[0]), "v"(v[1]), "v"(v[2]), "v"(v[3]), "v"(v[4]), "v"(v[5]), "v"(v[6]), "v"(v[7]));
            acc += v[0];
        } else if constexpr (W == 2) {
            float2 v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = *reinterpret_cast<const float2*>(p + u * S);
#pragma unroll
            for (int u = 0; u < 8; u += 4)
                asm volatile("" :: "v"(v[u].x), "v"(v[u].y), "v"(v[u + 1].x), "v"(v[u + 1].y), "v"(v[u + 2].x), "v"(v[u + 2].y), "v"(v[u + 3].x), "v"(v[u + 3].y));
            acc += v[0].x;
        } else {
            float4 v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = *reinterpret_cast<const float4*>(p + u * S);
#pragma unroll
            for (int u = 0; u < 8; u += 2)
                asm volatile("" :: "v"(v[u].x), "v"(v[u].y), "v"(v[u].z), "v"(v[u].w), "v"(v[u + 1].x), "v"(v[u + 1].y), "v"(v[u + 1].z), "v"(v[u + 1].w));
            acc += v[0].x;
        }
    }
    long long t1 = clock64();
    __syncthreads();
    if (threadIdx.x == 0) cycles[blockIdx.x] = t1 - t0;
    out[blockIdx.x * 1024 + threadIdx.x] = acc;
}
template <int W, int MODE, int P>
void run(const char* name, float* out, long long* cyc)
{
    // largest index touched
    const int maxk = MODE == 0 ? 7 * 512 * W : 15 * 8 * W;
    const int S = MODE == 0 ? 64 * W : W;
    const int maxi = (MODE == 0 ? 63 * W : 63 * P) + maxk + 7 * S + W - 1;
    if (maxi >= 16384) { printf("%s: out of bounds, skipped\n", name); return; }
    const int blocks = 256, iters = 4096;
    long long h[256];
    for (int rep = 0; rep < 2; ++rep) {
        hipLaunchKernelGGL((lds_read<W, MODE, P>), dim3(blocks), dim3(1024), 0, 0, out, cyc, iters);
        (void)hipDeviceSynchronize();
    }
    (void)hipMemcpy(h, cyc, sizeof(h), hipMemcpyDeviceToHost);
    double avg = 0; for (int i = 0; i < blocks; ++i) avg += h[i]; avg /= blocks;
    const double bytes = 1024.0 * iters * 8 * 4 * W;
    const double wave_loads = 16.0 * iters * 8;
    printf("%-18s W=%d: %.0f cycles, %.1f B/clk/CU, %.2f clk per wave-load\n", name, W, avg, bytes / avg, avg / wave_loads);
}
// The access pattern of one FFA merge level without its barriers: per
// element a 4-byte read of a head row, a 4-byte read of a tail row rotated by
// `sh` columns (wrap chosen per element by a compare and a select), their sum
// written back over the head row; lanes on consecutive columns, rows of 512
// floats, each wavefront on its own pair of rows (1024 floats of s[16384]).
__global__ __launch_bounds__(1024) void lds_merge(float* out, long long* cycles, int iters)
{
    __shared__ float s[16384];
    for (int i = threadIdx.x; i < 16384; i += 1024) s[i] = (float)(i & 255);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    float* head = s + wave * 1024;
    const float* tail = head + 512;
    long long t0 = clock64();
    for (int it = 0; it < iters; ++it) {
        const int sh = (it * 37) & 511;
        const int lim = 512 - sh;
        float v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int c = lane + 64 * q;
            const int t = c >= lim ? c + sh - 512 : c + sh;
            v[q] = head[c] + tail[t];
        }
        asm volatile("" ::: "memory");
#pragma unroll
        for (int q = 0; q < 8; ++q) head[lane + 64 * q] = v[q];
        asm volatile("" ::: "memory");
    }
    long long t1 = clock64();
    __syncthreads();
    if (threadIdx.x == 0) cycles[blockIdx.x] = t1 - t0;
    out[blockIdx.x * 1024 + threadIdx.x] = head[lane];
}

int main()
{
    float* out; long long* cyc;
    (void)hipMalloc(&out, 256 * 1024 * sizeof(float));
    (void)hipMalloc(&cyc, 256 * sizeof(long long));
    run<1, 0, 0>("contiguous", out, cyc);
    run<2, 0, 0>("contiguous", out, cyc);
    run<4, 0, 0>("contiguous", out, cyc);
    run<1, 1, 241>("rows pitch 241", out, cyc);
    run<1, 1, 243>("rows pitch 243", out, cyc);
    run<2, 1, 242>("rows pitch 242", out, cyc);
    run<2, 1, 246>("rows pitch 246", out, cyc);
    run<4, 1, 244>("rows pitch 244", out, cyc);
    run<4, 1, 236>("rows pitch 236", out, cyc);
    run<4, 1, 228>("rows pitch 228", out, cyc);
    run<4, 1, 240>("rows pitch 240", out, cyc);
    {
        // every index above stays inside s[16384]: wave*1024 + 512 + 511
        const int blocks = 256, iters = 4096;
        long long h[256];
        for (int rep = 0; rep < 2; ++rep) {
            hipLaunchKernelGGL(lds_merge, dim3(blocks), dim3(1024), 0, 0, out, cyc, iters);
            (void)hipDeviceSynchronize();
        }
        (void)hipMemcpy(h, cyc, sizeof(h), hipMemcpyDeviceToHost);
        double avg = 0; for (int i = 0; i < blocks; ++i) avg += h[i]; avg /= blocks;
        const double bytes = 1024.0 * iters * 8 * 3 * 4;  // two reads + one write per element
        printf("%-18s    : %.0f cycles, %.1f B/clk/CU (2 reads + 1 write per element)\n", "merge pattern", avg, bytes / avg);
    }
    printf("done\n");
    return 0;
}
