// Learned per-amplitude preconditioner: a 4 -> 128 -> 128 -> 1 network
// (SiLU) that scales the orbital-energy-denominator step of the amplitude
// solver, r = s * g / D with s = 1 + y(x), and its training.  No counterpart
// in the reference, whose step is r = g / D (baserout/diavc.f:31-35,
// contract/dia4op_ev.f).
//
// Parameters (float32, MLP_P = 17281, in this order): W1[128][4], b1[128],
// W2[128][128] (row = output unit), b2[128], w3[128], b3.
//   z1 = bf(W1) x + b1,  a = bf(silu(z1)),  z2 = bf(W2) a + b2,
//   y  = bf(w3) . silu(z2) + b3
// bf() rounds to bfloat16 (nearest, ties to even).  Inference keeps the
// weights in bfloat16 and runs the 128 x 128 layer on BF16 MFMA
// (v_mfma_f32_32x32x8_bf16, float32 accumulation); training keeps float32
// master weights, evaluates the same rounded network on the vector ALU and
// passes gradients straight through bf().
//
// Features of amplitude i: u = g / D' (D' as the solver's step),
//   x = [D/4 - 1, asinh(u / rms(u)), asinh(t / rms(t)) (0 if rms(t) = 0), min(it, 32)/32]
// (the asinh in float32 of the float64 ratio)
// Loss of the training pairs (x, u~, dt~): mean((1 + y) u~ + dt~)^2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "../../include/gecco_hip_mlp.h"

namespace gecco {
int set_error(int code, const char *msg);
void clear_error();
}  // namespace gecco

namespace {

constexpr int NH = 128;
constexpr int P = GECCO_HIP_MLP_P;
constexpr int OW1 = 0, OB1 = 512, OW2 = 640, OB2 = 17024, OW3 = 17152, OB3 = 17280;
constexpr int NT = 256;
constexpr int NORM_GROUPS = 512;           // workgroups of the norm pass
constexpr int GRAD_GROUPS = GECCO_HIP_MLP_GRAD_GROUPS;
constexpr int TB = 64;                     // samples per training tile
constexpr double DIA_THRESH = 1e-10;
constexpr int W2LD = NH + 4;               // padded LDS row of the bf16 W2

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t bf16_bits(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_round(float f) {
    return __uint_as_float((uint32_t)bf16_bits(f) << 16);
}
__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
// inference: hardware exp and reciprocal (a few ulp)
__device__ __forceinline__ float silu_fast(float z) {
    return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}
__device__ __forceinline__ float dsilu(float z) {
    const float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// ---------------------------------------------------------------- features
// partial sums of u^2 and t^2 per workgroup, then one workgroup adds them in
// workgroup order: work[0] = rms(u), work[1] = rms(t)
__global__ __launch_bounds__(NT) void mlp_norms_partial(const double *__restrict__ diag,
                                                        const double *__restrict__ rest,
                                                        const double *__restrict__ t, int64_t n,
                                                        double *__restrict__ work) {
    double su = 0.0, st = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
        const double ti = t[i], d = diag[i];
        const double dv = fabs(d) <= DIA_THRESH ? DIA_THRESH : d;
        const double u = (d * ti + rest[i]) / dv;
        su += u * u;
        st += ti * ti;
    }
    __shared__ double red[2][NT / 64];
    for (int o = 32; o > 0; o >>= 1) {
        su += __shfl_down(su, o, 64);
        st += __shfl_down(st, o, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) { red[0][wave] = su; red[1][wave] = st; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < NT / 64; ++w) { a += red[0][w]; b += red[1][w]; }
        work[2 + 2 * blockIdx.x] = a;
        work[3 + 2 * blockIdx.x] = b;
    }
}

__global__ __launch_bounds__(64) void mlp_norms_final(int ngroups, int64_t n, double *__restrict__ work) {
    if (threadIdx.x != 0) return;
    double a = 0.0, b = 0.0;
    for (int g = 0; g < ngroups; ++g) { a += work[2 + 2 * g]; b += work[3 + 2 * g]; }
    work[0] = sqrt(a / (double)n);
    work[1] = sqrt(b / (double)n);
}

// ---------------------------------------------------------------- inference
// Each wave evaluates 32 amplitudes at a time: lane l holds amplitude l % 32
// and, for every 8-wide slice ks of the hidden layer, the 4 activations
// k = 8 ks + 4 (l / 32) + q as the A operand of v_mfma_f32_32x32x8_bf16;
// the B operand is W2[j][k..k+3] for output j = 32 cb + l % 32.  Output
// element (i, j) of a 32 x 32 block sits in lane j + 32 ((i / 4) % 2),
// register (i % 4) + 4 (i / 8).
// mode 0: features from (diag, rest, t, rms), output scale = 1 + y (double)
// mode 1: features x[n][4] given, output y (float)
template <int MODE>
__global__ __launch_bounds__(NT) void mlp_infer_kernel(const float *__restrict__ prm,
                                                       const float *__restrict__ xin,
                                                       const double *__restrict__ diag,
                                                       const double *__restrict__ rest,
                                                       const double *__restrict__ t,
                                                       const double *__restrict__ work, double itf,
                                                       int64_t n, float *__restrict__ yout,
                                                       double *__restrict__ sout) {
    __shared__ uint16_t w2[NH * W2LD];
    __shared__ float w1[NH * 4], b1[NH], b2[NH], w3[NH];
    const int tid = threadIdx.x;
    for (int e = tid; e < NH * NH; e += NT) {
        const int j = e / NH, k = e - j * NH;
        w2[j * W2LD + k] = bf16_bits(prm[OW2 + e]);
    }
    for (int e = tid; e < NH * 4; e += NT) w1[e] = bf16_round(prm[OW1 + e]);
    for (int e = tid; e < NH; e += NT) {
        b1[e] = prm[OB1 + e];
        b2[e] = prm[OB2 + e];
        w3[e] = bf16_round(prm[OW3 + e]);
    }
    const float b3 = prm[OB3];
    double ru = 1.0, rt = 0.0;
    if (MODE == 0) { ru = work[0]; rt = work[1]; }
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6;
    const int row = lane & 31, half = lane >> 5;
    for (int64_t base = ((int64_t)blockIdx.x * (NT / 64) + wave) * 32; base < n;
         base += (int64_t)gridDim.x * (NT / 64) * 32) {
        // keeps the weight reads from LDS inside the loop (hoisted, they
        // would hold several hundred registers)
        asm volatile("" ::: "memory");
        const int64_t i = base + row;
        float x[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (i < n) {
            if (MODE == 1) {
#pragma unroll
                for (int m = 0; m < 4; ++m) x[m] = xin[i * 4 + m];
            } else {
                const double d = diag[i], ti = t[i];
                const double dv = fabs(d) <= DIA_THRESH ? DIA_THRESH : d;
                const double u = (d * ti + rest[i]) / dv;
                x[0] = (float)(d / 4.0 - 1.0);
                x[1] = ru > 0.0 ? asinhf((float)(u / ru)) : 0.0f;
                x[2] = rt > 0.0 ? asinhf((float)(ti / rt)) : 0.0f;
                x[3] = (float)itf;
            }
        }
        // first layer: this lane's 64 hidden units, rounded to bfloat16
        v4s a[16];
#pragma unroll
        for (int ks = 0; ks < 16; ++ks) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int j = 8 * ks + 4 * half + q;
                float z = b1[j];
#pragma unroll
                for (int m = 0; m < 4; ++m) z = fmaf(w1[j * 4 + m], x[m], z);
                a[ks][q] = (short)bf16_bits(silu_fast(z));
            }
        }
        // second layer on MFMA, then SiLU and the output weights
        float part[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) part[r] = 0.0f;
#pragma unroll 1
        for (int cb = 0; cb < 4; ++cb) {
            const int j = 32 * cb + row;
            v16f acc;
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
#pragma unroll
            for (int ks = 0; ks < 16; ++ks) {
                const uint16_t *wp = w2 + j * W2LD + 8 * ks + 4 * half;
                v4s b;
                b[0] = (short)wp[0];
                b[1] = (short)wp[1];
                b[2] = (short)wp[2];
                b[3] = (short)wp[3];
                acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a[ks], b, acc, 0, 0, 0);
            }
            const float bj = b2[j], wj = w3[j];
#pragma unroll
            for (int r = 0; r < 16; ++r) part[r] = fmaf(wj, silu_fast(acc[r] + bj), part[r]);
        }
        // sum over the 32 output units of each lane group (lanes 0-31, 32-63)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            float v = part[r];
#pragma unroll
            for (int o = 1; o < 32; o <<= 1) v += __shfl_xor(v, o, 64);
            part[r] = v;
        }
        // lane r (< 16) of each group writes row (r % 4) + 8 (r / 4) + 4 half
        float mine = 0.0f;
#pragma unroll
        for (int r = 0; r < 16; ++r)
            if (row == r) mine = part[r];
        if (row < 16) {
            const int ri = (row & 3) + 8 * (row >> 2) + 4 * half;
            const int64_t io = base + ri;
            if (io < n) {
                const float y = mine + b3;
                if (MODE == 1) yout[io] = y;
                else sout[io] = 1.0 + (double)y;
            }
        }
    }
}

// ---------------------------------------------------------------- training
// Persistent workgroups; each takes tiles of TB samples, evaluates the
// rounded network in float32 on the vector ALU and accumulates the gradient:
// dW2 in registers (thread t owns row t / 2, columns 64 (t % 2) .. + 64),
// the small parameter groups in LDS; partial[g][P] and lossp[g] at the end.
__global__ __launch_bounds__(NT) void mlp_grad_partial(const float *__restrict__ prm,
                                                       const float *__restrict__ x,
                                                       const float *__restrict__ ut,
                                                       const float *__restrict__ dt, int64_t n,
                                                       float *__restrict__ partial,
                                                       double *__restrict__ lossp) {
    __shared__ float W2s[NH * NH];          // bf-rounded W2, [j][k]
    __shared__ float As[TB * NH];            // activations a, [s][k]
    __shared__ float Zs[TB * NH];            // z2, then dz2, [s][j]
    __shared__ float w1s[NH * 4], b1s[NH], b2s[NH], w3s[NH];
    __shared__ float xs[TB * 4], us[TB], ds[TB], dys[TB];
    __shared__ float gW1[NH * 4], gb1[NH], gb2[NH], gw3[NH];
    __shared__ float gb3;
    __shared__ double lred[NT / 64];
    const int tid = threadIdx.x;
    for (int e = tid; e < NH * NH; e += NT) W2s[e] = bf16_round(prm[OW2 + e]);
    for (int e = tid; e < NH * 4; e += NT) { w1s[e] = bf16_round(prm[OW1 + e]); gW1[e] = 0.0f; }
    for (int e = tid; e < NH; e += NT) {
        b1s[e] = prm[OB1 + e];
        b2s[e] = prm[OB2 + e];
        w3s[e] = bf16_round(prm[OW3 + e]);
        gb1[e] = gb2[e] = gw3[e] = 0.0f;
    }
    if (tid == 0) gb3 = 0.0f;
    const float b3 = prm[OB3];
    const float inv_n = 1.0f / (float)n;
    const int jr = tid >> 1, kc = (tid & 1) * 64;          // dW2 ownership
    float gW2[64];
#pragma unroll
    for (int c = 0; c < 64; ++c) gW2[c] = 0.0f;
    double lacc = 0.0;
    const int s_me = tid >> 2, q_me = tid & 3;              // (sample, quarter)
    const int64_t ntile = (n + TB - 1) / TB;
    for (int64_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        const int64_t s0 = tile * TB;
        __syncthreads();
        for (int e = tid; e < TB * 4; e += NT) {
            const int64_t s = s0 + e / 4;
            xs[e] = s < n ? x[s * 4 + (e & 3)] : 0.0f;
        }
        for (int e = tid; e < TB; e += NT) {
            const int64_t s = s0 + e;
            us[e] = s < n ? ut[s] : 0.0f;       // u~ = 0: the sample adds nothing
            ds[e] = s < n ? dt[s] : 0.0f;
        }
        __syncthreads();
        // layer 1: sample s_me, hidden units 32 q_me .. + 32
        for (int c = 0; c < 32; ++c) {
            const int j = 32 * q_me + c;
            float z = b1s[j];
#pragma unroll
            for (int m = 0; m < 4; ++m) z = fmaf(w1s[j * 4 + m], xs[s_me * 4 + m], z);
            As[s_me * NH + j] = bf16_round(silu(z));
        }
        __syncthreads();
        // layer 2 and the output
        float ypart = 0.0f;
        for (int c = 0; c < 32; ++c) {
            const int j = 32 * q_me + c;
            float z = b2s[j];
            for (int k = 0; k < NH; ++k) z = fmaf(W2s[j * NH + k], As[s_me * NH + k], z);
            Zs[s_me * NH + j] = z;
            ypart = fmaf(w3s[j], silu(z), ypart);
        }
        ypart += __shfl_xor(ypart, 1, 64);
        ypart += __shfl_xor(ypart, 2, 64);
        if (q_me == 0) {
            const float y = ypart + b3;
            const float e = (1.0f + y) * us[s_me] + ds[s_me];
            lacc += (double)e * (double)e;
            dys[s_me] = 2.0f * e * us[s_me] * inv_n;
        }
        __syncthreads();
        // dz2 (in place of z2), and the w3 / b2 / b3 gradient
        const float dy = dys[s_me];
        for (int c = 0; c < 32; ++c) {
            const int j = 32 * q_me + c;
            const float z = Zs[s_me * NH + j];
            atomicAdd(&gw3[j], dy * silu(z));
            const float dz = dy * w3s[j] * dsilu(z);
            Zs[s_me * NH + j] = dz;
            atomicAdd(&gb2[j], dz);
        }
        if (q_me == 0) atomicAdd(&gb3, dy);
        __syncthreads();
        // dW2[jr][kc..kc+64] += sum_s dz2[s][jr] a[s][k]
        for (int s = 0; s < TB; ++s) {
            const float dz = Zs[s * NH + jr];
#pragma unroll
            for (int c = 0; c < 64; ++c) gW2[c] = fmaf(dz, As[s * NH + kc + c], gW2[c]);
        }
        // da = bf(W2)^T dz2, dz1 = da silu'(z1): W1 / b1 gradient
        for (int c = 0; c < 32; ++c) {
            const int k = 32 * q_me + c;
            float da = 0.0f;
            for (int j = 0; j < NH; ++j) da = fmaf(W2s[j * NH + k], Zs[s_me * NH + j], da);
            float z1 = b1s[k];
#pragma unroll
            for (int m = 0; m < 4; ++m) z1 = fmaf(w1s[k * 4 + m], xs[s_me * 4 + m], z1);
            const float dz1 = da * dsilu(z1);
            atomicAdd(&gb1[k], dz1);
#pragma unroll
            for (int m = 0; m < 4; ++m) atomicAdd(&gW1[k * 4 + m], dz1 * xs[s_me * 4 + m]);
        }
    }
    __syncthreads();
    float *out = partial + (int64_t)blockIdx.x * P;
#pragma unroll
    for (int c = 0; c < 64; ++c) out[OW2 + jr * NH + kc + c] = gW2[c];
    for (int e = tid; e < NH * 4; e += NT) out[OW1 + e] = gW1[e];
    for (int e = tid; e < NH; e += NT) {
        out[OB1 + e] = gb1[e];
        out[OB2 + e] = gb2[e];
        out[OW3 + e] = gw3[e];
    }
    if (tid == 0) out[OB3] = gb3;
    for (int o = 32; o > 0; o >>= 1) lacc += __shfl_down(lacc, o, 64);
    if ((tid & 63) == 0) lred[tid >> 6] = lacc;
    __syncthreads();
    if (tid == 0) {
        double l = 0.0;
        for (int w = 0; w < NT / 64; ++w) l += lred[w];
        lossp[blockIdx.x] = l;
    }
}

// grad[p] = sum over workgroups of partial[g][p], in workgroup order
__global__ __launch_bounds__(NT) void mlp_grad_reduce(const float *__restrict__ partial, int ngroups,
                                                      const double *__restrict__ lossp, int64_t n,
                                                      float *__restrict__ grad,
                                                      double *__restrict__ loss) {
    const int p = blockIdx.x * NT + threadIdx.x;
    if (p < P) {
        float s = 0.0f;
        for (int g = 0; g < ngroups; ++g) s += partial[(int64_t)g * P + p];
        grad[p] = s;
    }
    if (p == 0 && loss) {
        double l = 0.0;
        for (int g = 0; g < ngroups; ++g) l += lossp[g];
        loss[0] = l / (double)n;
    }
}

__global__ __launch_bounds__(NT) void mlp_adam_kernel(float *__restrict__ prm, float *__restrict__ m,
                                                      float *__restrict__ v,
                                                      const float *__restrict__ g, float lr,
                                                      float c1, float c2) {
    const int p = blockIdx.x * NT + threadIdx.x;
    if (p >= P) return;
    const float gp = g[p];
    const float mp = 0.9f * m[p] + 0.1f * gp;
    const float vp = 0.999f * v[p] + 0.001f * gp * gp;
    m[p] = mp;
    v[p] = vp;
    prm[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + 1e-8f);
}

int fail(int code, const char *msg) { return gecco::set_error(code, msg); }

int launched(const char *what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(GECCO_HIP_MLP_EDEVICE, what);
    return 0;
}

}  // namespace

extern "C" {

int gecco_hip_mlp_infer(const float *params, const float *x, int64_t n, float *y, void *stream) {
    gecco::clear_error();
    if (!params || (n > 0 && (!x || !y)) || n < 0) return fail(GECCO_HIP_MLP_EINVAL, "mlp_infer: bad arguments");
    if (n == 0) return 0;
    const int64_t grid = std::min<int64_t>(2048, (n + 127) / 128);
    hipLaunchKernelGGL(mlp_infer_kernel<1>, dim3((unsigned)grid), dim3(NT), 0, (hipStream_t)stream,
                       params, x, nullptr, nullptr, nullptr, nullptr, 0.0, n, y, nullptr);
    return launched("mlp_infer launch");
}

int gecco_hip_mlp_scale(const float *params, const double *diag, const double *rest, const double *t,
                        int64_t n, double it, double *scale, double *work, void *stream) {
    gecco::clear_error();
    if (!params || !work || n < 0 || (n > 0 && (!diag || !rest || !t || !scale)))
        return fail(GECCO_HIP_MLP_EINVAL, "mlp_scale: bad arguments");
    if (n == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    const int ng = (int)std::min<int64_t>(NORM_GROUPS, (n + NT - 1) / NT);
    hipLaunchKernelGGL(mlp_norms_partial, dim3(ng), dim3(NT), 0, s, diag, rest, t, n, work);
    hipLaunchKernelGGL(mlp_norms_final, dim3(1), dim3(64), 0, s, ng, n, work);
    const int64_t grid = std::min<int64_t>(2048, (n + 127) / 128);
    hipLaunchKernelGGL(mlp_infer_kernel<0>, dim3((unsigned)grid), dim3(NT), 0, s, params, nullptr, diag,
                       rest, t, work, std::min(it, 32.0) / 32.0, n, nullptr, scale);
    return launched("mlp_scale launch");
}

int gecco_hip_mlp_grad(const float *params, const float *x, const float *ut, const float *dt, int64_t n,
                       float *grad, double *loss, float *work, void *stream) {
    gecco::clear_error();
    if (!params || !grad || !work || n <= 0 || !x || !ut || !dt)
        return fail(GECCO_HIP_MLP_EINVAL, "mlp_grad: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    const int ng = (int)std::min<int64_t>(GRAD_GROUPS, (n + TB - 1) / TB);
    double *lossp = reinterpret_cast<double *>(work + (int64_t)GRAD_GROUPS * P);
    hipLaunchKernelGGL(mlp_grad_partial, dim3(ng), dim3(NT), 0, s, params, x, ut, dt, n, work, lossp);
    hipLaunchKernelGGL(mlp_grad_reduce, dim3((P + NT - 1) / NT), dim3(NT), 0, s, work, ng, lossp, n,
                       grad, loss);
    return launched("mlp_grad launch");
}

int gecco_hip_mlp_adam(float *params, float *m, float *v, const float *grad, int64_t step, double lr,
                       void *stream) {
    gecco::clear_error();
    if (!params || !m || !v || !grad || step < 1) return fail(GECCO_HIP_MLP_EINVAL, "mlp_adam: bad arguments");
    const float c1 = (float)(1.0 - std::pow(0.9, (double)step));
    const float c2 = (float)(1.0 - std::pow(0.999, (double)step));
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((P + NT - 1) / NT), dim3(NT), 0, (hipStream_t)stream, params,
                       m, v, grad, (float)lr, c1, c2);
    return launched("mlp_adam launch");
}

}  // extern "C"
