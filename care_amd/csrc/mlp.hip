// Training and stand-alone inference of the learned correction gate
// (mlp_device.h): BF16 MFMA forward and backward passes with FP32
// accumulation, deterministic gradient reduction and Adam on FP32 master
// weights.  No counterpart in the reference (its classifier is a scikit-learn
// random forest trained offline, ml/care.py:231-254).

#include "mlp_device.h"

namespace care_engine {

namespace {

constexpr int kRowsPerBlock = 64;  // 4 waves x 16 rows
constexpr int kPartialStride = kMlpParams + 1;  // gradient + loss

__device__ __forceinline__ void wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// sum of v over the four lanes l%16 + 16g (g = 0..3); every lane receives it
__device__ __forceinline__ float sum_over_groups(float v) {
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}

__device__ __forceinline__ ushort4v load4_strided(const bf16raw* p, int stride) {
    ushort4v v;
    v[0] = p[0];
    v[1] = p[stride];
    v[2] = p[2 * stride];
    v[3] = p[3 * stride];
    return v;
}

}  // namespace

// One wave per tile of 16 rows: forward, binary cross-entropy with mean over
// the batch, backward.  Every wave writes its gradient (and loss) to
// partial[(block * 4 + wave) * kPartialStride ...].  Rows index X through
// perm[first .. first + rows); rows beyond `rows` contribute nothing.
__global__ void __launch_bounds__(256) k_mlp_grad(const bf16raw* X, const u8* Y, const u32* perm, u64 first, int rows,
                                                  const bf16raw* w1b, const bf16raw* w2b, const bf16raw* w3b,
                                                  const float* master, float invBatch, float* partial) {
    __shared__ __attribute__((aligned(16))) bf16raw sX[4][16 * kMlpIn];
    __shared__ __attribute__((aligned(16))) char sScratch[4][kMlpTileScratch];
    __shared__ float sZ[4][16];
    __shared__ float sDz[4][16];
    __shared__ __attribute__((aligned(16))) bf16raw sD3[4][16 * kMlpH3];
    __shared__ __attribute__((aligned(16))) bf16raw sD2[4][16 * kMlpH2];
    __shared__ __attribute__((aligned(16))) bf16raw sD1[4][16 * kMlpH1];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int r = lane & 15, g = lane >> 4;
    const int tileRow0 = blockIdx.x * kRowsPerBlock + wave * 16;
    bf16raw* x = sX[wave];
    // gather the tile: lane handles (row lane/4, 8 features at 8*(lane%4))
    {
        const int row = lane >> 2, part = lane & 3;
        const int gr = tileRow0 + row;
        ushort8 v = {0, 0, 0, 0, 0, 0, 0, 0};
        if (gr < rows) v = load8(X + u64(perm[first + gr]) * kMlpIn + 8 * part);
        *reinterpret_cast<ushort8*>(x + row * kMlpIn + 8 * part) = v;
    }
    wave_fence();
    MlpWeights W;
    W.w1 = w1b;
    W.w2 = w2b;
    W.w3 = w3b;
    W.b1 = master + kMlpB1;
    W.b2 = master + kMlpB2;
    W.b3 = master + kMlpB3;
    W.w4 = master + kMlpW4;
    W.b4 = master[kMlpB4];
    mlp_tile_forward(W, x, sScratch[wave], sZ[wave]);
    const bf16raw* h1 = reinterpret_cast<const bf16raw*>(sScratch[wave]);
    const bf16raw* h2 = h1 + 16 * kMlpH1;
    const float* h3 = reinterpret_cast<const float*>(h2 + 16 * kMlpH2);
    float* gout = partial + u64(blockIdx.x * 4 + wave) * kPartialStride;

    // output layer: dz = (sigmoid(z) - y) / batch, loss = softplus(z) - y z
    float loss = 0.f;
    if (lane < 16) {
        const int gr = tileRow0 + lane;
        float dz = 0.f;
        if (gr < rows) {
            const float zz = sZ[wave][lane];
            const float y = float(Y[perm[first + gr]]);
            const float p = 1.0f / (1.0f + expf(-zz));
            dz = (p - y) * invBatch;
            loss = (fmaxf(zz, 0.f) - y * zz + log1pf(expf(-fabsf(zz)))) * invBatch;
        }
        sDz[wave][lane] = dz;
    }
    wave_fence();
    // dw4, db4 and dH3 = dz w4 relu'(h3): lane (r, g) covers rows 4g..4g+3 of column r
    {
        float dw4 = 0.f, db3 = 0.f;
        const float w4 = master[kMlpW4 + r];
        for (int i = 0; i < 4; i++) {
            const int row = 4 * g + i;
            const float dz = sDz[wave][row];
            const float hv = h3[row * kMlpH3 + r];
            dw4 = fmaf(dz, hv, dw4);
            const float d = hv > 0.f ? dz * w4 : 0.f;
            db3 += d;
            sD3[wave][row * kMlpH3 + r] = f2bf(d);
        }
        dw4 = sum_over_groups(dw4);
        db3 = sum_over_groups(db3);
        float db4 = lane < 16 ? sDz[wave][lane] : 0.f;
        for (int o = 8; o > 0; o >>= 1) db4 += __shfl_xor(db4, o, 64);
        float ls = loss;
        for (int o = 8; o > 0; o >>= 1) ls += __shfl_xor(ls, o, 64);
        if (lane < 16) {
            gout[kMlpW4 + r] = dw4;
            gout[kMlpB3 + r] = db3;
        }
        if (lane == 0) {
            gout[kMlpB4] = db4;
            gout[kMlpParams] = ls;
        }
    }
    wave_fence();
    // dW3[n][k] = sum_row dH3[row][n] h2[row][k]   (n 16, k 32)
    for (int kt = 0; kt < 2; kt++) {
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma16(load4_strided(sD3[wave] + (4 * g) * kMlpH3 + r, kMlpH3),
                     load4_strided(h2 + (4 * g) * kMlpH2 + 16 * kt + r, kMlpH2), acc);
        for (int i = 0; i < 4; i++) gout[kMlpW3 + (4 * g + i) * kMlpH2 + 16 * kt + r] = acc[i];
    }
    // dH2[row][k] = sum_n dH3[row][n] W3[n][k] relu'(h2)   (K = 16)
    for (int kt = 0; kt < 2; kt++) {
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
        ushort4v a;
        for (int j = 0; j < 4; j++) a[j] = sD3[wave][r * kMlpH3 + 4 * g + j];
        acc = mfma16(a, load4_strided(w3b + (4 * g) * kMlpH2 + 16 * kt + r, kMlpH2), acc);
        float db = 0.f;
        for (int i = 0; i < 4; i++) {
            const int row = 4 * g + i, col = 16 * kt + r;
            const float d = bf2f(h2[row * kMlpH2 + col]) > 0.f ? acc[i] : 0.f;
            db += d;
            sD2[wave][row * kMlpH2 + col] = f2bf(d);
        }
        db = sum_over_groups(db);
        if (lane < 16) gout[kMlpB2 + 16 * kt + r] = db;
    }
    wave_fence();
    // dW2[n][k] = sum_row dH2[row][n] h1[row][k]   (n 32, k 64)
    for (int nt = 0; nt < 2; nt++)
        for (int kt = 0; kt < 4; kt++) {
            floatx4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma16(load4_strided(sD2[wave] + (4 * g) * kMlpH2 + 16 * nt + r, kMlpH2),
                         load4_strided(h1 + (4 * g) * kMlpH1 + 16 * kt + r, kMlpH1), acc);
            for (int i = 0; i < 4; i++) gout[kMlpW2 + (16 * nt + 4 * g + i) * kMlpH1 + 16 * kt + r] = acc[i];
        }
    // dH1[row][k] = sum_n dH2[row][n] W2[n][k] relu'(h1)   (K = 32)
    for (int kt = 0; kt < 4; kt++) {
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
        ushort8 b;
        for (int j = 0; j < 8; j++) b[j] = w2b[(8 * g + j) * kMlpH1 + 16 * kt + r];
        acc = mfma32(load8(sD2[wave] + r * kMlpH2 + 8 * g), b, acc);
        float db = 0.f;
        for (int i = 0; i < 4; i++) {
            const int row = 4 * g + i, col = 16 * kt + r;
            const float d = bf2f(h1[row * kMlpH1 + col]) > 0.f ? acc[i] : 0.f;
            db += d;
            sD1[wave][row * kMlpH1 + col] = f2bf(d);
        }
        db = sum_over_groups(db);
        if (lane < 16) gout[kMlpB1 + 16 * kt + r] = db;
    }
    wave_fence();
    // dW1[n][k] = sum_row dH1[row][n] x[row][k]   (n 64, k 32)
    for (int nt = 0; nt < 4; nt++)
        for (int kt = 0; kt < 2; kt++) {
            floatx4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma16(load4_strided(sD1[wave] + (4 * g) * kMlpH1 + 16 * nt + r, kMlpH1),
                         load4_strided(x + (4 * g) * kMlpIn + 16 * kt + r, kMlpIn), acc);
            for (int i = 0; i < 4; i++) gout[kMlpW1 + (16 * nt + 4 * g + i) * kMlpIn + 16 * kt + r] = acc[i];
        }
}

// Sums the partial gradients in a fixed order: a block covers 32
// parameters; its 8 thread groups sum the partials p = g, g + 8, ... of one
// parameter each, and group sums are added in group order.  mode 0: writes
// the gradient (and the loss at index kMlpParams) to out; mode 1: Adam step
// on the FP32 master weights (t = step number from 1), refreshes the BF16
// copies and adds the loss to *lossAcc.
constexpr int kReduceParams = 32, kReduceGroups = 8;
__global__ void __launch_bounds__(kReduceParams * kReduceGroups)
    k_mlp_reduce(const float* partial, int nPartials, int mode, float* out, float* master, float* m1, float* m2,
                 float lr, float beta1, float beta2, float eps, float bc1, float bc2, bf16raw* w1b, bf16raw* w2b,
                 bf16raw* w3b, float* lossAcc) {
    __shared__ float sPart[kReduceGroups][kReduceParams];
    const int j = threadIdx.x % kReduceParams, grp = threadIdx.x / kReduceParams;
    const int i = blockIdx.x * kReduceParams + j;
    float s = 0.f;
    if (i <= kMlpParams)
        for (int p = grp; p < nPartials; p += kReduceGroups) s += partial[u64(p) * kPartialStride + i];
    sPart[grp][j] = s;
    __syncthreads();
    if (grp != 0 || i > kMlpParams) return;
    s = 0.f;
    for (int g = 0; g < kReduceGroups; g++) s += sPart[g][j];
    if (mode == 0) {
        out[i] = s;
        return;
    }
    if (i == kMlpParams) {
        lossAcc[0] += s;
        return;
    }
    const float mm = beta1 * m1[i] + (1.f - beta1) * s;
    const float vv = beta2 * m2[i] + (1.f - beta2) * s * s;
    m1[i] = mm;
    m2[i] = vv;
    const float w = master[i] - lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    master[i] = w;
    if (i < kMlpB1) w1b[i - kMlpW1] = f2bf(w);
    else if (i >= kMlpW2 && i < kMlpB2) w2b[i - kMlpW2] = f2bf(w);
    else if (i >= kMlpW3 && i < kMlpB3) w3b[i - kMlpW3] = f2bf(w);
}

// BF16 copies of the weight matrices of a master parameter vector
__global__ void k_mlp_cast(const float* master, bf16raw* w1b, bf16raw* w2b, bf16raw* w3b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kMlpH1 * kMlpIn) w1b[i] = f2bf(master[kMlpW1 + i]);
    if (i < kMlpH2 * kMlpH1) w2b[i] = f2bf(master[kMlpW2 + i]);
    if (i < kMlpH3 * kMlpH2) w3b[i] = f2bf(master[kMlpW3 + i]);
}

// Stand-alone forward pass: one wave per tile of 16 rows of X (bf16), z out.
__global__ void __launch_bounds__(256) k_mlp_forward(const bf16raw* X, u64 n, const bf16raw* w1b, const bf16raw* w2b,
                                                     const bf16raw* w3b, const float* master, float* zOut) {
    __shared__ __attribute__((aligned(16))) bf16raw sX[4][16 * kMlpIn];
    __shared__ __attribute__((aligned(16))) char sScratch[4][kMlpTileScratch];
    __shared__ float sZ[4][16];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const u64 row0 = (u64(blockIdx.x) * 4 + wave) * 16;
    {
        const int row = lane >> 2, part = lane & 3;
        ushort8 v = {0, 0, 0, 0, 0, 0, 0, 0};
        if (row0 + row < n) v = load8(X + (row0 + row) * kMlpIn + 8 * part);
        *reinterpret_cast<ushort8*>(sX[wave] + row * kMlpIn + 8 * part) = v;
    }
    wave_fence();
    MlpWeights W;
    W.w1 = w1b;
    W.w2 = w2b;
    W.w3 = w3b;
    W.b1 = master + kMlpB1;
    W.b2 = master + kMlpB2;
    W.b3 = master + kMlpB3;
    W.w4 = master + kMlpW4;
    W.b4 = master[kMlpB4];
    mlp_tile_forward(W, sX[wave], sScratch[wave], sZ[wave]);
    if (lane < 16 && row0 + lane < n) zOut[row0 + lane] = sZ[wave][lane];
}

void launch_mlp_grad(const bf16raw* X, const u8* Y, const u32* perm, u64 first, int rows, const bf16raw* w1b,
                     const bf16raw* w2b, const bf16raw* w3b, const float* master, float invBatch, float* partial,
                     hipStream_t s) {
    const int blocks = (rows + kRowsPerBlock - 1) / kRowsPerBlock;
    hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(256), 0, s, X, Y, perm, first, rows, w1b, w2b, w3b, master,
                       invBatch, partial);
}

int mlp_partials(int rows) { return 4 * ((rows + kRowsPerBlock - 1) / kRowsPerBlock); }
int mlp_partial_stride() { return kPartialStride; }

void launch_mlp_reduce(const float* partial, int nPartials, int mode, float* out, float* master, float* m1, float* m2,
                       float lr, float beta1, float beta2, float eps, float bc1, float bc2, bf16raw* w1b, bf16raw* w2b,
                       bf16raw* w3b, float* lossAcc, hipStream_t s) {
    const int blocks = (kMlpParams + 1 + kReduceParams - 1) / kReduceParams;
    hipLaunchKernelGGL(k_mlp_reduce, dim3(blocks), dim3(kReduceParams * kReduceGroups), 0, s, partial, nPartials, mode,
                       out, master, m1, m2, lr,
                       beta1, beta2, eps, bc1, bc2, w1b, w2b, w3b, lossAcc);
}

void launch_mlp_cast(const float* master, bf16raw* w1b, bf16raw* w2b, bf16raw* w3b, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_cast, dim3((kMlpH1 * kMlpIn + 255) / 256), dim3(256), 0, s, master, w1b, w2b, w3b);
}

void launch_mlp_forward(const bf16raw* X, u64 n, const bf16raw* w1b, const bf16raw* w2b, const bf16raw* w3b,
                        const float* master, float* zOut, hipStream_t s) {
    if (n == 0) return;
    const u64 blocks = (n + 63) / 64;
    hipLaunchKernelGGL(k_mlp_forward, dim3(unsigned(blocks)), dim3(256), 0, s, X, n, w1b, w2b, w3b, master, zOut);
}

}  // namespace care_engine
